"""Pins the CPU MLP oracle (oracle/mlp_oracle.py) by central finite
differences in FP64, before it is used to check the GPU training kernels."""
import numpy as np

from massivemds_amd import datagen
from oracle import mlp_oracle as mo


def _params(n, seed=3):
    rng = np.random.default_rng(seed)
    shapes = [(mo.IN, mo.HIDDEN), (mo.HIDDEN,), (mo.HIDDEN, mo.HIDDEN), (mo.HIDDEN,), (mo.HIDDEN, mo.OUT),
              (mo.OUT,), (n, mo.EMBED)]
    return [rng.normal(size=s) * (0.3 if len(s) == 2 else 0.1) for s in shapes]


def test_gradient_matches_finite_differences():
    n, p = 9, 4
    params = _params(n)
    x = datagen.normals(1, 1, n * p).reshape(n, p)
    g = datagen.normals(1, 2, n * p).reshape(n, p)
    xc = 0.5 * datagen.normals(1, 3, n * p).reshape(n, p)
    gc = 0.5 * datagen.normals(1, 4, n * p).reshape(n, p)
    sx, sg = np.full(16, 1.3), np.full(16, 0.7)
    loss, grads = mo.forward_backward(params, x, g, xc, gc, sx, sg, p)
    rng = np.random.default_rng(0)
    h = 1e-6
    for t, (w, gw) in enumerate(zip(params, grads)):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            if t == 0 and idx[0] >= p + mo.EMBED:
                continue  # padded input rows carry no signal
            if t in (4, 5) and idx[-1] >= p:
                continue  # padded outputs are not in the loss
            old = w[idx]
            w[idx] = old + h
            lp, _ = mo.forward_backward(params, x, g, xc, gc, sx, sg, p)
            w[idx] = old - h
            lm, _ = mo.forward_backward(params, x, g, xc, gc, sx, sg, p)
            w[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - gw[idx]) <= 1e-6 * max(abs(fd), 1e-3), (t, idx, fd, gw[idx])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.001953125, 1.01171875, -3.14159, 65504.0, 1e-30])
    r = mo.to_bf16(v)
    # 1 + 2^-8 and 1 + 3 2^-8 are ties: they go to the even neighbour (down, then up)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 and r[3] == 1.015625
    assert abs(r[4] + 3.140625) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)


def test_prediction_is_the_trained_map():
    """predict() returns gc + sigma_g * output, the loss's target map: a
    network whose loss is zero on (x, g) predicts g exactly."""
    n, p = 7, 3
    params = _params(n, seed=5)
    x = datagen.normals(2, 1, n * p).reshape(n, p)
    xc = datagen.normals(2, 2, n * p).reshape(n, p)
    gc = datagen.normals(2, 3, n * p).reshape(n, p)
    sx, sg = np.full(16, 0.9), np.full(16, 1.7)
    g = mo.predict(params, x, xc, gc, sx, sg, p)
    loss, _ = mo.forward_backward(params, x, g, xc, gc, sx, sg, p)
    assert loss < 1e-28
    loss2, _ = mo.forward_backward(params, x, g + 1.7 * 0.1, xc, gc, sx, sg, p)
    assert abs(loss2 - 0.01) < 1e-12  # every target shifted by 0.1 after scaling
