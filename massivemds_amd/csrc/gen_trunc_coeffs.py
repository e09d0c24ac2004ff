"""Generates mds_trunc_coeffs.h: the polynomial and rational coefficients the
kernels use for the truncated-normal terms.  Every approximation is fitted
here from its defining formula with mpmath (60 digits) by iteratively
reweighted linear least squares on Chebyshev nodes (Loeb / Sanathanan-Koerner
iteration for the rational fit), minimising relative error:

* Mills ratio m(z) = Q(z) / phi(z), Q(z) = erfc(z / sqrt 2) / 2, on
  z in [0, 12] as P_n(u) / Q_{n+1}(u) with u = z / 4 (a rational form that
  keeps m's 1/z decay; beyond z = 12, Q < 2e-33 and no longer matters in
  double precision, so z is clamped there);
* exp(-r) on |r| <= ln(2) / 2 (after the usual 2^k range reduction);
* atanh(s) / s as a polynomial in v = s^2 on v in [0, 1/9], for
  log(1 - Q) = 2 atanh(-Q / (2 - Q)) with Q in [0, 1/2];
* the FP64 interval table of log Phi(z): on each z interval
  [k / 64, (k + 1) / 64), k < 640, a degree-5 polynomial in t = 64 z - k,
  fitted by least squares to log Phi and, jointly, to its derivative
  phi / Phi (so that the kernels take phi / Phi as the polynomial's
  derivative); row 640 is zero (z >= 10: |log Phi| < 8e-24 and
  phi / Phi < 8e-23).  Rows of 6 values (48 bytes) are read as three 16-byte
  loads; that stride puts consecutive rows on all eight 16-byte bank groups
  of a 32-bank LDS, and the short intervals keep the loads per pair at three.

Run: python massivemds_amd/csrc/gen_trunc_coeffs.py  (rewrites the header and
prints the maximum relative errors of the double-rounded coefficients).
"""
from pathlib import Path

import mpmath as mp

mp.mp.dps = 60
OUT = Path(__file__).resolve().parent / "mds_trunc_coeffs.h"


def cheb_nodes(a, b, n):
    return [(mp.cos(mp.pi * (2 * k + 1) / (2 * n)) + 1) / 2 * (b - a) + a for k in range(n)] + [mp.mpf(a), mp.mpf(b)]


def fit_rational(f, a, b, n, m, scale, iters=12):
    xs = cheb_nodes(a, b, 8 * (n + m + 2))
    fs = [f(x) for x in xs]
    us = [x / scale for x in xs]
    qprev = [mp.mpf(1)] * len(xs)
    for _ in range(iters):
        A = mp.matrix(len(xs), n + 1 + m)
        rhs = mp.matrix(len(xs), 1)
        for r, (u, fx) in enumerate(zip(us, fs)):
            w = 1 / (fx * qprev[r])
            for k in range(n + 1):
                A[r, k] = u ** k * w
            for k in range(1, m + 1):
                A[r, n + k] = -fx * u ** k * w
            rhs[r] = fx * w
        sol = mp.qr_solve(A, rhs)[0]
        p = [sol[k] for k in range(n + 1)]
        q = [mp.mpf(1)] + [sol[n + k] for k in range(1, m + 1)]
        qprev = [sum(q[k] * u ** k for k in range(m + 1)) for u in us]
    return p, q


def fit_poly(f, a, b, n, iters=1):
    xs = cheb_nodes(a, b, 8 * (n + 2))
    A = mp.matrix(len(xs), n + 1)
    rhs = mp.matrix(len(xs), 1)
    for r, x in enumerate(xs):
        fx = f(x)
        for k in range(n + 1):
            A[r, k] = x ** k / fx
        rhs[r] = 1
    sol = mp.qr_solve(A, rhs)[0]
    return [sol[k] for k in range(n + 1)]


def horner(c, x):
    v = mp.mpf(0)
    for k in reversed(range(len(c))):
        v = v * x + c[k]
    return v


def mills(z):
    return mp.erfc(z / mp.sqrt(2)) / 2 / (mp.exp(-z * z / 2) / mp.sqrt(2 * mp.pi))


def max_rel_err(fn, ref, a, b, n=4001):
    return max(abs(fn(x) / ref(x) - 1) for x in mp.linspace(a, b, n))


def to_double(c):
    return [mp.mpf(float(v)) for v in c]


def emit(name, c, ctype="double"):
    vals = ",\n    ".join(repr(float(v)) for v in c)
    return f"constexpr {ctype} {name}[{len(c)}] = {{\n    {vals}}};\n"


TAB_STEP = 64       # intervals per unit of z
TAB_ROWS = 641      # 640 fitted intervals on [0, 10) and one zero row
TAB_DEG = 5
TAB_STRIDE = 6      # doubles per row (an even count: rows are read as 16-byte pairs)


def log_phi(z):
    return mp.log(mp.ncdf(z))


def phi_over_phi(z):
    return mp.npdf(z) / mp.ncdf(z)


def fit_interval(k):
    m = 4 * (TAB_DEG + 1)
    ts = [(mp.cos(mp.pi * (2 * i + 1) / (2 * m)) + 1) / 2 for i in range(m)] + [mp.mpf(0), mp.mpf(1)]
    rows, rhs = [], []
    wd = mp.mpf(1) / TAB_DEG
    for t in ts:
        z = (k + t) / TAB_STEP
        rows.append([t ** j for j in range(TAB_DEG + 1)])
        rhs.append(log_phi(z))
        rows.append([wd * j * t ** (j - 1) if j else mp.mpf(0) for j in range(TAB_DEG + 1)])
        rhs.append(wd * phi_over_phi(z) / TAB_STEP)
    sol = mp.qr_solve(mp.matrix(rows), mp.matrix(rhs))[0]
    return [float(sol[j]) for j in range(TAB_DEG + 1)]


def table_errors(tab, samples=8):
    """Max |error| of log Phi and phi / Phi with the double coefficients
    evaluated in double arithmetic (value and derivative Horner, as the
    kernels do), over random points of every fitted interval."""
    import numpy as np

    rng = np.random.default_rng(7)
    ev = eh = 0.0
    for k in range(TAB_ROWS - 1):
        c = tab[k]
        for t in rng.uniform(0.0, 1.0, samples):
            p, q = c[TAB_DEG], 0.0
            for j in range(TAB_DEG - 1, -1, -1):
                q = q * t + p
                p = p * t + c[j]
            z = mp.mpf((k + t) / TAB_STEP)
            ev = max(ev, abs(p - float(log_phi(z))))
            eh = max(eh, abs(q * TAB_STEP - float(phi_over_phi(z))))
    return ev, eh


def emit_table(tab):
    lines = []
    for k, c in enumerate(tab):
        vals = ", ".join(repr(float(v)) for v in c) + ", 0.0" * (TAB_STRIDE - len(c))
        lines.append(f"    {vals},  // row {k}")
    return (f"alignas(16) constexpr double kLnPhiTab[{TAB_ROWS * TAB_STRIDE}] = {{\n" + "\n".join(lines)
            + "\n};\n")


def main():
    # Mills ratio, FP64 (8, 9) and FP32 (4, 5)
    p64, q64 = fit_rational(mills, 0, 12, 8, 9, 4)
    p32, q32 = fit_rational(mills, 0, 12, 4, 5, 4)
    e_m64 = max_rel_err(lambda z: horner(to_double(p64), z / 4) / horner(to_double(q64), z / 4), mills, 0, 12)
    e_m32 = max_rel_err(lambda z: horner(p32, z / 4) / horner(q32, z / 4), mills, 0, 12, 1201)
    # exp(-r), |r| <= ln2/2
    h = mp.log(2) / 2
    ex64 = fit_poly(lambda r: mp.exp(-r), -h, h, 11)
    e_ex = max_rel_err(lambda r: horner(to_double(ex64), r), lambda r: mp.exp(-r), -h, h)
    # atanh(s)/s in v = s^2, v in [0, 1/9]
    at = lambda v: mp.atanh(mp.sqrt(v)) / mp.sqrt(v) if v > 0 else mp.mpf(1)
    at64 = fit_poly(at, mp.mpf(0), mp.mpf(1) / 9, 9)
    e_at = max_rel_err(lambda v: horner(to_double(at64), v), at, mp.mpf(1e-30), mp.mpf(1) / 9)
    at32 = fit_poly(at, mp.mpf(0), mp.mpf(1) / 9, 5)
    e_at32 = max_rel_err(lambda v: horner(at32, v), at, mp.mpf(1e-30), mp.mpf(1) / 9, 1201)

    tab = [fit_interval(k) for k in range(TAB_ROWS - 1)] + [[0.0] * (TAB_DEG + 1)]
    e_tv, e_th = table_errors(tab)

    # ln 2 = hi + lo with hi holding 32 significant bits, so k * hi is exact
    # for |k| < 2^21
    ln2 = mp.log(2)
    hi = float(mp.floor(ln2 * 2 ** 32) / 2 ** 32)
    lo = float(ln2 - mp.mpf(hi))
    body = [
        "// Generated by gen_trunc_coeffs.py -- do not edit by hand.",
        "// Fitted from the defining formulas in mpmath (see the generator).",
        f"// max relative error with these double coefficients: Mills (8,9) {float(e_m64):.2e}, "
        f"exp(-r) deg 11 {float(e_ex):.2e}, atanh(s)/s deg 9 {float(e_at):.2e};",
        f"// FP32 forms (exact arithmetic): Mills (4,5) {float(e_m32):.2e}, atanh(s)/s deg 5 {float(e_at32):.2e}.",
        f"// log Phi table, double evaluation: max |error| {e_tv:.2e} (log Phi), {e_th:.2e} (phi / Phi).",
        "#pragma once",
        "",
        "namespace mdship {",
        "namespace coeffs {",
        "",
        f"constexpr double kLn2Hi = {hi!r};",
        f"constexpr double kLn2Lo = {lo!r};",
        f"constexpr double kLog2e = {float(1 / ln2)!r};",
        "constexpr double kMillsScale = 0.25;  // u = z / 4",
        "constexpr double kMillsZMax = 12.0;",
        emit("kMillsP64", p64),
        emit("kMillsQ64", q64),
        emit("kMillsP32", p32, "float"),
        emit("kMillsQ32", q32, "float"),
        emit("kExpNeg64", ex64),
        emit("kAtanh64", at64),
        emit("kAtanh32", at32, "float"),
        f"constexpr int kLnPhiTabStep = {TAB_STEP};   // intervals per unit of z",
        f"constexpr int kLnPhiTabRows = {TAB_ROWS};",
        f"constexpr int kLnPhiTabDeg = {TAB_DEG};",
        f"constexpr int kLnPhiTabStride = {TAB_STRIDE};  // doubles per row ({8 * TAB_STRIDE} bytes)",
        f"constexpr double kLnPhiTabClamp = {TAB_ROWS - 1}.5;  // {TAB_STEP} z is clamped here: the zero row",
        emit_table(tab),
        "}  // namespace coeffs",
        "}  // namespace mdship",
        "",
    ]
    OUT.write_text("\n".join(body))
    print(OUT)
    print(f"mills64 {float(e_m64):.3e} mills32 {float(e_m32):.3e} exp {float(e_ex):.3e} "
          f"atanh64 {float(e_at):.3e} atanh32 {float(e_at32):.3e} table {e_tv:.3e} / {e_th:.3e}")


if __name__ == "__main__":
    main()
