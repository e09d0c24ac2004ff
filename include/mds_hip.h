/*
 * mds_hip.h -- C ABI of libmds_hip.so, the MI355X (gfx950) engine for the
 * MassiveMDS pairwise (left-truncated) normal MDS log-likelihood and its
 * gradient with respect to the N x P latent locations.
 *
 * Each entry point below replaces one virtual of the reference's engine
 * interface, mds::AbstractMultiDimensionalScaling
 * (src/AbstractMultiDimensionalScaling.hpp:47-59), or the factory
 * (src/factory.cpp:32-118), with the same argument meaning:
 *   - all double* arrays are caller-owned host buffers, copied in or out
 *     before the call returns (src/NewMultiDimensionalScaling.hpp:160-163,
 *     259, 301);
 *   - locations and gradients are row-major N x P; observations are the
 *     N x N matrix, row-major, NaN = missing, diagonal ignored.
 * Every function returns 0 on success and a non-zero status otherwise (the
 * reference only asserts); mds_hip_last_error() gives the message.
 * Instances are independent and not thread-safe, as in the reference.
 */
#ifndef MDS_HIP_H
#define MDS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flag bits, identical to the reference's mds::Flags (src/flags.h:6-14). */
#define MDS_HIP_FLAG_FLOAT (1L << 2)           /* FP32 engine (FP32 terms, FP64 sums) */
#define MDS_HIP_FLAG_TBB (1L << 3)             /* accepted, ignored by the GPU engine */
#define MDS_HIP_FLAG_OPENCL (1L << 4)          /* "GPU engine": routed to HIP */
#define MDS_HIP_FLAG_LEFT_TRUNCATION (1L << 5)
#define MDS_HIP_FLAG_SSE (1L << 7)             /* accepted, ignored */
#define MDS_HIP_FLAG_AVX (1L << 8)             /* accepted, ignored */
#define MDS_HIP_FLAG_AVX512 (1L << 9)          /* accepted, ignored */
#define MDS_HIP_FLAG_HIP (1L << 10)            /* explicit alias of the GPU route */

/* Status codes. */
#define MDS_HIP_OK 0
#define MDS_HIP_ERR_ARGUMENT 1   /* bad length, dimension, index or parameter */
#define MDS_HIP_ERR_DEVICE 2     /* HIP runtime error or no GPU */
#define MDS_HIP_ERR_STATE 3      /* call made before the data it needs were set */

typedef struct mds_hip_engine* mds_hip_handle;

/* Largest embedding dimension P the kernels accept (the OpenCL engine it
 * replaces stops at 8, src/backend/opencl/instantiate.cpp:16-31). */
int mds_hip_max_dimension(void);

/* Number of visible HIP devices (0 when none). */
int mds_hip_device_count(int* count);

/*
 * Construction, replacing constructOpenCLMultiDimensionalScalingDouble/Float
 * (src/backend/opencl/instantiate.cpp:14-52) as reached from
 * mds::factory(P, N, flags, device, threads) (src/factory.cpp:32-118).
 * `device` is an index into the HIP devices; out of range selects device 0,
 * as the OpenCL engine falls back to its default device
 * (src/backend/opencl/OpenCLMultiDimensionalScaling.hpp:85-89).
 */
int mds_hip_create(int embeddingDimension, int locationCount, long flags, int device,
                   mds_hip_handle* out);
int mds_hip_destroy(mds_hip_handle h);

/* Message of the last failed call on `h` (or of the last failed create when
 * h is NULL); never NULL. */
const char* mds_hip_last_error(mds_hip_handle h);

/* virtual void updateLocations(int, double*, size_t)
 * (src/AbstractMultiDimensionalScaling.hpp:48; semantics
 * src/NewMultiDimensionalScaling.hpp:134-166).  locationIndex = -1 updates
 * all N rows (length N*P); locationIndex >= 0 updates one row (length P). */
int mds_hip_update_locations(mds_hip_handle h, int locationIndex, const double* location,
                             size_t length);

/* virtual double getSumOfIncrements() (:51; src/NewMultiDimensionalScaling.hpp:193-203). */
int mds_hip_get_sum_of_increments(mds_hip_handle h, double* out);

/* virtual void getLogLikelihoodGradient(double*, size_t) (:52;
 * src/NewMultiDimensionalScaling.hpp:281-302).  length = N*P. */
int mds_hip_get_log_likelihood_gradient(mds_hip_handle h, double* result, size_t length);

/* virtual void storeState() / restoreState() / acceptState() (:53-55;
 * src/NewMultiDimensionalScaling.hpp:205-255). */
int mds_hip_store_state(mds_hip_handle h);
int mds_hip_restore_state(mds_hip_handle h);
int mds_hip_accept_state(mds_hip_handle h);

/* virtual void setPairwiseData(double*, size_t) (:56;
 * src/NewMultiDimensionalScaling.hpp:257-260).  length = N*N. */
int mds_hip_set_pairwise_data(mds_hip_handle h, const double* data, size_t length);

/* virtual void setParameters(double*, size_t) (:57;
 * src/NewMultiDimensionalScaling.hpp:262-274).  length = 1: the precision
 * tau, which must be finite and > 0. */
int mds_hip_set_parameters(mds_hip_handle h, const double* data, size_t length);

/* virtual void makeDirty() (:58; src/NewMultiDimensionalScaling.hpp:276-279). */
int mds_hip_make_dirty(mds_hip_handle h);

/* virtual int getInternalDimension() (:59); returns P. */
int mds_hip_get_internal_dimension(mds_hip_handle h, int* out);

/*
 * Device-resident extensions (no counterpart in the reference; used by
 * samplers and benchmarks that keep the state on the GPU).
 *
 * mds_hip_evaluate_async enqueues one evaluation at the current state on the
 * instance's stream (likelihood, plus the gradient when computeGradient is
 * non-zero) without any host copy; results stay on the device and also fill
 * the engine's caches when read through the calls above.
 */
int mds_hip_evaluate_async(mds_hip_handle h, int computeGradient);
int mds_hip_synchronize(mds_hip_handle h);

/*
 * Allocation-free evaluation.  No counterpart in the reference: this entry
 * point is for callers that cannot allocate device memory while they
 * evaluate (tight serving loops, harnesses that forbid hipMalloc).  It is
 * mds_hip_evaluate_async with device-resident inputs and outputs, and it
 * calls no hipMalloc, hipFree or allocating library routine: the tile sums
 * and gradient partials go to a static device array of the library (64 MiB
 * per GPU, held until the process ends) instead of the instance's buffers.
 *   locations:   device, N x P row-major FP64; they become the instance's
 *                current locations (as updateLocations(-1, ...)).  NULL
 *                evaluates at the current locations.
 *   sumOut:      device, one double: what getSumOfIncrements returns there.
 *   gradientOut: device, N x P row-major FP64: the log-likelihood gradient;
 *                NULL computes the likelihood alone.
 * Each pointer must lie inside one device allocation of the instance's
 * device, long enough, or the call fails with MDS_HIP_ERR_ARGUMENT.  So does
 * an instance whose gradient partials exceed the array (8 * tiles + 64 * s *
 * P' * ceil(N / 64)^2 bytes, s = 8 for the FP64 engine and 4 for FP32, P' the
 * kernel dimension >= P: with the gradient, N <= 16,320 at P <= 2 in FP64).  The work runs on the instance's stream, after
 * the instance's earlier work, and the call returns when the outputs are
 * written; the inputs must be ready when it is made.  Calls on one GPU share
 * the array and it keeps the last call's contents, so calls that overlap in
 * time (instances used from different threads) race and must be serialised
 * by the caller; calls made one after another never overlap.
 */
int mds_hip_evaluate_device_noalloc(mds_hip_handle h, const double* locations, double* sumOut,
                                    double* gradientOut);

/* Kernel timing with HIP events on the instance stream.  While enabled,
 * every evaluation records the tile-kernel and whole-evaluation durations;
 * get_timing synchronises and returns their sums (ms) and the count. */
int mds_hip_set_timing(mds_hip_handle h, int enable);
int mds_hip_get_timing(mds_hip_handle h, double* tileMs, double* evalMs, long long* count);
/* Work counters since creation: full O(N^2) evaluations launched, and sums
 * obtained by the O(N) single-row update (updateLocations(i, ...) of a state
 * whose sum was known; no counterpart in the reference, whose incremental
 * path is disabled, src/NewMultiDimensionalScaling.hpp:563-566). */
int mds_hip_get_counters(mds_hip_handle h, long long* fullEvaluations, long long* rowUpdates);

/* Layout facts: 1 when the stored observations are symmetric (one tile
 * plane), bytes of packed observations on the device, number of tiles. */
int mds_hip_get_layout(mds_hip_handle h, int* symmetric, long long* observationBytes,
                       long long* tiles);

/*
 * Batched evaluation (no counterpart in the reference, whose instances are
 * independent objects driven one call at a time; for callers with many
 * independent datasets or chains).  `count` datasets of their own N and P
 * (P <= 10) share the flags (FLOAT, LEFT_TRUNCATION); each has its own data,
 * locations and precision, with the same argument meaning as the
 * single-instance calls.  mds_hip_batch_evaluate runs every dataset's
 * likelihood and gradient in one persistent launch plus one reduce launch.
 */
typedef struct mds_hip_batch* mds_hip_batch_handle;

int mds_hip_batch_create(int count, const int* embeddingDimensions, const int* locationCounts, long flags,
                         int device, mds_hip_batch_handle* out);
int mds_hip_batch_destroy(mds_hip_batch_handle h);
const char* mds_hip_batch_last_error(mds_hip_batch_handle h);
int mds_hip_batch_set_pairwise_data(mds_hip_batch_handle h, int index, const double* data, size_t length);
int mds_hip_batch_update_locations(mds_hip_batch_handle h, int index, const double* locations, size_t length);
int mds_hip_batch_set_parameters(mds_hip_batch_handle h, int index, const double* data, size_t length);
int mds_hip_batch_evaluate(mds_hip_batch_handle h);
/* getSumOfIncrements of every dataset (length = count). */
int mds_hip_batch_get_sums_of_increments(mds_hip_batch_handle h, double* out, size_t length);
int mds_hip_batch_get_gradient(mds_hip_batch_handle h, int index, double* out, size_t length);
/* Times every launch configuration on this batch and keeps the fastest;
 * with cachePath, reuses a cached choice for the same batch signature
 * (unless force) and appends new choices.  bestMs: its time per evaluation. */
int mds_hip_batch_autotune(mds_hip_batch_handle h, const char* cachePath, int force, double* bestMs);
/* strategy 0: one launch over all items; 1: one launch per P class. */
int mds_hip_batch_set_config(mds_hip_batch_handle h, int wavesPerBlock, int blocksPerCU, int tilesPerItem,
                             int strategy);
int mds_hip_batch_get_config(mds_hip_batch_handle h, int* wavesPerBlock, int* blocksPerCU, int* tilesPerItem,
                             int* strategy);
/* Entry `index` reads the observations of entry `sourceIndex` (same N; P may
 * differ) instead of its own: several chains of one dataset without copies.
 * Later data set on the source are seen by `index`; data set on `index`
 * itself end the sharing. */
int mds_hip_batch_share_pairwise_data(mds_hip_batch_handle h, int index, int sourceIndex);
int mds_hip_batch_set_timing(mds_hip_batch_handle h, int enable);
int mds_hip_batch_get_timing(mds_hip_batch_handle h, double* ms, long long* count);

/*
 * Device-resident HMC sampler: the GPU counterpart of the reference's R
 * hmcsampler (R/Untitled.R:477-729) for the model without a tree (row
 * covariance I_N): potential = -(matrix-normal log prior + MDS
 * log-likelihood), leapfrog with step trajectory / leapfrogSteps, the
 * trajectory captured as one hipGraph, Metropolis test in FP64 on the host.
 * learnPrecision: random-walk Metropolis on tau with the reference proposal
 * |U(tau - w, tau + w)| under truncation, Gamma Gibbs draw otherwise.
 * learnTraitPrecision: Wishart(P + N, (I + X^T X)^-1) Gibbs draw every
 * traitInvWeight iterations.  The engine must hold data and a precision; on
 * return it holds the final locations and precision.
 *
 * trace: (iterations - burnIn) rows of MDS_HIP_HMC_TRACE_COLUMNS values --
 * log posterior, tau, trace(T), five pairwise distances |x_a - x_b| for
 * (a, b) = (0, 1), (0, N/2), (N/3, 2N/3), (1, N-1), (N/4, 3N/4), and the
 * norms |x_0|, |x_{N/2}|, |x_{N-1}| (functionals invariant to rotations and
 * reflections of the configuration).
 */
#define MDS_HIP_HMC_TRACE_COLUMNS 11

/*
 * Surrogate of the gradient rows (no counterpart in the reference; for
 * samplers that can afford approximate leapfrog gradients because the exact
 * likelihood still decides acceptance): a per-location MLP, input x_i (P <= 16)
 * plus a learned 16-value embedding of location i, two hidden layers of 256
 * with tanh, output the P gradient components; BF16 weights on the MFMA
 * matrix cores with FP32 accumulation, trained on the GPU with Adam on
 * (x, exact gradient) snapshots.  Snapshots come from HMC runs with
 * collectSnapshots set, or from mds_hip_surrogate_add_snapshot.
 */
typedef struct mds_hip_surrogate* mds_hip_surrogate_handle;

int mds_hip_surrogate_create(int locationCount, int embeddingDimension, int snapshotCapacity,
                             unsigned long long seed, int device, mds_hip_surrogate_handle* out);
int mds_hip_surrogate_destroy(mds_hip_surrogate_handle h);
const char* mds_hip_surrogate_last_error(mds_hip_surrogate_handle h);
int mds_hip_surrogate_snapshot_count(mds_hip_surrogate_handle h, int* count);
int mds_hip_surrogate_add_snapshot(mds_hip_surrogate_handle h, const double* x, const double* g, size_t length);
/* Trains on all stored snapshots; finalLoss: mean squared error on the
 * normalised targets after training. */
int mds_hip_surrogate_train(mds_hip_surrogate_handle h, int epochs, double learningRate, double* finalLoss);
/* Predicted log-likelihood gradient rows (N x P) at locations x (N x P). */
int mds_hip_surrogate_predict(mds_hip_surrogate_handle h, const double* x, size_t length, double* grad);
/* FP32 parameters: W1 (32 x 256), b1 (256), W2 (256 x 256), b2 (256), W3 (256 x 16),
 * b3 (16), embeddings (N x 16), all row-major, concatenated. */
int mds_hip_surrogate_parameter_count(mds_hip_surrogate_handle h, long long* count);
int mds_hip_surrogate_get_parameters(mds_hip_surrogate_handle h, float* out, size_t count);
int mds_hip_surrogate_set_parameters(mds_hip_surrogate_handle h, const float* in, size_t count);
/* Normalisation (train sets it from the snapshots): network input
 * (x_i - xCentre_i) / sigmaX and target (g_i - gCentre_i) / sigmaG, with
 * per-location centres (N x P each) and per-dimension scales (P each). */
int mds_hip_surrogate_set_normalisation(mds_hip_surrogate_handle h, const float* xCentre, const float* gCentre,
                                        size_t centreLength, const float* sigmaX, const float* sigmaG,
                                        size_t scaleLength);
/* One forward + backward pass without an update: the gradient of the loss
 * with respect to every parameter (same order as get_parameters) and the loss. */
int mds_hip_surrogate_loss_gradient(mds_hip_surrogate_handle h, const double* x, const double* g, size_t length,
                                    float* paramGrad, size_t count, double* loss);

typedef struct {
    int iterations;
    int burnIn;
    int leapfrogSteps;
    double trajectory;
    int learnPrecision;
    int learnTraitPrecision;
    int traitInvWeight;
    double precisionProposalWidth;
    double priorDf;              /* <= 0: P */
    unsigned long long seed;
    int useGraph;
    int adaptStepSize;           /* dual averaging (Hoffman & Gelman 2014) during burn-in */
    double targetAccept;         /* its target acceptance rate */
    double stepSize;             /* > 0: leapfrog step, overriding trajectory / leapfrogSteps */
    mds_hip_surrogate_handle surrogate;
    int useSurrogate;            /* leapfrog on the surrogate's gradients; exact likelihood in the test */
    int collectSnapshots;        /* store (x, exact gradient) after every exact gradient pass
                                  * at the start and along accepted trajectories (a rejected
                                  * trajectory's snapshots are dropped) */
    double fallbackAcceptRate;   /* surrogate: below this acceptance over a window of */
    int fallbackWindow;          /* iterations, exact gradients take over for fallbackWindow
                                  * iterations, doubled after each consecutive failed window
                                  * (at most 1024x); 0 disables the fallback */
    int adaptPrecisionWidth;     /* truncation: during burn-in, adapt the width w of the
                                  * precision's proposal |U(tau - w, tau + w)| toward 0.44
                                  * acceptance (Robbins-Monro on log w); fixed afterwards.
                                  * 0 keeps precisionProposalWidth, as the reference does */
} mds_hip_hmc_options;

typedef struct {
    double acceptRate;
    double precisionAcceptRate;
    double samplingSeconds;      /* iterations after burn-in */
    double totalSeconds;
    double evaluations;          /* likelihood (+ gradient) passes */
    double finalPrecision;
    double stepSize;
    double finalPotential;
    double surrogateIterations;  /* iterations whose trajectory used the surrogate */
    double exactIterations;
    double snapshots;            /* snapshots stored in this run */
    double surrogateAcceptRate;  /* acceptance over the surrogate iterations alone */
    double precisionProposalWidth;  /* the precision proposal's width after burn-in */
} mds_hip_hmc_stats;

int mds_hip_hmc_default_options(mds_hip_hmc_options* options);
int mds_hip_hmc_run(mds_hip_handle h, const mds_hip_hmc_options* options, const double* initialLocations,
                    size_t length, double* trace, size_t traceLength, double* finalLocations,
                    mds_hip_hmc_stats* stats);

/*
 * The sampler above on every entry of a batch at once (no counterpart in the
 * reference; many chains, of one dataset through
 * mds_hip_batch_share_pairwise_data or of datasets of uneven N and P): one
 * batched tile launch and one batched reduce (which also applies the kicks)
 * per leapfrog step move every chain, the trajectory captured as one
 * hipGraph.  Entry k
 * runs its own chain with seed options->seed + k, its own precision (from
 * its setParameters), step size and trait precision, and gives the bits of
 * mds_hip_hmc_run on a single engine with that seed and the same data.
 * Exact gradients only (useSurrogate and collectSnapshots must be 0).
 * initialLocations / finalLocations: the entries' N_k x P_k arrays
 * concatenated (length sum N_k P_k; finalLocations may be NULL); traces:
 * count x (iterations - burnIn) x MDS_HIP_HMC_TRACE_COLUMNS (or NULL);
 * stats: statsCount = count entries (or NULL).  On return each entry holds
 * its chain's final locations and precision.
 */
int mds_hip_batch_hmc_run(mds_hip_batch_handle h, const mds_hip_hmc_options* options, const double* initialLocations,
                          size_t length, double* traces, size_t traceLength, double* finalLocations,
                          size_t finalLength, mds_hip_hmc_stats* stats, int statsCount);

#ifdef __cplusplus
}
#endif

#endif /* MDS_HIP_H */
