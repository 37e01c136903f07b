/* batchfit_hip.h — C ABI of the MI355X (gfx950) batched bounded
 * Levenberg–Marquardt fitter, libbatchfit_hip.so.
 *
 * This is the drop-in seam below the reference's GPU backend: it takes the
 * place of metal::detail::run_gpu / submit_gpu (include/batchfit/metal/batch.hpp:90-92,
 * src/metal/runtime.cpp:165-316) with the same inputs (shared x[M], one row of
 * y per fit, initial parameters, the 108-byte Config of
 * include/batchfit/metal/types.hpp:7-15) and the same outputs (params, status,
 * iterations, chi2, optional covariance/stderr/uncertainty status). Callers:
 * the C++ façade batchfit::fit (include/batchfit/fit.hpp) and the Python
 * binding batchfit_amd.fit(..., device="hip").
 *
 * Conventions
 *  - Plain pointers and sizes only; no C++ types or exceptions cross the ABI.
 *  - Every entry returns BF_OK (0) or an error code and writes a message into
 *    err[errlen] when err is not NULL. Per-fit numerical problems are status
 *    codes, never errors, as in the reference's batch contract
 *    (cpu/batch.hpp:23-65: one FitResult per row, whatever the row holds).
 *  - Status codes are cpu/solver.hpp:16-23; uncertainty status codes are
 *    uncertainty_types.hpp:3-10; both int32. Iterations are uint32.
 *  - y is fit-major (row i = fit i, M contiguous floats) unless stated.
 *  - A NULL p0 runs the data-only initial guess on the device
 *    (cpu/initial_guess.hpp:120-249, called by python/bindings.cpp:69-80).
 */
#ifndef BATCHFIT_HIP_H
#define BATCHFIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BF_HIP_ABI_VERSION 1

enum bf_error {
    BF_OK = 0,
    BF_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    BF_RUNTIME_ERROR = 2,    /* a HIP call failed; message names it */
    BF_UNAVAILABLE = 3,      /* no gfx950 device visible */
    BF_TIMEOUT = 4           /* bf_serve_wait: not finished within the timeout */
};

enum bf_model { BF_GAUSSIAN = 0, BF_LORENTZIAN = 1, BF_PSEUDO_VOIGT = 2, BF_TWO_LORENTZIAN = 3 };

/* numerics bits (bit0/bit1 as Config::numerics, src/metal/lm.metal:8-10) */
enum bf_numerics {
    BF_COMPENSATED = 1u,    /* Kahan per-lane partials */
    BF_SCALED_CHOLESKY = 2u,
    BF_MIXED_FP64 = 4u      /* FP64 group combine, solve, accept test, covariance */
};

/* Byte layout of metal::detail::Config (include/batchfit/metal/types.hpp:7-15),
 * 108 bytes. minimum_spacing, varied_x and width_factor are derived by the
 * library from x and the model (batch.hpp:236-244) and ignored on input. */
typedef struct bf_config {
    uint32_t fits, points, max_iterations;
    float initial_lambda, gradient_tolerance, step_tolerance, cost_tolerance;
    float lower[8], upper[8];
    uint32_t numerics;
    float minimum_spacing;
    uint32_t varied_x;
    float width_factor;
} bf_config;

/* Fit-to-hardware mapping (replaces metal::KernelVariant, batch.hpp:19). */
enum bf_mapping {
    BF_MAP_AUTO = 0,  /* tuned table by (model, M) */
    BF_MAP_LANE = 1,  /* one lane per fit, y streamed (point-major when transposed) */
    BF_MAP_GROUP = 2, /* group_lanes lanes per fit, y in registers or streamed (see streamed) */
    BF_MAP_LATENCY = 3 /* by (model, M, N): a batch too small to fill the device in the
                          tuned mapping takes 64, 128 or 256 lanes per fit (shorter latency;
                          results then are those of that mapping), else as BF_MAP_AUTO */
};

typedef struct bf_launch {
    uint32_t mapping;     /* bf_mapping */
    uint32_t group_lanes; /* 4, 8, 16, 32 or 64 lanes of a wave for BF_MAP_GROUP (4 always
                             streams), or 128 / 256 for a workgroup (2 or 4 waves) per fit
                             (src/metal/cooperative.metal:16-32, KernelVariant::ThreadgroupPerFit,
                             include/batchfit/metal/batch.hpp:19); 0 = tuned */
    uint32_t point_major; /* BF_MAP_LANE only: 1 = transpose y on the device first */
    uint32_t streamed;    /* BF_MAP_GROUP only: 1 = read the points from memory on every
                             pass instead of holding them in registers (same results) */
    void *stream;         /* hipStream_t; NULL = the library's stream */
} bf_launch;

/* One fit of a ragged batch (mixed model, M and options in one launch). */
typedef struct bf_fit_desc {
    uint64_t y_offset; /* index of the fit's first observation in y */
    uint32_t points;   /* M of this fit; must equal its grid's points */
    uint32_t grid;     /* row of the grid table: the fit's x */
    uint32_t model;    /* bf_model */
    uint32_t config;   /* row of the config table (bounds, tolerances, numerics) */
    uint32_t p0_row;   /* row of p0 (rows of 8 floats), or 0xffffffff for the device guess */
    uint32_t out_row;  /* row of the outputs: params/stderr rows of 8, cov rows of 64 */
} bf_fit_desc;

/* One shared grid of a ragged batch: x[x_offset .. x_offset+points). */
typedef struct bf_grid_desc {
    uint64_t x_offset;
    uint32_t points;
    uint32_t reserved;
} bf_grid_desc;

/* Library and device facts. */
int bf_hip_abi_version(void);
/* 1 if a gfx950 device is usable (bindings.cpp:208-211 metal_available). */
int bf_hip_available(void);
/* Fill cfg with the reference defaults (cpu/solver.hpp:26-41, float): 100
 * iterations, lambda 1e-3, tolerances 1e-5/1e-6/1e-6, unbounded, numerics 0. */
void bf_hip_default_config(bf_config *cfg);

/* Synchronous fit on host arrays (python/bindings.cpp:60-121 compute(),
 * batchfit::fit_batch include/batchfit/metal/batch.hpp:268-284).
 * x[M], y[N*M] fit-major, p0[N*P] or NULL; outputs params[N*P], status[N],
 * iterations[N], chi2[N]; cov[N*P*P], stderr_[N*P], ustatus[N] all NULL or
 * all non-NULL (uncertainties, cpu/uncertainty.hpp:13-110). launch may be NULL. */
int bf_hip_fit(uint32_t model, const bf_config *cfg, const bf_launch *launch, const float *x,
               const float *y, const float *p0, float *params, int32_t *status,
               uint32_t *iterations, float *chi2, float *cov, float *stderr_, int32_t *ustatus,
               char *err, size_t errlen);

/* Asynchronous fit on device arrays already resident in HBM, enqueued on
 * launch->stream (submit_batch, include/batchfit/metal/batch.hpp:253-264).
 * Same arguments as bf_hip_fit with device pointers. Needs no host copy; its
 * scratch is taken stream-ordered from the device pool. */
int bf_hip_fit_device(uint32_t model, const bf_config *cfg, const bf_launch *launch,
                      const float *d_x, const float *d_y, const float *d_p0, float *d_params,
                      int32_t *d_status, uint32_t *d_iterations, float *d_chi2, float *d_cov,
                      float *d_stderr, int32_t *d_ustatus, char *err, size_t errlen);

/* bf_hip_fit_device without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same arguments and results as bf_hip_fit_device; its scratch (grid facts,
 * work counter, device-guess rows) is one static device array per GPU that
 * holds the guesses of at most 32 MiB / (4·P) fits (2,097,152 of P=4) when p0
 * is NULL, and any N when p0 is given. Larger N, and the point-major
 * transpose of BF_MAP_LANE, are rejected with BF_INVALID_ARGUMENT; no kernel
 * timing is recorded. Calls that overlap on one GPU share that array and race:
 * serialise them (one stream, or events). */
int bf_hip_fit_device_noalloc(uint32_t model, const bf_config *cfg, const bf_launch *launch,
                              const float *d_x, const float *d_y, const float *d_p0,
                              float *d_params, int32_t *d_status, uint32_t *d_iterations,
                              float *d_chi2, float *d_cov, float *d_stderr, int32_t *d_ustatus,
                              char *err, size_t errlen);

/* Ragged batches: fits of any model and M, each with its own config row, in
 * one hipGraph launch (persistent kernels fed by device work queues).
 * bf_hip_ragged_create takes HOST descriptor tables, validates them, buckets
 * the fits by (model, mapping class), orders every class longest-first and
 * uploads everything; numerics must be 0 (FP32) or 5 (compensated mixed).
 * bf_hip_ragged_run enqueues the plan on its stream over DEVICE data: x
 * (grids packed), y (observations packed), p0 rows of 8 (or NULL when no
 * descriptor names a p0 row) and outputs with rows of 8 (params, stderr), 64
 * (cov) or 1 (status, iterations, chi2, ustatus); cov/stderr/ustatus only when
 * the plan was created with uncertainties. The captured graph is reused while
 * the pointers stay the same. No counterpart in the reference (it fits one
 * uniform batch per call); each fit is solved as a uniform batch of one would be. */
typedef struct bf_ragged_plan_s *bf_ragged_plan;
int bf_hip_ragged_create(bf_ragged_plan *plan, uint32_t fits, const bf_fit_desc *fit_table,
                         uint32_t grids, const bf_grid_desc *grid_table, uint32_t configs,
                         const bf_config *config_table, uint32_t numerics, int uncertainties,
                         void *stream, char *err, size_t errlen);
int bf_hip_ragged_run(bf_ragged_plan plan, const float *d_x, const float *d_y, const float *d_p0,
                      float *d_params, int32_t *d_status, uint32_t *d_iterations, float *d_chi2,
                      float *d_cov, float *d_stderr, int32_t *d_ustatus, char *err,
                      size_t errlen);
/* Device time of the last run of the plan (ms), measured by events recorded
 * on the plan's stream around the graph launch: [0] whole graph; [1..] 0. */
int bf_hip_ragged_timing(bf_ragged_plan plan, double *times, int count);
int bf_hip_ragged_destroy(bf_ragged_plan plan);

/* Continuous-batching server. Any host thread submits requests (one model,
 * one grid x[M], N spectra y[N*M], optional p0[N*P], its own bounds and
 * options in cfg with cfg->fits = N, cfg->points = M); a scheduler thread
 * admits pending requests in arrival order into free device slots while the
 * bytes held by admitted batches stay within hbm_budget, and runs each
 * admitted batch as one ragged launch on that slot's stream, so up to
 * `slots` batches overlap on the device. Host-array requests are staged
 * through pinned memory (inputs are borrowed until bf_serve_wait returns);
 * on_device requests name device arrays and receive device outputs (no
 * uncertainties). A request's result equals the same request alone. No
 * counterpart in the reference (it serialises fitting behind one mutex,
 * python/bindings.cpp:13,66, and streams NPY through a 3-slot ring,
 * src/metal/pipeline.cpp:99-169). */
typedef struct bf_server_s *bf_server;
typedef struct bf_serve_options {
    uint64_t hbm_budget;     /* bytes admitted batches may hold; 0 = 64 GiB */
    uint32_t slots;          /* batches on the device at once; 0 = 3 (max 16) */
    uint32_t max_batch_fits; /* fits per admitted batch; 0 = 4,194,304 */
    uint32_t numerics;       /* 0 (FP32) or 5 (compensated mixed, default) */
    uint32_t reserved;
} bf_serve_options;
typedef struct bf_request_info {
    int32_t rc;        /* BF_OK or the batch's error */
    uint32_t batch;    /* admitted batch number */
    double queued_s;   /* submit → admission */
    double service_s;  /* admission → completion */
    double latency_s;  /* submit → completion */
    char message[256];
} bf_request_info;
typedef struct bf_serve_counters {
    uint64_t requests_done, fits_done, batches, pending;
    uint64_t bytes_in_flight, max_bytes_in_flight;
    uint32_t max_batches_in_flight, reserved;
} bf_serve_counters;
int bf_serve_create(bf_server *server, const bf_serve_options *options, char *err, size_t errlen);
int bf_serve_submit(bf_server server, uint32_t model, const bf_config *cfg, int on_device,
                    const float *x, const float *y, const float *p0, float *params,
                    int32_t *status, uint32_t *iterations, float *chi2, float *cov,
                    float *stderr_, int32_t *ustatus, uint64_t *ticket, char *err, size_t errlen);
/* Wait for a request (timeout_s < 0 waits forever); returns the request's rc
 * or BF_TIMEOUT. The ticket is released once it returns anything but BF_TIMEOUT. */
int bf_serve_wait(bf_server server, uint64_t ticket, double timeout_s, bf_request_info *info);
int bf_serve_stats(bf_server server, bf_serve_counters *stats);
/* Drain every admitted and pending request, then stop the threads. */
int bf_serve_destroy(bf_server server);

/* NPY inputs and the streaming pipeline (metal::fit_npy / stream_npy,
 * include/batchfit/metal/pipeline.hpp:49-67, src/metal/pipeline.cpp:70-265;
 * NPY rules of src/cpu/npy.cpp:51-84). Inputs are NPY v1.0 files, C order,
 * little-endian float32 or float64: x[M], y[N,M] and optionally p0[N,P].
 * Chunks of fits are mapped from the file, staged through a ring of pinned
 * slots, narrowed to float32 on the device when stored as float64, fitted
 * with the device guess (or p0) on the slot's stream, and handed to
 * `consume` in order of first_fit; a nonzero return from `consume` stops the
 * pipeline with BF_RUNTIME_ERROR. Output pointers passed to `consume` are
 * valid only during the call (params [fits,P], cov [fits,P,P] row-major). */
typedef struct bf_npy_options {
    uint64_t chunk_fits;    /* fits per chunk; 0 = from budget_bytes */
    uint32_t in_flight;     /* chunks in flight, 1..8 (the reference allows 1..3) */
    int32_t uncertainties;  /* also compute covariance, stderr, status */
    uint64_t budget_bytes;  /* bytes the ring may hold per fit column; 0 = 1 GiB */
} bf_npy_options;
typedef struct bf_npy_stats {
    uint64_t fits, chunks, chunk_fits, max_in_flight, max_mapped_bytes, shared_payload_bytes;
    double wall_seconds, prepare_seconds, device_seconds, readback_seconds;
} bf_npy_stats;
typedef int (*bf_npy_consumer)(void *user, uint64_t first_fit, uint64_t fits, const float *params,
                               const int32_t *status, const uint32_t *iterations, const float *chi2,
                               const float *cov, const float *stderr_, const int32_t *ustatus);
/* Header of an NPY file: shape[2] (shape[1] = 0 for rank 1), rank, dtype
 * (0 float32, 1 float64, 2 uint32), payload offset (npy::inspect,
 * include/batchfit/npy.hpp:19). Needs no device. */
int bf_npy_inspect(const char *path, uint64_t *shape, uint32_t *rank, uint32_t *dtype,
                   uint64_t *offset, char *err, size_t errlen);
/* Fits per chunk for these dimensions (choose_chunk_fits,
 * src/metal/pipeline.cpp:28-68): chunk_fits when set, else the budget over
 * the ring's bytes per fit, rounded down to a multiple of 256. 0 if invalid. */
uint64_t bf_npy_chunk_fits(uint64_t fits, uint64_t points, uint32_t parameters,
                           uint32_t input_bytes, uint32_t p0_bytes, const bf_npy_options *options);
int bf_hip_fit_npy(uint32_t model, const char *x_path, const char *y_path, const char *p0_path,
                   const bf_config *cfg, const bf_npy_options *options, bf_npy_consumer consume,
                   void *user, bf_npy_stats *stats, char *err, size_t errlen);

/* Noise of bf_hip_synthesize. The SNR modes are those of the reference's
 * generator (src/cpu/synthetic.cpp:83-108): Gaussian with σ = RMS of the
 * clean row / snr, and Poisson counts y = Poisson(gain·f)/gain with gain =
 * snr²·Σf/Σf² (a row whose clean signal is negative anywhere, or a point
 * whose expected count exceeds 1e12, is written as NaN where the reference
 * raises). */
enum bf_noise {
    BF_NOISE_ABSOLUTE = 0,     /* σ = noise */
    BF_NOISE_RELATIVE = 1,     /* σ = noise · amplitude (parameter 0) */
    BF_NOISE_SNR_GAUSSIAN = 2, /* noise = snr */
    BF_NOISE_SNR_POISSON = 3   /* noise = snr */
};

/* Seeded synthetic spectra written straight into HBM (benchmarks, serving
 * load): y[n*m] (fit-major) of `model` on the device grid d_x[m], parameters
 * per fit drawn uniformly from [lower[k], upper[k]) (8 entries), noise by
 * noise_mode (bf_noise; 0 with noise = 0 is the noise-free signal). layout 1
 * (two_lorentzian) draws one HWHM for both peaks and sets c2 = c1 +
 * U[sep_lo, sep_hi)·HWHM. Philox4x32-10 counters (seed, fit_offset + fit,
 * point) make every fit independent of how the batch is split. d_truth[n*8]
 * receives the drawn parameters when not NULL. Enqueued on `stream`. No
 * counterpart in the reference's hot path (its generator is
 * src/cpu/synthetic.cpp on the host). */
int bf_hip_synthesize(uint32_t model, uint32_t n, uint32_t m, const float *d_x,
                      const float *lower, const float *upper, uint32_t layout, float sep_lo,
                      float sep_hi, uint32_t noise_mode, float noise, uint64_t seed,
                      uint64_t fit_offset, float *d_y, float *d_truth, void *stream, char *err,
                      size_t errlen);

/* Data-only initial guess alone, host arrays (cpu/initial_guess.hpp:120-249):
 * params[N*P], baseline[N], gstatus[N] (GuessStatus codes 0..3). */
int bf_hip_guess(uint32_t model, uint32_t n, uint32_t m, const float *x, const float *y,
                 float *params, float *baseline, int32_t *gstatus, char *err, size_t errlen);

/* Guess kernels (every form gives the same bits): */
enum bf_guess_form {
    BF_GUESS_AUTO = 0, /* the fastest that applies (bf_hip_guess) */
    BF_GUESS_WAVE = 1, /* one wave per fit, the signal staged in LDS (M <= 16,384; 15,886 two-peak) */
    BF_GUESS_WIDE = 2  /* one wave per fit, the signal read from memory (any M) */
};
/* bf_hip_guess by a chosen kernel, for tests that hold the forms to each
 * other. No reference counterpart (the reference has one CPU guess,
 * cpu/initial_guess.hpp:120-249). */
int bf_hip_guess_form(uint32_t model, uint32_t form, uint32_t n, uint32_t m, const float *x,
                      const float *y, float *params, float *baseline, int32_t *gstatus, char *err,
                      size_t errlen);

/* Device model evaluation for parity tests (tests/models_metal.mm:19-60):
 * value[i] and jac[i*P..] of model at (x[i], params[i*8..]) for i < n. */
int bf_hip_model_eval(uint32_t model, uint32_t n, const float *x, const float *params,
                      float *value, float *jac, char *err, size_t errlen);

/* The mapping a bf_hip_fit / bf_hip_fit_device call with these arguments
 * uses, without launching anything (no device needed): lanes per fit (1 =
 * one lane per fit) and the register tile in points per lane (0 = the points
 * are read from memory on every pass). No reference counterpart: the
 * reference's kernel choice is internal (include/batchfit/metal/batch.hpp:21-50). */
int bf_hip_plan(uint32_t model, uint32_t points, uint32_t numerics, const bf_launch *launch,
                uint32_t *group_lanes, uint32_t *tile_points, char *err, size_t errlen);
/* bf_hip_plan for a batch of `fits` spectra (BF_MAP_LATENCY depends on it;
 * bf_hip_plan is this with fits = 0). */
int bf_hip_plan_batch(uint32_t model, uint32_t points, uint32_t fits, uint32_t numerics,
                      const bf_launch *launch, uint32_t *group_lanes, uint32_t *tile_points,
                      char *err, size_t errlen);

/* Enable (1) or disable (0) kernel timing with HIP events; also enabled by the
 * environment variable BF_HIP_TIMING. */
int bf_hip_set_timing(int enable);
/* Kernel timing of the most recent call on this thread, in milliseconds,
 * measured with HIP events on the stream the kernels ran on:
 * times[0] grid prologue, [1] guess, [2] transpose, [3] LM (+covariance). */
int bf_hip_last_timing(double *times, int count);

/* Device memory helpers so hosts without another HIP binding can stage
 * inputs in HBM: hipMalloc, hipFree and a synchronous copy (kind 1 H2D, 2 D2H,
 * 3 D2D) ordered after the work already enqueued on the library's stream. */
int bf_hip_malloc(void **ptr, size_t bytes, char *err, size_t errlen);
int bf_hip_free(void *ptr);
int bf_hip_memcpy(void *dst, const void *src, size_t bytes, int kind, char *err, size_t errlen);
int bf_hip_synchronize(char *err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
