// bf_runtime.hip — the C ABI of libbatchfit_hip.so (include/batchfit_hip.h):
// argument checks, stream-ordered scratch, the grid prologue, the data-only
// guess, the optional point-major transpose and the LM launch.
//
// Replaces the reference's Metal host runtime (src/metal/runtime.cpp:26-316)
// and the binding's host preparation (python/bindings.cpp:60-121): the guess
// runs on the device instead of a 2-worker CPU pool, and y stays fit-major in
// HBM for the group mappings instead of a host 16×16 transpose.
#include "bf_host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace bfhip {

template <int MODEL> hipError_t launch_lm_model(const Plan &, const FitArgs &, hipStream_t);
template <int MODEL> hipError_t launch_guess_model(const GuessArgs &, size_t, uint32_t, hipStream_t);
template <int MODEL> hipError_t launch_eval_model(const EvalArgs &, hipStream_t);

namespace {

// ---------------------------------------------------------------------------
// Device kernels owned by the runtime.

// Grid facts of x, and the zeroed work counter of the LM launch that follows
// (pool may be null).
__global__ void __launch_bounds__(256) grid_kernel(const float *x, uint32_t m, GridInfo *out,
                                                   uint32_t *pool) {
    grid_block(x, m, out);
    if (pool && threadIdx.x == 0)
        *pool = 0;
}

// Fit-major [n,m] → point-major [m,n] (the layout of the reference GPU
// kernels, batch.hpp:119), 64×64 tiles through LDS.
__global__ void __launch_bounds__(256) transpose_kernel(const float *in, float *out, uint32_t n,
                                                        uint32_t m) {
    __shared__ float tile[64][65];
    const uint32_t fit0 = blockIdx.x * 64, pt0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int k = ty; k < 64; k += 4) {
        const uint32_t i = fit0 + k, j = pt0 + tx;
        if (i < n && j < m)
            tile[k][tx] = in[size_t(i) * m + j];
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const uint32_t j = pt0 + k, i = fit0 + tx;
        if (i < n && j < m)
            out[size_t(j) * n + i] = tile[tx][k];
    }
}

// ---------------------------------------------------------------------------
// Host side.

std::atomic<bool> g_timing{std::getenv("BF_HIP_TIMING") != nullptr};

struct Timing {
    hipEvent_t ev[5] = {};
    bool recorded = false;
    bool phase[4] = {};
};
thread_local Timing t_timing;

hipError_t record(int i, hipStream_t stream) {
    if (!t_timing.ev[i]) {
        const hipError_t e = hipEventCreate(&t_timing.ev[i]);
        if (e != hipSuccess)
            return e;
    }
    return hipEventRecord(t_timing.ev[i], stream);
}

// Lanes per fit of the latency mapping (BF_MAP_LATENCY) for a batch of n fits
// of M points, or 0 to keep the throughput table's choice. The table is
// measured on batches that fill the device (262,144 fits per shape); a small
// batch in small groups leaves most of the device idle, so it takes the
// smallest group of 64, 128 or 256 lanes whose register tile holds M (at
// most 32 points per lane; 256 lanes streamed beyond 8192 points) while the
// batch stays within 2^19 lanes (8 waves per SIMD). Measured on pseudo-Voigt
// and two-Lorentzian fits, M = 1024 … 8192, N = 32 … 2048: 1.5-10x shorter
// LM time than the throughput mapping.
int latency_lanes(int model, uint32_t m, uint32_t n) {
    const TunedRange tuned = tuned_mapping(model, m);
    if (uint64_t(n) * uint64_t(tuned.lanes) >= (uint64_t(1) << 18))
        return 0;  // the throughput mapping fills the device
    for (int g : {64, 128, 256})
        if (m <= 32u * uint32_t(g) || g == 256)
            return uint64_t(n) * uint64_t(g) <= (uint64_t(1) << 19) ? g : 0;
    return 0;
}

// Tuned default mapping by (model, M) (replaces metal::dispatch_variant,
// include/batchfit/metal/batch.hpp:21-50); with BF_MAP_LATENCY also by the
// batch size n.
int choose_plan(uint32_t model, uint32_t m, uint32_t n, uint32_t numerics, const bf_launch *launch,
                Plan &plan, char *err, size_t errlen) {
    plan.comp = (numerics & BF_COMPENSATED) != 0;
    plan.mixed = (numerics & BF_MIXED_FP64) != 0;
    const uint32_t mapping = launch ? launch->mapping : BF_MAP_AUTO;
    if (mapping == BF_MAP_LANE) {
        plan.G = 1;
        plan.R = 0;
        return BF_OK;
    }
    if (mapping == BF_MAP_LATENCY) {
        if (const int g = latency_lanes(int(model), m, n)) {
            bf_launch wide{};
            wide.mapping = BF_MAP_GROUP;
            wide.group_lanes = uint32_t(g);
            wide.streamed = 0;
            if (plan_instantiated(g, g > 64 ? workgroup_tile(m, g) : 2, plan.comp, plan.mixed) ||
                plan_instantiated(g, 0, plan.comp, plan.mixed))
                return choose_plan(model, m, n, numerics, &wide, plan, err, errlen);
        }
        return choose_plan(model, m, n, numerics, nullptr, plan, err, errlen);
    }
    const TunedRange tuned = tuned_mapping(int(model), m);
    int g = tuned.lanes;
    bool streamed = mapping == BF_MAP_AUTO && tuned.streamed;
    if (mapping == BF_MAP_GROUP && launch->group_lanes)
        g = int(launch->group_lanes);
    if (mapping == BF_MAP_GROUP)
        streamed = launch->streamed != 0;
    else if (mapping != BF_MAP_GROUP && mapping != BF_MAP_AUTO)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "unknown mapping");
    if (mapping == BF_MAP_AUTO && g != 64 &&
        !plan_instantiated(g, g == 4 ? 0 : 2, plan.comp, plan.mixed))
        g = 64;  // 4/8/16/32-lane groups exist for numerics 0 and 5 only
    if (g != 4 && g != 8 && g != 16 && g != 32 && g != 64 && g != 128 && g != 256)
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "group_lanes must be 4, 8, 16, 32, 64 (lanes of a wave) or 128, 256 "
                    "(a workgroup per fit)");
    const uint32_t per_lane = (m + uint32_t(g) - 1) / uint32_t(g);
    int r = pow2_ceil(per_lane < 2 ? 2 : per_lane);
    if (g > 64)
        r = workgroup_tile(m, g);  // tiles of 8 or 32 points per lane
    if (r > kMaxTile || streamed || g == 4)  // 4-lane groups always stream
        r = 0;
    plan.G = g;
    plan.R = r;
    if (!plan_instantiated(plan.G, plan.R, plan.comp, plan.mixed))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "this group size supports numerics 0-3 (FP32) and 5/7 (compensated mixed)");
    return BF_OK;
}

hipError_t launch_lm(uint32_t model, const Plan &plan, const FitArgs &args, hipStream_t s) {
    switch (model) {
    case 0: return launch_lm_model<0>(plan, args, s);
    case 1: return launch_lm_model<1>(plan, args, s);
    case 2: return launch_lm_model<2>(plan, args, s);
    case 3: return launch_lm_model<3>(plan, args, s);
    }
    return hipErrorInvalidValue;
}
hipError_t launch_guess(uint32_t model, const GuessArgs &args, size_t lds, hipStream_t s,
                        uint32_t form = GUESS_AUTO) {
    switch (model) {
    case 0: return launch_guess_model<0>(args, lds, form, s);
    case 1: return launch_guess_model<1>(args, lds, form, s);
    case 2: return launch_guess_model<2>(args, lds, form, s);
    case 3: return launch_guess_model<3>(args, lds, form, s);
    }
    return hipErrorInvalidValue;
}

bool options_valid(const bf_config &c, int p) {
    if (!std::isfinite(c.initial_lambda) || c.initial_lambda <= 0.0f)
        return false;
    for (float tol : {c.gradient_tolerance, c.step_tolerance, c.cost_tolerance})
        if (!std::isfinite(tol) || tol < 0.0f)
            return false;
    for (int k = 0; k < p; ++k)
        if (std::isnan(c.lower[k]) || std::isnan(c.upper[k]) || c.lower[k] > c.upper[k] ||
            c.lower[k] == INFINITY || c.upper[k] == -INFINITY)
            return false;
    return true;
}

bool shape_free(uint32_t model, const bf_config &c) {
    bool free = c.lower[1] != c.upper[1] || c.lower[2] != c.upper[2];
    if (model == BF_PSEUDO_VOIGT)
        free = free || c.lower[3] != c.upper[3];
    if (model == BF_TWO_LORENTZIAN)
        free = free || c.lower[4] != c.upper[4] || c.lower[5] != c.upper[5];
    return free;
}

float width_factor_of(uint32_t model) {
    return model == BF_GAUSSIAN ? 2.354820045030949382f : model == BF_PSEUDO_VOIGT ? 1.0f : 2.0f;
}

// Stream-ordered scratch released at the end of the enqueuing call.
struct Scratch {
    hipStream_t stream;
    void *ptrs[4] = {};
    int used = 0;
    explicit Scratch(hipStream_t s) : stream(s) {}
    hipError_t get(void **out, size_t bytes) {
        const hipError_t e = hipMallocAsync(out, bytes ? bytes : 1, stream);
        if (e == hipSuccess)
            ptrs[used++] = *out;
        return e;
    }
    ~Scratch() {
        for (int i = 0; i < used; ++i)
            (void)hipFreeAsync(ptrs[i], stream);
    }
};

int check_common(uint32_t model, const bf_config *cfg, char *err, size_t errlen) {
    if (model > 3)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "unknown fit model");
    if (!cfg)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "config is required");
    if (cfg->points == 0)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "expected x[M] and y[N,M] with M>0");
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP fitting is unavailable: " + rt.why);
    return BF_OK;
}

// Scratch of the no-allocation entry (bf_hip_fit_device_noalloc): grid facts
// and the retiring kernel's work counter in the first 512 bytes, then the
// device guess rows (N·P floats). Every GPU the library runs on holds its own
// copy until the process ends; calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocGuessBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[512 + kNoAllocGuessBytes];

// The device address of `scratch` on the current device, looked up once per
// device (the array's name is a host-side stand-in, not a device address).
int noalloc_scratch(char **out, char *err, size_t errlen) {
    static std::mutex mu;
    static char *address[64] = {};
    int device = 0;
    BF_TRY(hipGetDevice(&device), "hipGetDevice");
    if (device < 0 || device >= 64)
        return fail(err, errlen, BF_RUNTIME_ERROR, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(mu);
    if (!address[device]) {
        void *ptr = nullptr;
        BF_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        address[device] = static_cast<char *>(ptr);
    }
    *out = address[device];
    return BF_OK;
}

// fixed: null → scratch is taken stream-ordered from the device pool;
// otherwise the no-allocation scratch above (no transpose, no timing events).
int fit_device(uint32_t model, const bf_config *cfg, const bf_launch *launch, const float *x,
               const float *y, const float *p0, float *params, int32_t *status,
               uint32_t *iterations, float *chi2, float *cov, float *se, int32_t *us, char *err,
               size_t errlen, char *fixed = nullptr) {
    if (const int rc = check_common(model, cfg, err, errlen))
        return rc;
    const uint32_t n = cfg->fits, m = cfg->points;
    const int P = kParams[model];
    const bool unc = cov || se || us;
    if (unc && !(cov && se && us))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "covariance, stderr and uncertainty status are requested together");
    if (n == 0)
        return BF_OK;
    if (!x || !y || !params || !status || !iterations || !chi2)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "x, y and every output must be non-null");
    Plan plan;
    if (const int rc = choose_plan(model, m, n, cfg->numerics, launch, plan, err, errlen))
        return rc;
    hipStream_t stream =
        launch && launch->stream ? static_cast<hipStream_t>(launch->stream) : Runtime::get().stream;
    if (fixed && !p0 && size_t(n) * P * sizeof(float) > kNoAllocGuessBytes)
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "the no-allocation entry holds device guesses for at most " +
                        std::to_string(kNoAllocGuessBytes / (P * sizeof(float))) +
                        " fits of this model; pass p0 or use bf_hip_fit_device");
    if (fixed && plan.G == 1 && launch && launch->point_major)
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "the no-allocation entry does not transpose y to point-major");
    const bool timing = !fixed && g_timing.load(std::memory_order_relaxed);
    Timing &tm = t_timing;
    tm.recorded = false;
    for (bool &b : tm.phase)
        b = false;
    Scratch temp(stream);
    GridInfo *grid = nullptr;
    // GridInfo, then the work counter of the retiring kernel at a 256-byte
    // offset; grid_kernel writes both (the counter is cleared by the kernel,
    // not by a memset).
    if (fixed)
        grid = reinterpret_cast<GridInfo *>(fixed);
    else
        BF_TRY(temp.get(reinterpret_cast<void **>(&grid), 512), "hipMallocAsync(grid)");
    static_assert(sizeof(GridInfo) <= 256, "grid facts fit before the counter");
    uint32_t *pool = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(grid) + 256);
    if (timing)
        BF_TRY(record(0, stream), "hipEventRecord");
    hipLaunchKernelGGL(grid_kernel, dim3(1), dim3(256), 0, stream, x, m, grid, pool);
    BF_TRY(hipGetLastError(), "grid prologue launch");
    if (timing)
        BF_TRY(record(1, stream), "hipEventRecord");
    tm.phase[0] = true;

    const float *init = p0;
    if (!init) {
        const uint32_t extra_n = guess_extra_floats(int(model), m);
        // The single-peak models' guess streams the row (the one-wave kernel
        // for edge windows of 129–230 values); the two-peak guess holds the
        // signal in LDS up to kGuessWaveMaxPoints and reads it from memory
        // beyond (no limit on M either way).
        const size_t lds = size_t(m + extra_n) * sizeof(float);
        float *guess = nullptr;
        if (fixed)
            guess = reinterpret_cast<float *>(fixed + 512);  // every row written by the guess
        else
            BF_TRY(temp.get(reinterpret_cast<void **>(&guess), size_t(n) * P * sizeof(float)),
                   "hipMallocAsync(guess)");
        GuessArgs ga{x, y, (long long)m, 1, grid, n, m, m, extra_n, guess, nullptr, nullptr};
        BF_TRY(launch_guess(model, ga, lds, stream), "initial guess launch");
        init = guess;
        tm.phase[1] = true;
    }
    if (timing)
        BF_TRY(record(2, stream), "hipEventRecord");

    const float *ydata = y;
    long long ys_fit = m, ys_pt = 1;
    if (plan.G == 1 && launch && launch->point_major) {
        float *yt = nullptr;
        BF_TRY(temp.get(reinterpret_cast<void **>(&yt), size_t(n) * m * sizeof(float)),
               "hipMallocAsync(point-major y)");
        hipLaunchKernelGGL(transpose_kernel, dim3((n + 63) / 64, (m + 63) / 64), dim3(256), 0,
                           stream, y, yt, n, m);
        BF_TRY(hipGetLastError(), "transpose launch");
        ydata = yt;
        ys_fit = 1;
        ys_pt = n;
        tm.phase[2] = true;
    }
    if (timing)
        BF_TRY(record(3, stream), "hipEventRecord");

    FitArgs a{};
    a.x = x;
    a.y = ydata;
    a.ys_fit = ys_fit;
    a.ys_pt = ys_pt;
    a.p0 = init;
    a.params = params;
    a.status = status;
    a.iterations = iterations;
    a.chi2 = chi2;
    a.cov = unc ? cov : nullptr;
    a.stderr_ = se;
    a.ustatus = us;
    a.grid = grid;
    a.pool = pool;
    a.flat_check = p0 != nullptr;
    a.n = n;
    a.m = m;
    a.max_iterations = cfg->max_iterations;
    a.initial_lambda = cfg->initial_lambda;
    a.gradient_tolerance = cfg->gradient_tolerance;
    a.step_tolerance = cfg->step_tolerance;
    a.cost_tolerance = cfg->cost_tolerance;
    for (int k = 0; k < 8; ++k) {
        a.lower[k] = cfg->lower[k];
        a.upper[k] = cfg->upper[k];
    }
    a.numerics = cfg->numerics;
    a.options_ok = options_valid(*cfg, P);
    a.free_shape = shape_free(model, *cfg);
    a.width_factor = width_factor_of(model);
    BF_TRY(launch_lm(model, plan, a, stream), "LM kernel launch");
    tm.phase[3] = true;
    if (timing) {
        BF_TRY(record(4, stream), "hipEventRecord");
        tm.recorded = true;
    }
    return BF_OK;
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

namespace {

// Host threads that copy one buffer into another together, each its own
// piece: pageable rows into a pinned staging slot, at several times the rate
// of one thread, so the DMA from the slot is what sets the pace.
class CopyCrew {
  public:
    // Throws std::system_error if a thread cannot be started (after joining
    // the ones that were).
    explicit CopyCrew(unsigned threads) : threads_(threads) {
        try {
            for (unsigned k = 0; k < threads_; ++k)
                pool_.emplace_back([this, k] { run(k); });
        } catch (...) {
            halt();
            throw;
        }
    }
    ~CopyCrew() { halt(); }
    // dst[0..bytes) = src[0..bytes); returns when every piece is copied.
    void copy(void *dst, const void *src, size_t bytes) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            bytes_ = bytes;
            left_ = threads_;
            ++generation_;
        }
        work_.notify_all();
        std::unique_lock<std::mutex> lock(mu_);
        done_.wait(lock, [&] { return left_ == 0; });
    }

  private:
    void halt() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        work_.notify_all();
        for (auto &t : pool_)
            t.join();
        pool_.clear();
    }
    void run(unsigned k) {
        uint64_t seen = 0;
        for (;;) {
            char *dst;
            const char *src;
            size_t bytes;
            {
                std::unique_lock<std::mutex> lock(mu_);
                work_.wait(lock, [&] { return stop_ || generation_ != seen; });
                if (stop_)
                    return;
                seen = generation_;
                dst = dst_;
                src = src_;
                bytes = bytes_;
            }
            // Pieces of whole 4 KiB pages.
            const size_t step = ((bytes + threads_ - 1) / threads_ + 4095) & ~size_t(4095);
            const size_t a = std::min(bytes, size_t(k) * step), b = std::min(bytes, a + step);
            if (a < b)
                std::memcpy(dst + a, src + a, b - a);
            std::lock_guard<std::mutex> lock(mu_);
            if (--left_ == 0)
                done_.notify_one();
        }
    }
    const unsigned threads_;
    std::vector<std::thread> pool_;
    std::mutex mu_;
    std::condition_variable work_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
    unsigned left_ = 0;
    uint64_t generation_ = 0;
    bool stop_ = false;
};

}  // namespace

extern "C" {

int bf_hip_abi_version(void) { return BF_HIP_ABI_VERSION; }

int bf_hip_available(void) { return Runtime::get().status == BF_OK ? 1 : 0; }

void bf_hip_default_config(bf_config *cfg) {
    if (!cfg)
        return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->max_iterations = 100;
    cfg->initial_lambda = 1e-3f;
    cfg->gradient_tolerance = 1e-5f;
    cfg->step_tolerance = 1e-6f;
    cfg->cost_tolerance = 1e-6f;
    for (int k = 0; k < 8; ++k) {
        cfg->lower[k] = -INFINITY;
        cfg->upper[k] = INFINITY;
    }
}

int bf_hip_plan(uint32_t model, uint32_t points, uint32_t numerics, const bf_launch *launch,
                uint32_t *group_lanes, uint32_t *tile_points, char *err, size_t errlen) {
    return bf_hip_plan_batch(model, points, 0, numerics, launch, group_lanes, tile_points, err,
                             errlen);
}

int bf_hip_plan_batch(uint32_t model, uint32_t points, uint32_t fits, uint32_t numerics,
                      const bf_launch *launch, uint32_t *group_lanes, uint32_t *tile_points,
                      char *err, size_t errlen) {
    if (model > 3)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "unknown model");
    Plan plan{};
    if (const int rc = choose_plan(model, points, fits, numerics, launch, plan, err, errlen))
        return rc;
    if (group_lanes)
        *group_lanes = uint32_t(plan.G);
    if (tile_points)
        *tile_points = uint32_t(plan.R);
    return BF_OK;
}

int bf_hip_set_timing(int enable) {
    g_timing.store(enable != 0);
    return BF_OK;
}

int bf_hip_last_timing(double *times, int count) {
    Timing &tm = t_timing;
    for (int i = 0; i < count; ++i)
        times[i] = 0.0;
    if (!tm.recorded)
        return BF_INVALID_ARGUMENT;
    if (hipEventSynchronize(tm.ev[4]) != hipSuccess)
        return BF_RUNTIME_ERROR;
    for (int i = 0; i < count && i < 4; ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, tm.ev[i], tm.ev[i + 1]) != hipSuccess)
            return BF_RUNTIME_ERROR;
        times[i] = ms;
    }
    return BF_OK;
}

int bf_hip_fit_device(uint32_t model, const bf_config *cfg, const bf_launch *launch,
                      const float *d_x, const float *d_y, const float *d_p0, float *d_params,
                      int32_t *d_status, uint32_t *d_iterations, float *d_chi2, float *d_cov,
                      float *d_stderr, int32_t *d_ustatus, char *err, size_t errlen) {
    return fit_device(model, cfg, launch, d_x, d_y, d_p0, d_params, d_status, d_iterations,
                      d_chi2, d_cov, d_stderr, d_ustatus, err, errlen);
}

int bf_hip_fit_device_noalloc(uint32_t model, const bf_config *cfg, const bf_launch *launch,
                              const float *d_x, const float *d_y, const float *d_p0,
                              float *d_params, int32_t *d_status, uint32_t *d_iterations,
                              float *d_chi2, float *d_cov, float *d_stderr, int32_t *d_ustatus,
                              char *err, size_t errlen) {
    if (const int rc = check_common(model, cfg, err, errlen))
        return rc;
    char *fixed = nullptr;
    if (const int rc = noalloc_scratch(&fixed, err, errlen))
        return rc;
    return fit_device(model, cfg, launch, d_x, d_y, d_p0, d_params, d_status, d_iterations,
                      d_chi2, d_cov, d_stderr, d_ustatus, err, errlen, fixed);
}

int bf_hip_fit(uint32_t model, const bf_config *cfg, const bf_launch *launch, const float *x,
               const float *y, const float *p0, float *params, int32_t *status,
               uint32_t *iterations, float *chi2, float *cov, float *stderr_, int32_t *ustatus,
               char *err, size_t errlen) {
    if (const int rc = check_common(model, cfg, err, errlen))
        return rc;
    const bool unc = cov || stderr_ || ustatus;
    if (unc && !(cov && stderr_ && ustatus))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "covariance, stderr and uncertainty status are requested together");
    const size_t n = cfg->fits, m = cfg->points, P = size_t(kParams[model]);
    if (n == 0)
        return BF_OK;
    if (!x || !y || !params || !status || !iterations || !chi2)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "x, y and every output must be non-null");
    Runtime &rt = Runtime::get();
    std::lock_guard<std::mutex> lock(rt.host_calls);
    // The batch goes through the device in chunks of about 64 MiB of y on two
    // streams: while the device fits chunk k, the host copies chunk k + 1 in
    // on the other stream. A fit's result depends on its own row, model, M
    // and mapping only, so chunking changes no result.
    for (hipStream_t &side : rt.side)
        if (!side)
            BF_TRY(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate");
    const size_t chunk = std::min(n, std::max<size_t>(256, ((size_t(64) << 20) / (m * 4)) & ~size_t(255)));
    const size_t chunks = (n + chunk - 1) / chunk, slots = chunks > 1 ? 2 : 1;
    hipStream_t s = rt.stream;
    Scratch scratch(s);
    float *dx;
    BF_TRY(scratch.get(reinterpret_cast<void **>(&dx), m * 4), "hipMallocAsync(x)");
    BF_TRY(hipMemcpyAsync(dx, x, m * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync(x)");
    // Per slot: y | p0 | params | chi2 | status | iterations | cov | stderr | ustatus.
    const size_t out_per_fit = (P * 4 + 4 + 4 + 4) + (unc ? (P * P * 4 + P * 4 + 4) : 0);
    const size_t slot_bytes = chunk * (m * 4 + (p0 ? P * 4 : 0) + out_per_fit);
    char *dslot[2] = {nullptr, nullptr};
    for (size_t k = 0; k < slots; ++k)
        BF_TRY(scratch.get(reinterpret_cast<void **>(&dslot[k]), slot_bytes),
               "hipMallocAsync(chunk buffers)");
    BF_TRY(hipStreamSynchronize(s), "chunk buffer allocation");  // usable on the side streams
    // Several chunks: y goes through two pinned staging slots, filled by a
    // crew of host threads while the DMA engine moves the other slot, so the
    // copy-in runs at the link's pinned rate instead of the runtime's
    // single-threaded staging of pageable memory (chunks of at most 256 MiB;
    // larger rows keep the pageable copies).
    // A chunk's results come back in one copy into a pinned slot and are
    // unpacked on the host (a copy to pageable memory costs as much as a
    // chunk's fit, and there are up to seven per chunk).
    const size_t y_chunk_bytes = chunk * m * 4, out_chunk_bytes = chunk * out_per_fit;
    std::unique_ptr<CopyCrew> crew;
    if (chunks > 1 && y_chunk_bytes <= (size_t(256) << 20)) {
        try {
            const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
            crew = std::make_unique<CopyCrew>(std::min(8u, hw));
        } catch (const std::exception &) {
            crew.reset();  // no threads to be had: the runtime's pageable copies
        }
    }
    // Pinned slots of at least `need` bytes (kept for later calls); false if
    // they cannot be had, and then the runtime's pageable copies serve.
    auto ensure = [&](char *(&slots_)[2], size_t &have, size_t need) -> bool {
        if (have >= need)
            return true;
        have = 0;  // before any slot is freed: a failure below leaves no slot counted
        for (char *&slot : slots_)
            if (slot) {
                if (hipHostFree(slot) != hipSuccess)
                    return false;
                slot = nullptr;
            }
        for (char *&slot : slots_)
            if (hipHostMalloc(reinterpret_cast<void **>(&slot), need) != hipSuccess) {
                slot = nullptr;
                (void)hipGetLastError();
                return false;
            }
        have = need;
        return true;
    };
    const bool staged = crew && ensure(rt.staging, rt.staging_bytes, y_chunk_bytes) &&
                        ensure(rt.staging_out, rt.staging_out_bytes, out_chunk_bytes);
    struct Uploads {
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Uploads() {
            for (hipEvent_t e : ev)
                if (e)
                    (void)hipEventDestroy(e);
        }
    } uploads;
    if (staged)
        for (hipEvent_t &e : uploads.ev)
            BF_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    // Whatever way this call returns, the side streams finish with the chunk
    // buffers before `scratch` frees them on the library stream.
    struct SideDone {
        Runtime &rt;
        size_t slots;
        ~SideDone() {
            for (size_t k = 0; k < slots; ++k)
                (void)hipStreamSynchronize(rt.side[k]);
        }
    } side_done{rt, slots};
    struct Slot {
        float *y, *p0, *params, *chi2, *cov, *se;
        int32_t *status, *us;
        uint32_t *it;
    };
    auto carve = [&](char *base) {
        Slot v{};
        v.y = reinterpret_cast<float *>(base);
        v.p0 = v.y + chunk * m;
        v.params = v.p0 + (p0 ? chunk * P : 0);
        v.chi2 = v.params + chunk * P;
        v.status = reinterpret_cast<int32_t *>(v.chi2 + chunk);
        v.it = reinterpret_cast<uint32_t *>(v.status + chunk);
        v.cov = reinterpret_cast<float *>(v.it + chunk);
        v.se = v.cov + (unc ? chunk * P * P : 0);
        v.us = reinterpret_cast<int32_t *>(v.se + (unc ? chunk * P : 0));
        return v;
    };
    // Copies chunk c's results out (host arrays); ordered after its fit on its stream.
    auto drain = [&](size_t c) -> int {
        const Slot v = carve(dslot[c % slots]);
        hipStream_t st = rt.side[c % slots];
        const size_t f0 = c * chunk, nc = std::min(chunk, n - f0);
        if (staged) {
            char *h = rt.staging_out[c % slots];
            BF_TRY(hipMemcpyAsync(h, v.params, out_chunk_bytes, hipMemcpyDeviceToHost, st),
                   "copy results");
            BF_TRY(hipStreamSynchronize(st), "results of a chunk");
            // The slot's arrays as carved on the device: params | chi2 |
            // status | iterations | cov | stderr | ustatus, `chunk` rows each.
            const char *hp = h, *hc = hp + chunk * P * 4, *hs = hc + chunk * 4,
                       *hi = hs + chunk * 4, *hv = hi + chunk * 4;
            std::memcpy(params + f0 * P, hp, nc * P * 4);
            std::memcpy(chi2 + f0, hc, nc * 4);
            std::memcpy(status + f0, hs, nc * 4);
            std::memcpy(iterations + f0, hi, nc * 4);
            if (unc) {
                const char *he = hv + chunk * P * P * 4, *hu = he + chunk * P * 4;
                std::memcpy(cov + f0 * P * P, hv, nc * P * P * 4);
                std::memcpy(stderr_ + f0 * P, he, nc * P * 4);
                std::memcpy(ustatus + f0, hu, nc * 4);
            }
            return BF_OK;
        }
        BF_TRY(hipMemcpyAsync(params + f0 * P, v.params, nc * P * 4, hipMemcpyDeviceToHost, st),
               "copy params");
        BF_TRY(hipMemcpyAsync(chi2 + f0, v.chi2, nc * 4, hipMemcpyDeviceToHost, st), "copy chi2");
        BF_TRY(hipMemcpyAsync(status + f0, v.status, nc * 4, hipMemcpyDeviceToHost, st),
               "copy status");
        BF_TRY(hipMemcpyAsync(iterations + f0, v.it, nc * 4, hipMemcpyDeviceToHost, st),
               "copy iterations");
        if (unc) {
            BF_TRY(hipMemcpyAsync(cov + f0 * P * P, v.cov, nc * P * P * 4, hipMemcpyDeviceToHost,
                                  st),
                   "copy cov");
            BF_TRY(hipMemcpyAsync(stderr_ + f0 * P, v.se, nc * P * 4, hipMemcpyDeviceToHost, st),
                   "copy stderr");
            BF_TRY(hipMemcpyAsync(ustatus + f0, v.us, nc * 4, hipMemcpyDeviceToHost, st),
                   "copy ustatus");
        }
        return BF_OK;
    };
    for (size_t c = 0; c < chunks; ++c) {
        if (c >= slots)  // the slot's previous chunk leaves before this one enters
            if (const int rc = drain(c - slots))
                return rc;
        const Slot v = carve(dslot[c % slots]);
        hipStream_t st = rt.side[c % slots];
        const size_t f0 = c * chunk, nc = std::min(chunk, n - f0);
        if (staged) {
            // The slot's previous upload (chunk c − 2) has left it.
            if (c >= slots)
                BF_TRY(hipEventSynchronize(uploads.ev[c % slots]), "staging slot reuse");
            char *pinned = rt.staging[c % slots];
            crew->copy(pinned, y + f0 * m, nc * m * 4);
            BF_TRY(hipMemcpyAsync(v.y, pinned, nc * m * 4, hipMemcpyHostToDevice, st),
                   "hipMemcpyAsync(y)");
            BF_TRY(hipEventRecord(uploads.ev[c % slots], st), "hipEventRecord(upload)");
        } else {
            BF_TRY(hipMemcpyAsync(v.y, y + f0 * m, nc * m * 4, hipMemcpyHostToDevice, st),
                   "hipMemcpyAsync(y)");
        }
        if (p0)
            BF_TRY(hipMemcpyAsync(v.p0, p0 + f0 * P, nc * P * 4, hipMemcpyHostToDevice, st),
                   "hipMemcpyAsync(p0)");
        bf_config part = *cfg;
        part.fits = uint32_t(nc);
        bf_launch own{};
        if (launch)
            own = *launch;
        own.stream = st;
        const int rc = fit_device(model, &part, &own, dx, v.y, p0 ? v.p0 : nullptr, v.params,
                                  v.status, v.it, v.chi2, unc ? v.cov : nullptr,
                                  unc ? v.se : nullptr, unc ? v.us : nullptr, err, errlen);
        if (rc != BF_OK)
            return rc;
    }
    for (size_t c = chunks > slots ? chunks - slots : 0; c < chunks; ++c)
        if (const int rc = drain(c))
            return rc;
    for (size_t k = 0; k < slots; ++k)
        BF_TRY(hipStreamSynchronize(rt.side[k]), "fit completion");
    return BF_OK;
}

int bf_hip_guess(uint32_t model, uint32_t n, uint32_t m, const float *x, const float *y,
                 float *params, float *baseline, int32_t *gstatus, char *err, size_t errlen) {
    return bf_hip_guess_form(model, BF_GUESS_AUTO, n, m, x, y, params, baseline, gstatus, err,
                             errlen);
}

int bf_hip_guess_form(uint32_t model, uint32_t form, uint32_t n, uint32_t m, const float *x,
                      const float *y, float *params, float *baseline, int32_t *gstatus, char *err,
                      size_t errlen) {
    bf_config cfg;
    bf_hip_default_config(&cfg);
    cfg.points = m;
    cfg.fits = n;
    if (const int rc = check_common(model, &cfg, err, errlen))
        return rc;
    if (form > BF_GUESS_WIDE)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "unknown guess form");
    if (n == 0)
        return BF_OK;
    if (!x || !y || !params || !baseline || !gstatus)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "every array must be non-null");
    const uint32_t extra_n = guess_extra_floats(int(model), m);
    const size_t lds = size_t(m + extra_n) * sizeof(float);
    if (form == BF_GUESS_WAVE && lds > kGuessLdsBytes) {
        char msg[96];
        std::snprintf(msg, sizeof msg, "the one-wave guess of this model holds M up to %u points in LDS",
                      guess_lds_max_points(int(model)));
        return fail(err, errlen, BF_INVALID_ARGUMENT, msg);
    }
    Runtime &rt = Runtime::get();
    std::lock_guard<std::mutex> lock(rt.host_calls);
    hipStream_t s = rt.stream;
    const size_t P = size_t(kParams[model]);
    Scratch scratch(s);
    float *dx, *dy, *dp, *db;
    int32_t *dg;
    GridInfo *grid;
    BF_TRY(scratch.get(reinterpret_cast<void **>(&dx), size_t(m) * 4), "hipMallocAsync");
    BF_TRY(scratch.get(reinterpret_cast<void **>(&dy), size_t(n) * m * 4), "hipMallocAsync");
    BF_TRY(scratch.get(reinterpret_cast<void **>(&dp), size_t(n) * (P + 2) * 4), "hipMallocAsync");
    BF_TRY(scratch.get(reinterpret_cast<void **>(&grid), sizeof(GridInfo)), "hipMallocAsync");
    db = dp + size_t(n) * P;
    dg = reinterpret_cast<int32_t *>(db + n);
    BF_TRY(hipMemcpyAsync(dx, x, size_t(m) * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    BF_TRY(hipMemcpyAsync(dy, y, size_t(n) * m * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hipLaunchKernelGGL(grid_kernel, dim3(1), dim3(256), 0, s, dx, m, grid, nullptr);
    BF_TRY(hipGetLastError(), "grid prologue launch");
    GuessArgs ga{dx, dy, (long long)m, 1, grid, n, m, m, extra_n, dp, db, dg};
    BF_TRY(launch_guess(model, ga, lds, s, form), "initial guess launch");
    BF_TRY(hipMemcpyAsync(params, dp, size_t(n) * P * 4, hipMemcpyDeviceToHost, s), "copy");
    BF_TRY(hipMemcpyAsync(baseline, db, size_t(n) * 4, hipMemcpyDeviceToHost, s), "copy");
    BF_TRY(hipMemcpyAsync(gstatus, dg, size_t(n) * 4, hipMemcpyDeviceToHost, s), "copy");
    BF_TRY(hipStreamSynchronize(s), "guess completion");
    return BF_OK;
}

int bf_hip_model_eval(uint32_t model, uint32_t n, const float *x, const float *params, float *value,
                      float *jac, char *err, size_t errlen) {
    bf_config cfg;
    bf_hip_default_config(&cfg);
    cfg.points = 1;
    if (const int rc = check_common(model, &cfg, err, errlen))
        return rc;
    if (n == 0)
        return BF_OK;
    Runtime &rt = Runtime::get();
    std::lock_guard<std::mutex> lock(rt.host_calls);
    hipStream_t s = rt.stream;
    const size_t P = size_t(kParams[model]);
    Scratch scratch(s);
    float *buf;
    BF_TRY(scratch.get(reinterpret_cast<void **>(&buf), size_t(n) * (1 + 8 + 1 + P) * 4),
           "hipMallocAsync");
    float *dx = buf, *dp = dx + n, *dv = dp + size_t(n) * 8, *dj = dv + n;
    BF_TRY(hipMemcpyAsync(dx, x, size_t(n) * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    BF_TRY(hipMemcpyAsync(dp, params, size_t(n) * 32, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    EvalArgs ea{dx, dp, dv, dj, n};
    hipError_t e = hipErrorInvalidValue;
    switch (model) {
    case 0: e = launch_eval_model<0>(ea, s); break;
    case 1: e = launch_eval_model<1>(ea, s); break;
    case 2: e = launch_eval_model<2>(ea, s); break;
    case 3: e = launch_eval_model<3>(ea, s); break;
    }
    BF_TRY(e, "model evaluation launch");
    BF_TRY(hipMemcpyAsync(value, dv, size_t(n) * 4, hipMemcpyDeviceToHost, s), "copy");
    BF_TRY(hipMemcpyAsync(jac, dj, size_t(n) * P * 4, hipMemcpyDeviceToHost, s), "copy");
    BF_TRY(hipStreamSynchronize(s), "model evaluation completion");
    return BF_OK;
}

int bf_hip_malloc(void **ptr, size_t bytes, char *err, size_t errlen) {
    if (Runtime::get().status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP is unavailable: " + Runtime::get().why);
    BF_TRY(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
    return BF_OK;
}

int bf_hip_free(void *ptr) { return hipFree(ptr) == hipSuccess ? BF_OK : BF_RUNTIME_ERROR; }

int bf_hip_memcpy(void *dst, const void *src, size_t bytes, int kind, char *err, size_t errlen) {
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP is unavailable: " + rt.why);
    const hipMemcpyKind k = kind == 1   ? hipMemcpyHostToDevice
                            : kind == 2 ? hipMemcpyDeviceToHost
                                        : hipMemcpyDeviceToDevice;
    // Ordered after the work already on the library stream (which is
    // non-blocking, so a null-stream copy would not wait for it).
    BF_TRY(hipMemcpyAsync(dst, src, bytes, k, rt.stream), "hipMemcpyAsync");
    BF_TRY(hipStreamSynchronize(rt.stream), "hipStreamSynchronize");
    return BF_OK;
}

int bf_hip_synchronize(char *err, size_t errlen) {
    BF_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return BF_OK;
}

}  // extern "C"
