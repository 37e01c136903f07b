// bf_ragged.hpp — fits of mixed model, M and options in one captured launch.
//
// A ragged batch is described per fit (bf_fit_desc: observation offset, grid,
// M, model, config row, p0 row, output row). The host plan buckets fits into
// classes (model × mapping) and orders each class longest-first; per class,
// a persistent kernel keeps its lane groups busy by pulling the next fit from
// a device counter (the work queue) until the class is exhausted. Every fit
// is solved by the same solve_fit used for uniform batches, with the mapping
// chosen from (model, M) only, so a fit's result does not depend on the rest
// of the batch.
#pragma once

#include "bf_launch.hpp"

#include "../../include/batchfit_hip.h"

#include <string>
#include <vector>

namespace bfhip {

struct RaggedArgs {
    const bf_fit_desc *fits;
    const bf_grid_desc *grids;
    const GridInfo *grid_info;  // one per grid
    const bf_config *configs;
    const uint32_t *order;      // this class's fits, longest first
    uint32_t count;             // fits in this class
    uint32_t *cursor;           // this class's work-queue counter (reset before the launch)
    uint32_t flat_check;        // some fit of the class starts from a given p0
    const float *x, *y, *p0;    // p0 rows of 8, or the guess buffer
    const float *guess;         // rows of 8 written by the guess pass
    float *params;              // rows of 8
    int32_t *status;
    uint32_t *iterations;
    float *chi2;
    float *cov;                 // rows of 64 or null
    float *stderr_;             // rows of 8
    int32_t *ustatus;
    const uint32_t *narrow;     // 1 when every grid of the batch lies within kFixedBaselineMaxX
    uint32_t fixed_baseline;    // two-peak class whose every fit fixes both baseline parameters
    uint32_t wide_only;         // the full form of such a class: runs only when !*narrow
};

// The value of lane 0 in every lane, held in scalar registers (for data a
// whole wave shares: with 64-lane groups a descriptor and its config are
// per-wave, and scalar copies keep them out of the vector register budget).
template <typename T> __device__ __forceinline__ T wave_uniform(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i)
        w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// Per-fit FitArgs over pre-offset pointers, so solve_fit works on fit 0.
template <int G>
__device__ __forceinline__ FitArgs ragged_fit_args(const RaggedArgs &r, const bf_fit_desc &d,
                                                   uint32_t model, int p) {
    bf_config c = r.configs[d.config];
    if constexpr (G == 64)
        c = wave_uniform(c);
    FitArgs a;
    a.pool = nullptr;
    a.flat_check = 1;
    a.x = r.x + r.grids[d.grid].x_offset;
    a.y = r.y + d.y_offset;
    a.ys_fit = 0;
    a.ys_pt = 1;
    a.p0 = d.p0_row == 0xffffffffu ? r.guess + size_t(d.out_row) * 8 : r.p0 + size_t(d.p0_row) * 8;
    a.params = r.params + size_t(d.out_row) * 8;
    a.status = r.status + d.out_row;
    a.iterations = r.iterations + d.out_row;
    a.chi2 = r.chi2 + d.out_row;
    a.cov = r.cov ? r.cov + size_t(d.out_row) * 64 : nullptr;
    a.stderr_ = r.stderr_ ? r.stderr_ + size_t(d.out_row) * 8 : nullptr;
    a.ustatus = r.ustatus ? r.ustatus + d.out_row : nullptr;
    a.grid = r.grid_info + d.grid;
    a.n = 1;
    a.m = d.points;
    a.max_iterations = c.max_iterations;
    a.initial_lambda = c.initial_lambda;
    a.gradient_tolerance = c.gradient_tolerance;
    a.step_tolerance = c.step_tolerance;
    a.cost_tolerance = c.cost_tolerance;
    bool ok = finite_f(c.initial_lambda) && c.initial_lambda > 0.0f;
    for (float t : {c.gradient_tolerance, c.step_tolerance, c.cost_tolerance})
        ok = ok && finite_f(t) && t >= 0.0f;
    for (int k = 0; k < 8; ++k) {
        a.lower[k] = c.lower[k];
        a.upper[k] = c.upper[k];
        if (k < p)
            ok = ok && !nan_f(c.lower[k]) && !nan_f(c.upper[k]) && !(c.lower[k] > c.upper[k]) &&
                 __float_as_uint(c.lower[k]) != 0x7f800000u &&
                 __float_as_uint(c.upper[k]) != 0xff800000u;
    }
    a.numerics = c.numerics;
    a.options_ok = ok;
    bool free = c.lower[1] != c.upper[1] || c.lower[2] != c.upper[2];
    if (model == PSEUDO_VOIGT)
        free = free || c.lower[3] != c.upper[3];
    if (model == TWO_LORENTZIAN)
        free = free || c.lower[4] != c.upper[4] || c.lower[5] != c.upper[5];
    a.free_shape = free;
    a.width_factor = model == GAUSSIAN ? 2.354820045030949382f : model == PSEUDO_VOIGT ? 1.0f : 2.0f;
    return a;
}

// Persistent LM kernel of one class. The work queue is taken per wave: lane 0
// of the wave adds 1 to the class counter and the wave's 64/G groups solve
// tasks base·(64/G) … base·(64/G)+64/G-1. The queue read is branch-free (every
// lane issues the add, only lane 0 adds 1) and the loop exit is a scalar
// branch on a wave-uniform value, so no lane of a group can reach a group
// reduction on a path that skipped the queue read: with a per-group exit, the
// optimiser may thread the lanes with gl != 0 straight back to the group
// broadcast, which then reads lane 0 while it is inactive.
// Orders the LDS accesses of one wave around a hand-over between its lanes
// (the LDS unit runs one wave's instructions in order; this keeps the
// compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Occupancy floor of the persistent kernels: the work-queue loop and the
// per-fit descriptor cost a few registers more than the uniform kernel, so
// where that kernel asks for three waves per SIMD this one asks for two
// (three would spill).
constexpr int lm_ragged_min_waves(int model, int g, int r) {
    return lm_min_waves(model, g, r) > 2 ? 2 : lm_min_waves(model, g, r);
}

template <int MODEL, int G, int R, bool COMP, bool MIXED>
__global__ void __launch_bounds__(256, lm_ragged_min_waves(MODEL, G, R))
    lm_ragged_kernel(RaggedArgs r) {
    using S = typename std::conditional<MIXED, double, float>::type;
    constexpr int P = Params<MODEL>::P;
    constexpr uint32_t per_wave = 64 / G;
    const int lane = int(threadIdx.x % 64);
    const int gl = lane % G;
    const uint32_t g = uint32_t(lane / G);
    // With a register tile, each group copies its fit's grid into its own LDS
    // slice of G·R floats (zero beyond M), as the uniform kernel does for the
    // shared grid: the points then read x from LDS, not from memory.
    __shared__ float xs[R > 0 ? 256 * R : 1];
    float *const my_x = xs + (R > 0 ? (threadIdx.x / G) * G * R : 0);
    for (;;) {
        const uint32_t ticket = atomicAdd(r.cursor, lane == 0 ? 1u : 0u);
        const uint32_t base = __builtin_amdgcn_readfirstlane(ticket) * per_wave;
        if (base >= r.count)
            break;
        const uint32_t task = base + g;
        if (task < r.count) {
            bf_fit_desc d = r.fits[r.order[task]];
            if constexpr (G == 64)
                d = wave_uniform(d);
            const FitArgs a = ragged_fit_args<G>(r, d, MODEL, P);
            if constexpr (R > 0) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int j = gl + G * k;
                    my_x[j] = j < int(a.m) ? a.x[j] : 0.0f;
                }
                wave_lds_sync();
                solve_fit<MODEL, G, R, COMP, S>(a, 0, gl, my_x);
                wave_lds_sync();  // every lane has read the slice before it is refilled
            } else {
                solve_fit<MODEL, G, R, COMP, S>(a, 0, gl, a.x);
            }
        }
    }
}

// The fits of one streamed class of a ragged batch as a source for
// solve_pool: tasks are the class's slice of the work order; each group holds
// its current fit's arguments (scalar for whole-wave groups).
template <int MODEL, int G> struct RaggedPool : PoolQueue<G> {
    const RaggedArgs &r;
    FitArgs fa;
    GridInfo info;
    __device__ RaggedPool(const RaggedArgs &args, int lane_, int gl_)
        : PoolQueue<G>(args.cursor, args.count, lane_, gl_), r(args) {
        fa.pool = nullptr;
    }
    template <typename Pts> __device__ void begin(uint32_t task, Pts &pts) {
        bf_fit_desc d = r.fits[r.order[task]];
        if constexpr (G == 64)
            d = wave_uniform(d);
        fa = ragged_fit_args<G>(r, d, MODEL, Params<MODEL>::P);
        info = *fa.grid;
        pts.load(fa.x, fa.y, 1, this->gl, int(fa.m));
    }
    __device__ const FitArgs &args() const { return fa; }
    __device__ const GridInfo &grid() const { return info; }
    __device__ size_t row() const { return 0; }
};

// A streamed class without uncertainties: persistent waves whose groups
// retire fits one by one (solve_pool) from the class's counter. A two-peak
// class with a fixed baseline launches the FIXB form and then the full form
// with wide_only set, on the same counter: exactly one of them runs (the FIXB
// form when every grid of the batch lies within kFixedBaselineMaxX).
template <int MODEL, int G, bool MIXED, bool FLAT, bool FIXB = false>
__global__ void __launch_bounds__(256, lm_ragged_min_waves(MODEL, G, 0))
    lm_ragged_pool_kernel(RaggedArgs r) {
    using S = typename std::conditional<MIXED, double, float>::type;
    if (FIXB ? !*r.narrow : (r.wide_only && *r.narrow))
        return;  // the whole launch: the other form serves this class
    const int lane = int(threadIdx.x & 63);
    RaggedPool<MODEL, G> src(r, lane, lane % G);
    solve_pool<MODEL, G, MIXED, FLAT, S, FIXB>(lane % G, src);
}

// Uncertainties of a streamed class after its retiring launch: one group per
// fit, from the fit's written parameters, status and χ². FIXB and wide_only
// as for lm_ragged_pool_kernel: exactly one of the two forms runs.
template <int MODEL, int G, bool MIXED, bool FIXB = false>
__global__ void __launch_bounds__(256) lm_ragged_cov_kernel(RaggedArgs r) {
    using S = typename std::conditional<MIXED, double, float>::type;
    constexpr int P = Params<MODEL>::P;
    const uint32_t task = blockIdx.x * (256 / G) + threadIdx.x / G;
    const int gl = int(threadIdx.x % G);
    if (task >= r.count || (FIXB ? !*r.narrow : (r.wide_only && *r.narrow)))
        return;
    bf_fit_desc d = r.fits[r.order[task]];
    if constexpr (G == 64)
        d = wave_uniform(d);
    const FitArgs a = ragged_fit_args<G>(r, d, MODEL, P);
    float p[P];
#pragma unroll
    for (int k = 0; k < P; ++k)
        p[k] = a.params[k];
    Points<G, 0, stream_chunk(MODEL)> pts;
    pts.load(a.x, a.y, 1, gl, int(a.m));
    __shared__ S cl[256 / G][P * P];
    covariance<MODEL, G, 0, S, FIXB, true>(a, pts, 0, gl, p, a.status[0], a.chi2[0],
                                           cl[threadIdx.x / G]);
}

// Data-only guess of one class of a ragged batch, one wave per fit; LDS sized
// by the host for the class's largest M.
template <int MODEL> __global__ void __launch_bounds__(64) guess_ragged_kernel(
    const bf_fit_desc *fits, const uint32_t *order, uint32_t count, const bf_grid_desc *grids,
    const GridInfo *grid_info, const float *x, const float *y, float *guess, uint32_t m_pad,
    uint32_t extra_n) {
    const uint32_t task = blockIdx.x;
    if (task >= count)
        return;
    const bf_fit_desc d = fits[order[task]];
    if (d.p0_row != 0xffffffffu)
        return;
    GuessArgs g;
    g.x = x + grids[d.grid].x_offset;
    g.y = y + d.y_offset;
    g.ys_fit = 0;
    g.ys_pt = 1;
    g.grid = grid_info + d.grid;
    g.n = 1;
    g.m = d.points;
    g.m_pad = m_pad;
    g.extra_n = extra_n;
    g.params = guess + size_t(d.out_row) * 8;
    g.baseline = nullptr;
    g.gstatus = nullptr;
    guess_fit<MODEL>(g, 0);
}

// Mapping classes of a ragged batch: lanes per fit G and whether the points
// stream (R = 0) or sit in a register tile of R = 16 or 32 per lane, chosen
// from the measured table as for uniform batches. The arithmetic of a fit
// depends on (G, M) only, so the class's R does not change its result.
struct RaggedClass {
    int G, R;
    uint32_t max_points;
};
constexpr int kRaggedClassCount = 13;
constexpr RaggedClass kRaggedClasses[kRaggedClassCount] = {
    {16, 16, 256},  {16, 32, 512},        {32, 16, 512},        {32, 32, 1024},
    {64, 16, 1024}, {64, 32, 2048},       {64, 0, 0xffffffffu}, {16, 0, 0xffffffffu},
    {32, 0, 0xffffffffu}, {8, 16, 128},   {8, 32, 256},         {8, 0, 0xffffffffu},
    {4, 0, 0xffffffffu}};
// Enough queue counters for every run: one per (model, class) bucket, plus
// at most 8 more from splitting buckets heavier than an eighth of the batch.
// Queue counters: one per (model, class) bucket and one per fixed-baseline
// two-peak class, plus at most 8 more from splitting heavy buckets; the word
// after them is the batch's narrow-grid flag.
constexpr int kRaggedCounters = 80;
static_assert(5 * kRaggedClassCount + 8 <= kRaggedCounters, "one counter per run");
constexpr int kRaggedCounterWords = kRaggedCounters + 1;
inline int ragged_class_of(uint32_t model, uint32_t m) {
    const TunedRange t = tuned_mapping(int(model), m);
    // Beyond 32 points per lane the points stream, as in a uniform launch.
    const bool streamed = t.streamed || m > 32u * uint32_t(t.lanes);
    for (int c = 0; c < kRaggedClassCount; ++c) {
        const RaggedClass &k = kRaggedClasses[c];
        if (k.G == t.lanes && (streamed ? k.R == 0 : (k.R > 0 && m <= k.max_points)))
            return c;
    }
    return 6;  // 64 lanes, streamed (not reached for a valid table)
}

// One class of a ragged batch: a slice of the work order.
// Fits of a run whose M lies in one octave (2^k, 2^(k+1)] (all M <= 256
// together), for the guess: each launch sizes its LDS for its own longest
// fit, so short fits are not held to the occupancy of the longest.
struct GuessSlice {
    uint32_t offset, count;  // within the run's slice of the order array
    uint32_t max_points;
};

struct RaggedRun {
    uint32_t model;
    int cls;
    bool fixed_baseline;  // two-peak fits that all fix both baseline parameters
    uint32_t offset, count;  // slice of the order array
    uint32_t guesses;        // fits in the slice that start from the device guess
    uint32_t max_points;
    double cost;             // relative work, for ordering and branch balance
    std::vector<GuessSlice> guess_slices;  // octaves that hold a guessed fit
};

// Whether a fit of `model` under config c runs in the fixed-baseline form:
// two-peak, both baseline parameters fixed by equal bounds.
inline bool fixed_baseline(uint32_t model, const bf_config &c) {
    return model == TWO_LORENTZIAN && c.lower[6] == c.upper[6] && c.lower[7] == c.upper[7];
}

// Relative work per point of a fit by model (points × parameter-weighted
// passes: a P=8 fit costs about three times a P=3 one per point), for run
// ordering and branch balance.
constexpr double kRaggedWeight[4] = {1.0, 1.0, 1.5, 3.5};

// Lower end of the guess slice that holds a run's fits of at most `top`
// points, longest first: the octave (2^k, 2^(k+1)] of M (all M <= 256
// together), and single-peak fits of at most kGuessSortedMaxPoints apart
// from longer ones (they take the sorting guess).
inline uint32_t guess_slice_floor(uint32_t model, uint32_t top) {
    uint32_t below = 0;  // the octave is (below, 2·below]
    if (top > 256)
        for (below = 256; 2 * below < top; below *= 2) {
        }
    if (model != TWO_LORENTZIAN && top > kGuessSortedMaxPoints)
        below = below > kGuessSortedMaxPoints ? below : kGuessSortedMaxPoints;
    return below;
}

// Host view of a ragged batch: validated descriptors bucketed by (model,
// class), each bucket longest-first and, when asked, cut into runs of at
// most about an eighth of the batch's work.
struct RaggedLayout {
    std::vector<uint32_t> order;
    std::vector<RaggedRun> runs;
    std::vector<size_t> cost_order;  // runs, heaviest first
    bool any_guess = false, any_p0 = false;
    uint64_t out_rows = 0;  // 1 + the largest output row (rows are distinct, < 2^32 - 1)
};

// Device tables of a ragged batch.
struct RaggedDevice {
    const bf_fit_desc *fits;
    const bf_grid_desc *grids;
    const bf_config *configs;
    GridInfo *info;
    const uint32_t *order;
    uint32_t *cursors;  // kRaggedCounterWords words: the runs' counters, then the narrow-grid flag
    float *guess;       // out_rows × 8 when any fit starts from the guess
    uint32_t grid_count;
};

// Validate and bucket; returns an empty string or the reason for rejection.
// `split`: cut buckets heavier than an eighth of the batch into several runs
// (for the branches of a large batch's graph; the server's batches stay
// whole, which kept their tail latency lower).
std::string ragged_layout(uint32_t fits, const bf_fit_desc *fit_table, uint32_t grids,
                          const bf_grid_desc *grid_table, uint32_t configs,
                          const bf_config *config_table, RaggedLayout &out, bool split);

// Device inputs and outputs of one ragged batch (rows of 8 for params and
// stderr, 64 for cov; cov/stderr/ustatus null without uncertainties).
struct RaggedIO {
    const float *x, *y, *p0;
    float *params;
    int32_t *status;
    uint32_t *iterations;
    float *chi2, *cov, *stderr_;
    int32_t *ustatus;
};

// Side streams and events for running the runs of a batch as parallel
// branches (stream capture or eager).
struct RaggedFork {
    int streams;
    hipStream_t *side;
    hipEvent_t fork;
    hipEvent_t *join;
};

// Enqueue a whole ragged batch on `stream` (no synchronisation, no capture):
// queue reset and grid table on `stream`, then each run's guess and LM —
// on `stream` in order of decreasing cost, or spread over fork->side
// branches that join back into `stream`.
hipError_t ragged_enqueue(const RaggedLayout &layout, const RaggedDevice &dev, bool mixed,
                          const RaggedIO &io, hipStream_t stream, const RaggedFork *fork);

// Load the kernels a layout launches (outside any capture).
hipError_t ragged_prepare(const RaggedLayout &layout, bool mixed);

template <int MODEL> hipError_t prepare_ragged(int cls, bool mixed, bool guess);
template <int MODEL>
hipError_t launch_ragged_lm(int cls, bool mixed, const RaggedArgs &args, uint32_t blocks,
                            hipStream_t stream);
template <int MODEL>
hipError_t launch_ragged_guess(const bf_fit_desc *fits, const uint32_t *order, uint32_t count,
                               const bf_grid_desc *grids, const GridInfo *grid_info,
                               const float *x, const float *y, float *guess, uint32_t m_pad,
                               uint32_t extra_n, hipStream_t stream);

}  // namespace bfhip
