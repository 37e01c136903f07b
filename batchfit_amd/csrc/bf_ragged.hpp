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

namespace bfhip {

struct RaggedArgs {
    const bf_fit_desc *fits;
    const bf_grid_desc *grids;
    const GridInfo *grid_info;  // one per grid
    const bf_config *configs;
    const uint32_t *order;      // this class's fits, longest first
    uint32_t count;             // fits in this class
    uint32_t *cursor;           // this class's work-queue counter (reset before the launch)
    const float *x, *y, *p0;    // p0 rows of 8, or the guess buffer
    const float *guess;         // rows of 8 written by the guess pass
    float *params;              // rows of 8
    int32_t *status;
    uint32_t *iterations;
    float *chi2;
    float *cov;                 // rows of 64 or null
    float *stderr_;             // rows of 8
    int32_t *ustatus;
};

// Per-fit FitArgs over pre-offset pointers, so solve_fit works on fit 0.
__device__ __forceinline__ FitArgs ragged_fit_args(const RaggedArgs &r, const bf_fit_desc &d,
                                                   uint32_t model, int p) {
    const bf_config &c = r.configs[d.config];
    FitArgs a;
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

// Broadcast lane 0 of each G-lane group to the group.
template <int G> __device__ __forceinline__ uint32_t group_first(uint32_t v, int gl) {
    if constexpr (G == 64)
        return __builtin_amdgcn_readfirstlane(v);
    else
        return uint32_t(group_max_i<G>(gl == 0 ? int(v) : -1));
}

template <int MODEL, int G, int R, bool COMP, bool MIXED>
__global__ void __launch_bounds__(256, lm_min_waves(MODEL, G, R)) lm_ragged_kernel(RaggedArgs r) {
    using S = typename std::conditional<MIXED, double, float>::type;
    constexpr int P = Params<MODEL>::P;
    const int gl = int(threadIdx.x % G);
    for (;;) {
        uint32_t task = 0;
        if (gl == 0)
            task = atomicAdd(r.cursor, 1u);
        task = group_first<G>(task, gl);
        if (task >= r.count)
            break;
        const bf_fit_desc d = r.fits[r.order[task]];
        const FitArgs a = ragged_fit_args(r, d, MODEL, P);
        solve_fit<MODEL, G, R, COMP, S>(a, 0, gl, a.x);
    }
}

// Data-only guess of one class of a ragged batch, one wave per fit; LDS sized
// by the host for the class's largest M.
template <int MODEL> __global__ void __launch_bounds__(64) guess_ragged_kernel(
    const bf_fit_desc *fits, const uint32_t *order, uint32_t count, const bf_grid_desc *grids,
    const GridInfo *grid_info, const float *x, const float *y, float *guess, uint32_t m_pad,
    uint32_t sort_n) {
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
    g.sort_n = sort_n;
    g.params = guess + size_t(d.out_row) * 8;
    g.baseline = nullptr;
    g.gstatus = nullptr;
    guess_fit<MODEL>(g, 0);
}

// Mapping classes of a ragged batch by M (the tuned table of uniform
// batches: 16-lane groups for short spectra, 64 lanes with a 16- or 32-point
// register tile, and the streamed form beyond 2048 points).
struct RaggedClass {
    int G, R;
    uint32_t max_points;
};
constexpr RaggedClass kRaggedClasses[4] = {{16, 16, 256}, {64, 16, 1024}, {64, 32, 2048},
                                           {64, 0, 0xffffffffu}};
inline int ragged_class_of(uint32_t m) {
    for (int c = 0; c < 4; ++c)
        if (m <= kRaggedClasses[c].max_points)
            return c;
    return 3;
}

template <int MODEL>
hipError_t launch_ragged_lm(int cls, bool mixed, const RaggedArgs &args, uint32_t blocks,
                            hipStream_t stream);
template <int MODEL>
hipError_t launch_ragged_guess(const bf_fit_desc *fits, const uint32_t *order, uint32_t count,
                               const bf_grid_desc *grids, const GridInfo *grid_info,
                               const float *x, const float *y, float *guess, uint32_t m_pad,
                               uint32_t sort_n, hipStream_t stream);

}  // namespace bfhip
