// bf_launch.hpp — launch tables for the LM and guess kernels. Each model is
// instantiated in its own translation unit (bf_model_<name>.hip) so the
// library builds in parallel; bf_runtime.hip dispatches through launch_lm /
// launch_guess / launch_model_eval.
#pragma once

#include "bf_lm.hpp"
#include "bf_tuned.hpp"

namespace bfhip {

// Fit-to-hardware plan: G lanes per fit, R points per lane in registers
// (0 = stream from memory), numerics (compensated partials, mixed FP64).
struct Plan {
    int G;
    int R;
    bool comp;
    bool mixed;
};

// Register tiles that are instantiated for each G (powers of two; a tile of R
// slots serves every M ≤ G·R, unused slots are skipped by uniform branches).
constexpr int kMaxTile = 32;

struct EvalArgs {
    const float *x;
    const float *params;  // [n,8]
    float *value;
    float *jac;  // [n,P]
    uint32_t n;
};

template <int MODEL> __global__ void __launch_bounds__(256) model_eval_kernel(EvalArgs a) {
    constexpr int P = Params<MODEL>::P;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n)
        return;
    ModelConsts<MODEL> mc;
    mc.set(a.params + size_t(i) * 8);
    float jac[P];
    a.value[i] = mc.eval(a.x[i], jac);
    for (int k = 0; k < P; ++k)
        a.jac[size_t(i) * P + k] = jac[k] * jac_scale<MODEL>(k);
}

// Shared-grid facts (cpu/solver.hpp:58-69, batch.hpp:236-240,
// initial_guess.hpp:131-153) for one 256-thread block. Minima are exact, so the
// parallel order gives the reference's sequential values.
__device__ inline void grid_block(const float *x, uint32_t m, GridInfo *out) {
    __shared__ float s_abs[256], s_gap[256], s_max[256];
    __shared__ int s_flags[256];
    const int t = threadIdx.x;
    const float x0 = x[0];
    float amin = __uint_as_float(0x7f800000u), gmin = 3.40282346638528859812e+38f;
    float xmax = 0.0f;  // largest finite |x| (a non-finite x sets flag 1)
    int fl = 0;  // 1 non-finite, 2 not rising, 4 not falling, 8 varied, 16 not strictly rising
    for (uint32_t j = t; j < m; j += 256) {
        const float xj = x[j];
        xmax = fmaxf(xmax, fabsf(xj));
        fl |= finite_f(xj) ? 0 : 1;
        fl |= (xj != x0) ? 8 : 0;
        if (j > 0) {
            const float prev = x[j - 1];
            const float d = xj - prev;
            fl |= (d >= 0.0f) ? 0 : 2;
            fl |= (d <= 0.0f) ? 0 : 4;
            if (d != 0.0f)
                amin = fminf(amin, fabsf(d));
            fl |= (xj <= prev) ? 16 : 0;
            gmin = fminf(gmin, d);
        }
    }
    s_abs[t] = amin;
    s_gap[t] = gmin;
    s_max[t] = xmax;
    s_flags[t] = fl;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            s_abs[t] = fminf(s_abs[t], s_abs[t + o]);
            s_gap[t] = fminf(s_gap[t], s_gap[t + o]);
            s_max[t] = fmaxf(s_max[t], s_max[t + o]);
            s_flags[t] |= s_flags[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        const int f = s_flags[0];
        const bool monotone = !(f & 2) || !(f & 4);
        GridInfo g;
        g.min_spacing = monotone && finite_f(s_abs[0]) ? s_abs[0] : 0.0f;
        g.varied_x = (f & 8) ? 1 : 0;
        g.x_finite = (f & 1) ? 0 : 1;
        g.guess_grid_ok = !(f & 1) && !(f & 16) && finite_f(x[m - 1] - x[0]);
        g.guess_spacing = s_gap[0];
        const uint32_t count = m / 10 > 1 ? m / 10 : 1;
        g.count = int(count);
        g.left_x = std_midpoint(x[(count - 1) / 2], x[count / 2]);
        g.right_x = std_midpoint(x[m - count + (count - 1) / 2], x[m - count + count / 2]);
        g.max_abs_x = (f & 1) ? __uint_as_float(0x7f800000u) : s_max[0];
        *out = g;
    }
}

template <int MODEL> hipError_t launch_lm_model(const Plan &plan, const FitArgs &args,
                                                hipStream_t stream);
// One (model, lanes-per-fit) class; each lives in its own translation unit
// (bf_lm_<model>_g<G>.hip) so the library compiles in parallel.
template <int MODEL, int G> hipError_t launch_lm_group(const Plan &plan, const FitArgs &args,
                                                       hipStream_t stream);
template <int MODEL> hipError_t launch_guess_model(const GuessArgs &args, size_t lds_bytes,
                                                   uint32_t form, hipStream_t stream);
template <int MODEL> hipError_t launch_eval_model(const EvalArgs &args, hipStream_t stream);

// Lanes per fit and whether the points stream, by (model, M), from the
// measured table (bf_tuned.hpp). The
// same table drives uniform and ragged batches, so a fit's arithmetic does
// not depend on how it was batched.
inline TunedRange tuned_mapping(int model, uint32_t m) {
    for (const TunedRange &r : kTuned[model & 3])
        if (m <= r.max_points)
            return r;
    return {0xffffffffu, 64, 0};
}
inline int tuned_group(int model, uint32_t m) { return tuned_mapping(model, m).lanes; }

// Instantiated plans: lane-per-fit with every numerics form; 64-lane groups
// with tiles 0 (stream), 2 … 32 and every numerics form; 8-, 16- and 32-lane
// groups streamed or with tiles 2 … 32, 4-lane groups streamed, and workgroups
// of 128 or 256 lanes streamed or with tiles of 8 or 32 points, in the plain
// FP32 and compensated-mixed forms.
constexpr bool plan_instantiated(int G, int R, bool comp, bool mixed) {
    const bool tile = R == 2 || R == 4 || R == 8 || R == 16 || R == 32;
    if (G == 1)
        return R == 0;
    if (G == 64)
        return R == 0 || tile;
    if (G == 4)
        return R == 0 && comp == mixed;
    if (G == 8 || G == 16 || G == 32)
        return (R == 0 || tile) && comp == mixed;
    if (G == 128 || G == 256)
        return (R == 0 || R == 8 || R == 32) && comp == mixed;
    return false;
}

// The register tile a workgroup mapping (G = 128, 256) uses for M points: the
// smallest instantiated tile (8 or 32 points per lane) that holds them, or 0
// (streamed) beyond.
constexpr int workgroup_tile(uint32_t m, int g) {
    const uint32_t per_lane = (m + uint32_t(g) - 1) / uint32_t(g);
    return per_lane <= 8 ? 8 : per_lane <= 32 ? 32 : 0;
}

}  // namespace bfhip
