// bf_launch.hpp — launch tables for the LM and guess kernels. Each model is
// instantiated in its own translation unit (bf_model_<name>.hip) so the
// library builds in parallel; bf_runtime.hip dispatches through launch_lm /
// launch_guess / launch_model_eval.
#pragma once

#include "bf_guess.hpp"

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
        a.jac[size_t(i) * P + k] = jac[k];
}

template <int MODEL> hipError_t launch_lm_model(const Plan &plan, const FitArgs &args,
                                                hipStream_t stream);
template <int MODEL> hipError_t launch_guess_model(const GuessArgs &args, size_t lds_bytes,
                                                   hipStream_t stream);
template <int MODEL> hipError_t launch_eval_model(const EvalArgs &args, hipStream_t stream);

// Instantiated plans: lane-per-fit with every numerics form; 64-lane groups
// with tiles 0 (stream), 1 … 32 and every numerics form; 16- and 32-lane
// groups with tiles 2 … 32 in the plain FP32 and compensated-mixed forms.
constexpr bool plan_instantiated(int G, int R, bool comp, bool mixed) {
    const bool tile = R == 1 || R == 2 || R == 4 || R == 8 || R == 16 || R == 32;
    if (G == 1)
        return R == 0;
    if (G == 64)
        return R == 0 || tile;
    if (G == 16 || G == 32)
        return R >= 2 && tile && comp == mixed;
    return false;
}

}  // namespace bfhip
