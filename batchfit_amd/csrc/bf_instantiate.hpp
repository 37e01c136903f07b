// bf_instantiate.hpp — included once per model translation unit with
// BF_MODEL set to the model id; defines that model's launchers.
#pragma once

#include "bf_launch.hpp"

#ifndef BF_MODEL
#error "define BF_MODEL before including bf_instantiate.hpp"
#endif

namespace bfhip {

namespace {
template <int MODEL, int G, int R, bool COMP, bool MIXED>
hipError_t go(const FitArgs &args, hipStream_t stream) {
    if constexpr (!plan_instantiated(G, R, COMP, MIXED)) {
        return hipErrorInvalidValue;
    } else {
        constexpr int per_block = 256 / G;
        const uint32_t blocks = (args.n + per_block - 1) / per_block;
        hipLaunchKernelGGL((lm_kernel<MODEL, G, R, COMP, MIXED>), dim3(blocks), dim3(256), 0,
                           stream, args);
        return hipGetLastError();
    }
}

template <int MODEL, int G, bool COMP, bool MIXED>
hipError_t by_tile(int r, const FitArgs &args, hipStream_t stream) {
    switch (r) {
    case 0:
        return go<MODEL, G, 0, COMP, MIXED>(args, stream);
    case 1:
        return go<MODEL, G, 1, COMP, MIXED>(args, stream);
    case 2:
        return go<MODEL, G, 2, COMP, MIXED>(args, stream);
    case 4:
        return go<MODEL, G, 4, COMP, MIXED>(args, stream);
    case 8:
        return go<MODEL, G, 8, COMP, MIXED>(args, stream);
    case 16:
        return go<MODEL, G, 16, COMP, MIXED>(args, stream);
    case 32:
        return go<MODEL, G, 32, COMP, MIXED>(args, stream);
    }
    return hipErrorInvalidValue;
}

template <int MODEL, int G>
hipError_t by_numerics(const Plan &plan, const FitArgs &args, hipStream_t stream) {
    if (plan.mixed)
        return plan.comp ? by_tile<MODEL, G, true, true>(plan.R, args, stream)
                         : by_tile<MODEL, G, false, true>(plan.R, args, stream);
    return plan.comp ? by_tile<MODEL, G, true, false>(plan.R, args, stream)
                     : by_tile<MODEL, G, false, false>(plan.R, args, stream);
}
}  // namespace

template <> hipError_t launch_lm_model<BF_MODEL>(const Plan &plan, const FitArgs &args,
                                                 hipStream_t stream) {
    constexpr int MODEL = BF_MODEL;
    switch (plan.G) {
    case 1: {
        constexpr int per_block = 256;
        const uint32_t blocks = (args.n + per_block - 1) / per_block;
        if (plan.mixed) {
            if (plan.comp)
                hipLaunchKernelGGL((lm_kernel<MODEL, 1, 0, true, true>), dim3(blocks), dim3(256), 0,
                                   stream, args);
            else
                hipLaunchKernelGGL((lm_kernel<MODEL, 1, 0, false, true>), dim3(blocks), dim3(256),
                                   0, stream, args);
        } else {
            if (plan.comp)
                hipLaunchKernelGGL((lm_kernel<MODEL, 1, 0, true, false>), dim3(blocks), dim3(256),
                                   0, stream, args);
            else
                hipLaunchKernelGGL((lm_kernel<MODEL, 1, 0, false, false>), dim3(blocks), dim3(256),
                                   0, stream, args);
        }
        return hipGetLastError();
    }
    case 16:
        return by_numerics<MODEL, 16>(plan, args, stream);
    case 32:
        return by_numerics<MODEL, 32>(plan, args, stream);
    case 64:
        return by_numerics<MODEL, 64>(plan, args, stream);
    }
    return hipErrorInvalidValue;
}

template <> hipError_t launch_guess_model<BF_MODEL>(const GuessArgs &args, size_t lds_bytes,
                                                    hipStream_t stream) {
    hipLaunchKernelGGL((guess_kernel<BF_MODEL>), dim3(args.n), dim3(64), lds_bytes, stream, args);
    return hipGetLastError();
}

template <> hipError_t launch_eval_model<BF_MODEL>(const EvalArgs &args, hipStream_t stream) {
    hipLaunchKernelGGL((model_eval_kernel<BF_MODEL>), dim3((args.n + 255) / 256), dim3(256), 0,
                       stream, args);
    return hipGetLastError();
}

}  // namespace bfhip
