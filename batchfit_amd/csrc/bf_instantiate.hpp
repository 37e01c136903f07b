// bf_instantiate.hpp — included by each bf_lm_<model>_g<G>.hip with BF_MODEL
// and BF_GROUP set; defines launch_lm_group<BF_MODEL, BF_GROUP>.
#pragma once

#include "bf_launch.hpp"

#include <algorithm>

#if !defined(BF_MODEL) || !defined(BF_GROUP)
#error "define BF_MODEL and BF_GROUP before including bf_instantiate.hpp"
#endif

namespace bfhip {

namespace {
template <int MODEL, int G, int R, bool COMP, bool MIXED>
hipError_t go(const Plan &, const FitArgs &args, hipStream_t stream) {
    if constexpr (!plan_instantiated(G, R, COMP, MIXED)) {
        return hipErrorInvalidValue;
    } else if constexpr (G > 64) {
        // One fit per workgroup at a time; at most 2^20 workgroups, each
        // taking every gridDim-th fit after its first.
        const uint32_t groups = std::min<uint32_t>(args.n, 1u << 20);
        hipLaunchKernelGGL((lm_block_kernel<MODEL, G, R, COMP, MIXED>), dim3(groups), dim3(G), 0,
                           stream, args);
        return hipGetLastError();
    } else {
        constexpr int per_block = 256 / G;
        const uint32_t blocks = (args.n + per_block - 1) / per_block;
        if constexpr (R == 0 && G > 1) {
            if (args.pool) {
                // Persistent waves that retire fits one by one (solve_pool).
                // One refill per wave to start with, at most 4096 waves (more
                // than the device holds at once); uncertainties follow in a
                // second launch, one group per fit.
                const uint32_t waves = (args.n + pool_refill(G) - 1) / pool_refill(G);
                const uint32_t pool_blocks = std::min<uint32_t>((waves + 3) / 4, 1024u);
                FitArgs full = args;
                if constexpr (MODEL == TWO_LORENTZIAN) {
                    // Fixed baseline: the form that sums only the free
                    // parameters' rows, then the full form for wide grids.
                    if (args.lower[6] == args.upper[6] && args.lower[7] == args.upper[7]) {
                        if (args.flat_check)
                            hipLaunchKernelGGL((lm_pool_kernel<MODEL, G, COMP, MIXED, true, true>),
                                               dim3(pool_blocks), dim3(256), 0, stream, args);
                        else
                            hipLaunchKernelGGL((lm_pool_kernel<MODEL, G, COMP, MIXED, false, true>),
                                               dim3(pool_blocks), dim3(256), 0, stream, args);
                        if (hipError_t e = hipGetLastError(); e != hipSuccess)
                            return e;
                        full.wide_grid_only = 1;
                    }
                }
                if (args.flat_check)
                    hipLaunchKernelGGL((lm_pool_kernel<MODEL, G, COMP, MIXED, true>),
                                       dim3(pool_blocks), dim3(256), 0, stream, full);
                else
                    hipLaunchKernelGGL((lm_pool_kernel<MODEL, G, COMP, MIXED, false>),
                                       dim3(pool_blocks), dim3(256), 0, stream, full);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess || !args.cov)
                    return e;
                if constexpr (MODEL == TWO_LORENTZIAN) {
                    if (full.wide_grid_only) {
                        hipLaunchKernelGGL((lm_cov_kernel<MODEL, G, MIXED, true>), dim3(blocks),
                                           dim3(256), 0, stream, args);
                        if ((e = hipGetLastError()) != hipSuccess)
                            return e;
                    }
                }
                hipLaunchKernelGGL((lm_cov_kernel<MODEL, G, MIXED>), dim3(blocks), dim3(256), 0,
                                   stream, full);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((lm_kernel<MODEL, G, R, COMP, MIXED>), dim3(blocks), dim3(256), 0,
                           stream, args);
        return hipGetLastError();
    }
}

template <int MODEL, int G, bool COMP, bool MIXED>
hipError_t by_tile(const Plan &plan, const FitArgs &args, hipStream_t stream) {
    switch (plan.R) {
    case 0:
        return go<MODEL, G, 0, COMP, MIXED>(plan, args, stream);
    }
    if constexpr (G > 64) {
        switch (plan.R) {
        case 8:
            return go<MODEL, G, 8, COMP, MIXED>(plan, args, stream);
        case 32:
            return go<MODEL, G, 32, COMP, MIXED>(plan, args, stream);
        }
        return hipErrorInvalidValue;
    } else {
        switch (plan.R) {
        case 2:
            return go<MODEL, G, 2, COMP, MIXED>(plan, args, stream);
        case 4:
            return go<MODEL, G, 4, COMP, MIXED>(plan, args, stream);
        case 8:
            return go<MODEL, G, 8, COMP, MIXED>(plan, args, stream);
        case 16:
            return go<MODEL, G, 16, COMP, MIXED>(plan, args, stream);
        case 32:
            return go<MODEL, G, 32, COMP, MIXED>(plan, args, stream);
        }
        return hipErrorInvalidValue;
    }
}
}  // namespace

template <> hipError_t launch_lm_group<BF_MODEL, BF_GROUP>(const Plan &plan, const FitArgs &args,
                                                           hipStream_t stream) {
    constexpr int MODEL = BF_MODEL, G = BF_GROUP;
    if (plan.mixed)
        return plan.comp ? by_tile<MODEL, G, true, true>(plan, args, stream)
                         : by_tile<MODEL, G, false, true>(plan, args, stream);
    return plan.comp ? by_tile<MODEL, G, true, false>(plan, args, stream)
                     : by_tile<MODEL, G, false, false>(plan, args, stream);
}

}  // namespace bfhip
