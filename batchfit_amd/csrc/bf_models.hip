// bf_models.hip — per-model dispatch over the (model, group) launchers, the
// guess kernels and the model-evaluation test kernels.
#include "bf_launch.hpp"
#include "bf_guess_wide.hpp"

namespace bfhip {

#define BF_DECLARE_GROUPS(M)                                                                   \
    template <> hipError_t launch_lm_group<M, 1>(const Plan &, const FitArgs &, hipStream_t);  \
    template <> hipError_t launch_lm_group<M, 8>(const Plan &, const FitArgs &, hipStream_t);  \
    template <> hipError_t launch_lm_group<M, 4>(const Plan &, const FitArgs &, hipStream_t);  \
    template <> hipError_t launch_lm_group<M, 16>(const Plan &, const FitArgs &, hipStream_t); \
    template <> hipError_t launch_lm_group<M, 32>(const Plan &, const FitArgs &, hipStream_t); \
    template <> hipError_t launch_lm_group<M, 64>(const Plan &, const FitArgs &, hipStream_t); \
    template <> hipError_t launch_lm_group<M, 128>(const Plan &, const FitArgs &, hipStream_t); \
    template <> hipError_t launch_lm_group<M, 256>(const Plan &, const FitArgs &, hipStream_t);
BF_DECLARE_GROUPS(0)
BF_DECLARE_GROUPS(1)
BF_DECLARE_GROUPS(2)
BF_DECLARE_GROUPS(3)
#undef BF_DECLARE_GROUPS

#define BF_MODEL_LAUNCHERS(M)                                                                  \
    template <> hipError_t launch_lm_model<M>(const Plan &plan, const FitArgs &args,           \
                                              hipStream_t stream) {                            \
        switch (plan.G) {                                                                      \
        case 1:                                                                                \
            return launch_lm_group<M, 1>(plan, args, stream);                                  \
        case 4:                                                                                \
            return launch_lm_group<M, 4>(plan, args, stream);                                  \
        case 8:                                                                                \
            return launch_lm_group<M, 8>(plan, args, stream);                                  \
        case 16:                                                                               \
            return launch_lm_group<M, 16>(plan, args, stream);                                 \
        case 32:                                                                               \
            return launch_lm_group<M, 32>(plan, args, stream);                                 \
        case 64:                                                                               \
            return launch_lm_group<M, 64>(plan, args, stream);                                 \
        case 128:                                                                              \
            return launch_lm_group<M, 128>(plan, args, stream);                                \
        case 256:                                                                              \
            return launch_lm_group<M, 256>(plan, args, stream);                                \
        }                                                                                      \
        return hipErrorInvalidValue;                                                           \
    }                                                                                          \
    template <> hipError_t launch_guess_model<M>(const GuessArgs &args, size_t lds_bytes,      \
                                                 uint32_t form, hipStream_t stream) {          \
        hipError_t e = hipSuccess;                                                             \
        if (form == GUESS_AUTO && launch_guess_group<M>(args, stream, e))                      \
            return e;                                                                          \
        if (form == GUESS_WIDE ||                                                              \
            (form == GUESS_AUTO && (lds_bytes > kGuessLdsBytes ||                              \
                                    (args.m > kGuessWaveMaxPoints && args.ys_pt == 1)))) {     \
            if (args.ys_pt != 1)                                                               \
                return hipErrorInvalidValue;                                                   \
            hipLaunchKernelGGL((guess_wide_kernel<M>), dim3((args.n + kWideWaves - 1) /        \
                                                            kWideWaves),                       \
                               dim3(64 * kWideWaves), 0, stream, args);                        \
            return hipGetLastError();                                                          \
        }                                                                                      \
        if (lds_bytes > kGuessLdsBytes)                                                        \
            return hipErrorInvalidValue;                                                       \
        hipLaunchKernelGGL((guess_kernel<M>), dim3(args.n), dim3(64), lds_bytes, stream, args); \
        return hipGetLastError();                                                              \
    }                                                                                          \
    template <> hipError_t launch_eval_model<M>(const EvalArgs &args, hipStream_t stream) {    \
        hipLaunchKernelGGL((model_eval_kernel<M>), dim3((args.n + 255) / 256), dim3(256), 0,   \
                           stream, args);                                                      \
        return hipGetLastError();                                                              \
    }
BF_MODEL_LAUNCHERS(0)
BF_MODEL_LAUNCHERS(1)
BF_MODEL_LAUNCHERS(2)
BF_MODEL_LAUNCHERS(3)
#undef BF_MODEL_LAUNCHERS

}  // namespace bfhip
