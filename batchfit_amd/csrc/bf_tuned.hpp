// bf_tuned.hpp — generated by `python -m batchfit_amd.autotune --write`: lanes per
// fit by (model, M), measured on gfx950 (LM kernel time of the production numerics
// from true parameters; see batchfit_amd/autotune.py). Rows: gaussian, lorentzian,
// pseudo_voigt, two_lorentzian; each entry {largest M, lanes}, the last one open.
#pragma once

#include <stdint.h>

namespace bfhip {

struct TunedRange {
    uint32_t max_points;
    int lanes;
};
constexpr int kTunedRanges = 3;
constexpr TunedRange kTuned[4][kTunedRanges] = {
    {{512, 16}, {1024, 32}, {0xffffffffu, 64}},  // gaussian
    {{512, 16}, {1024, 32}, {0xffffffffu, 64}},  // lorentzian
    {{512, 16}, {0xffffffffu, 64}, {0xffffffffu, 64}},  // pseudo_voigt
    {{512, 16}, {1024, 32}, {0xffffffffu, 64}},  // two_lorentzian
};

}  // namespace bfhip
