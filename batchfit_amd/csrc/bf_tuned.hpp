// bf_tuned.hpp — generated by `python -m batchfit_amd.autotune --write`: lanes per
// fit by (model, M), measured on gfx950 (LM kernel time of the production numerics
// from true parameters; see batchfit_amd/autotune.py). Rows: gaussian, lorentzian,
// pseudo_voigt, two_lorentzian; each entry {largest M, lanes, streamed}, the last
// one open.
#pragma once

#include <stdint.h>

namespace bfhip {

struct TunedRange {
    uint32_t max_points;
    int lanes;
    int streamed;  // 1: points read from memory on every pass
};
constexpr int kTunedRanges = 5;
constexpr TunedRange kTuned[4][kTunedRanges] = {
    {{56, 8, 1}, {256, 8, 0}, {1792, 8, 1}, {0xffffffffu, 32, 1}, {0xffffffffu, 32, 1}},  // gaussian
    {{40, 8, 1}, {256, 8, 0}, {512, 16, 0}, {1280, 8, 1}, {0xffffffffu, 16, 1}},  // lorentzian
    {{80, 8, 1}, {224, 16, 0}, {256, 8, 0}, {0xffffffffu, 8, 1}, {0xffffffffu, 8, 1}},  // pseudo_voigt
    {{80, 8, 1}, {2560, 16, 1}, {0xffffffffu, 32, 1}, {0xffffffffu, 32, 1}, {0xffffffffu, 32, 1}},  // two_lorentzian
};

}  // namespace bfhip
