// bf_tuned.hpp — generated by `python -m batchfit_amd.autotune --write`: lanes per
// fit by (model, M), measured on gfx950 (LM kernel time of the production numerics
// from device-guess starts; see batchfit_amd/autotune.py). Rows: gaussian, lorentzian,
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
constexpr int kTunedRanges = 3;
constexpr TunedRange kTuned[4][kTunedRanges] = {
    {{1280, 4, 1}, {0xffffffffu, 16, 1}, {0xffffffffu, 16, 1}},  // gaussian
    {{1792, 4, 1}, {0xffffffffu, 16, 1}, {0xffffffffu, 16, 1}},  // lorentzian
    {{448, 4, 1}, {1792, 8, 1}, {0xffffffffu, 16, 1}},  // pseudo_voigt
    {{160, 4, 1}, {448, 8, 1}, {0xffffffffu, 16, 1}},  // two_lorentzian
};

}  // namespace bfhip
