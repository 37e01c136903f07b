// bf_guess_wide.hpp — the data-only initial guess (cpu/initial_guess.hpp:
// 120-249) of spectra too long for the one-wave guess's LDS, for every model:
// one wave per fit with nothing staged. The edge levels are the memory-read
// searches of stream_levels; the baseline-removed signal is formed from y and
// x at every read with the same FP32 operations (contraction off) as the
// staged signal, so each value has the same bits; the suffix minima of the
// second-peak search cover runs of 2^k 64-point steps in a fixed LDS table.
// The result equals the one-wave guess and the FP32 reference guess bit for
// bit, and M has no limit (the reference keeps its scratch on the heap,
// initial_guess.hpp:143).
#pragma once

#include "bf_guess_group.hpp"

#pragma clang fp contract(off)

namespace bfhip {

// The signal y[j] − (level + slope·(x[j] − left_x)) of guess_core, or
// y[j] − level where guess_core does not read the grid, formed at every read.
struct MemSignal {
    const float *y, *x;
    float level, slope, left_x;
    bool grid_term;
    __device__ __forceinline__ float operator[](int j) const {
        const float v = y[j];
        return grid_term ? v - (level + slope * (x[j] - left_x)) : v - level;
    }
};

// Run minima per wave of the wide guess (floats): runs of 2^k steps keep
// two per run, ((M − 2) / 64 >> k) + 1 runs, within this table.
constexpr int kWideSuffix = 1024;
constexpr int kWideWaves = 4;  // waves (fits) per block

// The guess of the fit whose row is y[0..m) by one whole wave; sfx holds
// kWideSuffix floats of the wave's own.
template <int MODEL>
__device__ GuessResult guess_core_wide(const float *y, const float *x, int m, const GridInfo &grid,
                                       float *sfx, int lane) {
    constexpr bool paired = MODEL == TWO_LORENTZIAN;
    GuessResult res;
    for (int k = 0; k < 8; ++k)
        res.p[k] = qnan_f();
    res.baseline = qnan_f();
    res.status = GS_INVALID;
    if (m < 5 || !grid.guess_grid_ok)
        return res;
    int bad = 0;
    for (int j = lane; j < m; j += 64)
        bad |= !finite_f(y[j]);
    if (wave_or_i(bad))
        return res;
    const EdgeLevels lv = stream_levels(y, m, grid.count, lane);
    res.baseline = lv.both;
    const float left_x = grid.left_x, right_x = grid.right_x;
    const float slope = paired ? (lv.right - lv.left) / (right_x - left_x) : 0.0f;
    const float level = paired ? lv.left : lv.both;
    if (!finite_f(slope))
        return res;
    const float noise = 1.4826f * lv.spread;
    const MemSignal sig{y, x, level, slope, left_x,
                        paired || __float_as_uint(level) == 0x80000000u};
    for (int j = lane; j < m; j += 64)
        bad |= !finite_f(sig[j]);
    if (wave_or_i(bad) || !finite_f(noise))
        return res;
    int shift = 0;
    while (((m - 2) / 64 >> shift) >= kWideSuffix / 2)
        ++shift;
    return guess_tail<MODEL>(sig, sfx, shift, x, m, grid, lv.left, slope, noise, lane, res);
}

// Uniform batches (fit-major rows, ys_pt = 1): fit = 4·block + wave.
template <int MODEL> __global__ void __launch_bounds__(256) guess_wide_kernel(GuessArgs a) {
    constexpr int P = Params<MODEL>::P;
    __shared__ float sfx[kWideWaves][MODEL == TWO_LORENTZIAN ? kWideSuffix : 1];
    const int wave = int(threadIdx.x / 64), lane = int(threadIdx.x & 63);
    const uint32_t fit = blockIdx.x * kWideWaves + uint32_t(wave);
    if (fit >= a.n)
        return;  // the whole wave: the waves of a block share no barrier
    const GuessResult g = guess_core_wide<MODEL>(a.y + size_t(fit) * size_t(a.ys_fit), a.x,
                                                 int(a.m), *a.grid, sfx[wave], lane);
    if (lane == 0) {
        for (int k = 0; k < P; ++k)
            a.params[size_t(fit) * P + k] = g.p[k];
        if (a.baseline)
            a.baseline[fit] = g.baseline;
        if (a.gstatus)
            a.gstatus[fit] = g.status;
    }
}

}  // namespace bfhip

// The code that follows this header keeps hipcc's default contraction.
#pragma clang fp contract(fast)
