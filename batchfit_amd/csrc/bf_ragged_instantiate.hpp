// bf_ragged_instantiate.hpp — included by bf_ragged_<model>.hip with BF_MODEL
// set; defines that model's ragged launchers (four mapping classes × the
// plain FP32 and compensated-mixed numerics).
#pragma once

#include "bf_guess_wide.hpp"
#include "bf_ragged.hpp"

#include <algorithm>

#ifndef BF_MODEL
#error "define BF_MODEL before including bf_ragged_instantiate.hpp"
#endif

namespace bfhip {

// Data-only guess of single-peak fits of a ragged batch whose edge windows
// hold at most 8·S values: one fit per group of 8 lanes, eight per wave, the
// stages of guess_single_kernel on each fit's own row, grid and M.
template <int MODEL, int S>
__global__ void __launch_bounds__(256) guess_ragged_single_kernel(
    const bf_fit_desc *fits, const uint32_t *order, uint32_t count, const bf_grid_desc *grids,
    const GridInfo *grid_info, const float *x, const float *y, float *guess) {
    const uint32_t task = blockIdx.x * 32 + threadIdx.x / 8;
    if (task >= count)
        return;  // the whole group: groups share no lanes' results
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
    g.m_pad = 0;
    g.extra_n = 0;
    g.params = guess + size_t(d.out_row) * 8;
    g.baseline = nullptr;
    g.gstatus = nullptr;
    float level, noise;
    sort_levels<S>(g, 0, int(threadIdx.x & 7), level, noise);
    peak_stage<MODEL, 8>(g, 0, level, noise);
}

// Data-only guess of ragged fits too long for the one-wave guess's LDS: the
// wide guess, one wave per fit, four per block.
template <int MODEL>
__global__ void __launch_bounds__(64 * kWideWaves) guess_ragged_wide_kernel(
    const bf_fit_desc *fits, const uint32_t *order, uint32_t count, const bf_grid_desc *grids,
    const GridInfo *grid_info, const float *x, const float *y, float *guess) {
    constexpr int P = Params<MODEL>::P;
    __shared__ float sfx[kWideWaves][MODEL == TWO_LORENTZIAN ? kWideSuffix : 1];
    const int wave = int(threadIdx.x / 64), lane = int(threadIdx.x & 63);
    const uint32_t task = blockIdx.x * kWideWaves + uint32_t(wave);
    if (task >= count)
        return;  // the whole wave
    const bf_fit_desc d = fits[order[task]];
    if (d.p0_row != 0xffffffffu)
        return;
    const GuessResult g = guess_core_wide<MODEL>(y + d.y_offset, x + grids[d.grid].x_offset,
                                                 int(d.points), grid_info[d.grid], sfx[wave],
                                                 lane);
    if (lane == 0)
        for (int k = 0; k < P; ++k)
            guess[size_t(d.out_row) * 8 + k] = g.p[k];
}

namespace {
template <int MODEL, int CLS, bool MIXED>
hipError_t ragged_go(const RaggedArgs &args, uint32_t blocks, hipStream_t stream) {
    constexpr int G = kRaggedClasses[CLS].G, R = kRaggedClasses[CLS].R;
    if constexpr (R == 0) {
        // Streamed class: retiring groups (one refill per wave to start with,
        // at most 4096 waves), then uncertainties one group per fit.
        const uint32_t waves = (args.count + pool_refill(G) - 1) / pool_refill(G);
        const uint32_t pool_blocks = std::min<uint32_t>((waves + 3) / 4, 1024u);
        RaggedArgs full = args;
        if constexpr (MODEL == TWO_LORENTZIAN) {
            // Fixed baseline: the form that sums only the free parameters'
            // rows, then the full form for batches with a wide grid.
            if (args.fixed_baseline) {
                if (args.flat_check)
                    hipLaunchKernelGGL((lm_ragged_pool_kernel<MODEL, G, MIXED, true, true>),
                                       dim3(pool_blocks), dim3(256), 0, stream, args);
                else
                    hipLaunchKernelGGL((lm_ragged_pool_kernel<MODEL, G, MIXED, false, true>),
                                       dim3(pool_blocks), dim3(256), 0, stream, args);
                if (hipError_t e = hipGetLastError(); e != hipSuccess)
                    return e;
                full.wide_only = 1;
            }
        }
        if (args.flat_check)
            hipLaunchKernelGGL((lm_ragged_pool_kernel<MODEL, G, MIXED, true>), dim3(pool_blocks),
                               dim3(256), 0, stream, full);
        else
            hipLaunchKernelGGL((lm_ragged_pool_kernel<MODEL, G, MIXED, false>), dim3(pool_blocks),
                               dim3(256), 0, stream, full);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || !args.cov)
            return e;
        constexpr uint32_t per_block = 256 / G;
        const dim3 cov_blocks((args.count + per_block - 1) / per_block);
        if constexpr (MODEL == TWO_LORENTZIAN) {
            if (args.fixed_baseline) {
                hipLaunchKernelGGL((lm_ragged_cov_kernel<MODEL, G, MIXED, true>), cov_blocks,
                                   dim3(256), 0, stream, args);
                if ((e = hipGetLastError()) != hipSuccess)
                    return e;
            }
        }
        hipLaunchKernelGGL((lm_ragged_cov_kernel<MODEL, G, MIXED>), cov_blocks, dim3(256), 0,
                           stream, full);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((lm_ragged_kernel<MODEL, G, R, MIXED, MIXED>), dim3(blocks), dim3(256), 0,
                       stream, args);
    return hipGetLastError();
}
template <int MODEL, bool MIXED>
hipError_t ragged_by_class(int cls, const RaggedArgs &args, uint32_t blocks, hipStream_t stream) {
    switch (cls) {
    case 0: return ragged_go<MODEL, 0, MIXED>(args, blocks, stream);
    case 1: return ragged_go<MODEL, 1, MIXED>(args, blocks, stream);
    case 2: return ragged_go<MODEL, 2, MIXED>(args, blocks, stream);
    case 3: return ragged_go<MODEL, 3, MIXED>(args, blocks, stream);
    case 4: return ragged_go<MODEL, 4, MIXED>(args, blocks, stream);
    case 5: return ragged_go<MODEL, 5, MIXED>(args, blocks, stream);
    case 6: return ragged_go<MODEL, 6, MIXED>(args, blocks, stream);
    case 7: return ragged_go<MODEL, 7, MIXED>(args, blocks, stream);
    case 8: return ragged_go<MODEL, 8, MIXED>(args, blocks, stream);
    case 9: return ragged_go<MODEL, 9, MIXED>(args, blocks, stream);
    case 10: return ragged_go<MODEL, 10, MIXED>(args, blocks, stream);
    case 11: return ragged_go<MODEL, 11, MIXED>(args, blocks, stream);
    case 12: return ragged_go<MODEL, 12, MIXED>(args, blocks, stream);
    }
    return hipErrorInvalidValue;
}
template <int MODEL, int CLS, bool MIXED> const void *ragged_symbol() {
    constexpr int G = kRaggedClasses[CLS].G, R = kRaggedClasses[CLS].R;
    return reinterpret_cast<const void *>(&lm_ragged_kernel<MODEL, G, R, MIXED, MIXED>);
}
// The retiring kernel of a streamed class (null for a tile class).
template <int MODEL, int CLS, bool MIXED> const void *pool_symbol() {
    constexpr int G = kRaggedClasses[CLS].G, R = kRaggedClasses[CLS].R;
    if constexpr (R == 0)
        return reinterpret_cast<const void *>(&lm_ragged_pool_kernel<MODEL, G, MIXED, true>);
    else
        return nullptr;
}
template <int MODEL, int CLS, bool MIXED> const void *pool_symbol_guessed() {
    constexpr int G = kRaggedClasses[CLS].G, R = kRaggedClasses[CLS].R;
    if constexpr (R == 0)
        return reinterpret_cast<const void *>(&lm_ragged_pool_kernel<MODEL, G, MIXED, false>);
    else
        return nullptr;
}
// Its uncertainty kernel (null for a tile class).
template <int MODEL, int CLS, bool MIXED> const void *cov_symbol() {
    constexpr int G = kRaggedClasses[CLS].G, R = kRaggedClasses[CLS].R;
    if constexpr (R == 0)
        return reinterpret_cast<const void *>(&lm_ragged_cov_kernel<MODEL, G, MIXED>);
    else
        return nullptr;
}
}  // namespace

// Loads the code object of a class's kernels (hipFuncGetAttributes) so that
// the first use inside a stream capture does not have to.
template <> hipError_t prepare_ragged<BF_MODEL>(int cls, bool mixed, bool guess) {
    static constexpr const void *(*table[kRaggedClassCount][2])() = {
        {&ragged_symbol<BF_MODEL, 0, false>, &ragged_symbol<BF_MODEL, 0, true>},
        {&ragged_symbol<BF_MODEL, 1, false>, &ragged_symbol<BF_MODEL, 1, true>},
        {&ragged_symbol<BF_MODEL, 2, false>, &ragged_symbol<BF_MODEL, 2, true>},
        {&ragged_symbol<BF_MODEL, 3, false>, &ragged_symbol<BF_MODEL, 3, true>},
        {&ragged_symbol<BF_MODEL, 4, false>, &ragged_symbol<BF_MODEL, 4, true>},
        {&ragged_symbol<BF_MODEL, 5, false>, &ragged_symbol<BF_MODEL, 5, true>},
        {&ragged_symbol<BF_MODEL, 6, false>, &ragged_symbol<BF_MODEL, 6, true>},
        {&ragged_symbol<BF_MODEL, 7, false>, &ragged_symbol<BF_MODEL, 7, true>},
        {&ragged_symbol<BF_MODEL, 8, false>, &ragged_symbol<BF_MODEL, 8, true>},
        {&ragged_symbol<BF_MODEL, 9, false>, &ragged_symbol<BF_MODEL, 9, true>},
        {&ragged_symbol<BF_MODEL, 10, false>, &ragged_symbol<BF_MODEL, 10, true>},
        {&ragged_symbol<BF_MODEL, 11, false>, &ragged_symbol<BF_MODEL, 11, true>},
        {&ragged_symbol<BF_MODEL, 12, false>, &ragged_symbol<BF_MODEL, 12, true>}};
    if (cls < 0 || cls >= kRaggedClassCount)
        return hipErrorInvalidValue;
    const void *sym = table[cls][mixed ? 1 : 0]();
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, sym);
    if (e == hipSuccess && kRaggedClasses[cls].R == 0) {
        static constexpr const void *(*pools[kRaggedClassCount][2])() = {
            {&pool_symbol<BF_MODEL, 0, false>, &pool_symbol<BF_MODEL, 0, true>},
            {&pool_symbol<BF_MODEL, 1, false>, &pool_symbol<BF_MODEL, 1, true>},
            {&pool_symbol<BF_MODEL, 2, false>, &pool_symbol<BF_MODEL, 2, true>},
            {&pool_symbol<BF_MODEL, 3, false>, &pool_symbol<BF_MODEL, 3, true>},
            {&pool_symbol<BF_MODEL, 4, false>, &pool_symbol<BF_MODEL, 4, true>},
            {&pool_symbol<BF_MODEL, 5, false>, &pool_symbol<BF_MODEL, 5, true>},
            {&pool_symbol<BF_MODEL, 6, false>, &pool_symbol<BF_MODEL, 6, true>},
            {&pool_symbol<BF_MODEL, 7, false>, &pool_symbol<BF_MODEL, 7, true>},
            {&pool_symbol<BF_MODEL, 8, false>, &pool_symbol<BF_MODEL, 8, true>},
            {&pool_symbol<BF_MODEL, 9, false>, &pool_symbol<BF_MODEL, 9, true>},
            {&pool_symbol<BF_MODEL, 10, false>, &pool_symbol<BF_MODEL, 10, true>},
            {&pool_symbol<BF_MODEL, 11, false>, &pool_symbol<BF_MODEL, 11, true>},
        {&pool_symbol<BF_MODEL, 12, false>, &pool_symbol<BF_MODEL, 12, true>}};
        e = hipFuncGetAttributes(&attr, pools[cls][mixed ? 1 : 0]());
        static constexpr const void *(*guessed[kRaggedClassCount][2])() = {
            {&pool_symbol_guessed<BF_MODEL, 0, false>, &pool_symbol_guessed<BF_MODEL, 0, true>},
            {&pool_symbol_guessed<BF_MODEL, 1, false>, &pool_symbol_guessed<BF_MODEL, 1, true>},
            {&pool_symbol_guessed<BF_MODEL, 2, false>, &pool_symbol_guessed<BF_MODEL, 2, true>},
            {&pool_symbol_guessed<BF_MODEL, 3, false>, &pool_symbol_guessed<BF_MODEL, 3, true>},
            {&pool_symbol_guessed<BF_MODEL, 4, false>, &pool_symbol_guessed<BF_MODEL, 4, true>},
            {&pool_symbol_guessed<BF_MODEL, 5, false>, &pool_symbol_guessed<BF_MODEL, 5, true>},
            {&pool_symbol_guessed<BF_MODEL, 6, false>, &pool_symbol_guessed<BF_MODEL, 6, true>},
            {&pool_symbol_guessed<BF_MODEL, 7, false>, &pool_symbol_guessed<BF_MODEL, 7, true>},
            {&pool_symbol_guessed<BF_MODEL, 8, false>, &pool_symbol_guessed<BF_MODEL, 8, true>},
            {&pool_symbol_guessed<BF_MODEL, 9, false>, &pool_symbol_guessed<BF_MODEL, 9, true>},
            {&pool_symbol_guessed<BF_MODEL, 10, false>, &pool_symbol_guessed<BF_MODEL, 10, true>},
            {&pool_symbol_guessed<BF_MODEL, 11, false>, &pool_symbol_guessed<BF_MODEL, 11, true>},
        {&pool_symbol_guessed<BF_MODEL, 12, false>, &pool_symbol_guessed<BF_MODEL, 12, true>}};
        if (e == hipSuccess)
            e = hipFuncGetAttributes(&attr, guessed[cls][mixed ? 1 : 0]());
        static constexpr const void *(*covs[kRaggedClassCount][2])() = {
            {&cov_symbol<BF_MODEL, 0, false>, &cov_symbol<BF_MODEL, 0, true>},
            {&cov_symbol<BF_MODEL, 1, false>, &cov_symbol<BF_MODEL, 1, true>},
            {&cov_symbol<BF_MODEL, 2, false>, &cov_symbol<BF_MODEL, 2, true>},
            {&cov_symbol<BF_MODEL, 3, false>, &cov_symbol<BF_MODEL, 3, true>},
            {&cov_symbol<BF_MODEL, 4, false>, &cov_symbol<BF_MODEL, 4, true>},
            {&cov_symbol<BF_MODEL, 5, false>, &cov_symbol<BF_MODEL, 5, true>},
            {&cov_symbol<BF_MODEL, 6, false>, &cov_symbol<BF_MODEL, 6, true>},
            {&cov_symbol<BF_MODEL, 7, false>, &cov_symbol<BF_MODEL, 7, true>},
            {&cov_symbol<BF_MODEL, 8, false>, &cov_symbol<BF_MODEL, 8, true>},
            {&cov_symbol<BF_MODEL, 9, false>, &cov_symbol<BF_MODEL, 9, true>},
            {&cov_symbol<BF_MODEL, 10, false>, &cov_symbol<BF_MODEL, 10, true>},
            {&cov_symbol<BF_MODEL, 11, false>, &cov_symbol<BF_MODEL, 11, true>},
        {&cov_symbol<BF_MODEL, 12, false>, &cov_symbol<BF_MODEL, 12, true>}};
        if (e == hipSuccess)
            e = hipFuncGetAttributes(&attr, covs[cls][mixed ? 1 : 0]());
    }
    if (e == hipSuccess && guess)
        e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&guess_ragged_kernel<BF_MODEL>));
    if (e == hipSuccess && guess)
        e = hipFuncGetAttributes(
            &attr, reinterpret_cast<const void *>(&guess_ragged_wide_kernel<BF_MODEL>));
    if constexpr (BF_MODEL != TWO_LORENTZIAN) {
        if (e == hipSuccess && guess)
            e = hipFuncGetAttributes(
                &attr, reinterpret_cast<const void *>(&guess_ragged_single_kernel<BF_MODEL, 4>));
        if (e == hipSuccess && guess)
            e = hipFuncGetAttributes(
                &attr, reinterpret_cast<const void *>(&guess_ragged_single_kernel<BF_MODEL, 8>));
        if (e == hipSuccess && guess)
            e = hipFuncGetAttributes(
                &attr, reinterpret_cast<const void *>(&guess_ragged_single_kernel<BF_MODEL, 16>));
    }
    return e;
}

template <> hipError_t launch_ragged_lm<BF_MODEL>(int cls, bool mixed, const RaggedArgs &args,
                                                  uint32_t blocks, hipStream_t stream) {
    return mixed ? ragged_by_class<BF_MODEL, true>(cls, args, blocks, stream)
                 : ragged_by_class<BF_MODEL, false>(cls, args, blocks, stream);
}

template <>
hipError_t launch_ragged_guess<BF_MODEL>(const bf_fit_desc *fits, const uint32_t *order,
                                         uint32_t count, const bf_grid_desc *grids,
                                         const GridInfo *grid_info, const float *x, const float *y,
                                         float *guess, uint32_t m_pad, uint32_t extra_n,
                                         hipStream_t stream) {
    if constexpr (BF_MODEL != TWO_LORENTZIAN) {
        // Single-peak fits of at most kGuessSortedMaxPoints: eight per wave
        // (edge sorting networks, then the streamed peak), as a uniform batch.
        if (m_pad <= kGuessSortedMaxPoints) {
            const uint32_t edge = m_pad / 10 > 1 ? m_pad / 10 : 1, blocks = (count + 31) / 32;
            if (edge <= 32)
                hipLaunchKernelGGL((guess_ragged_single_kernel<BF_MODEL, 4>), dim3(blocks),
                                   dim3(256), 0, stream, fits, order, count, grids, grid_info, x,
                                   y, guess);
            else if (edge <= 64)
                hipLaunchKernelGGL((guess_ragged_single_kernel<BF_MODEL, 8>), dim3(blocks),
                                   dim3(256), 0, stream, fits, order, count, grids, grid_info, x,
                                   y, guess);
            else
                hipLaunchKernelGGL((guess_ragged_single_kernel<BF_MODEL, 16>), dim3(blocks),
                                   dim3(256), 0, stream, fits, order, count, grids, grid_info, x,
                                   y, guess);
            return hipGetLastError();
        }
    }
    const size_t lds = size_t(m_pad + extra_n) * sizeof(float);
    if (m_pad > kGuessWaveMaxPoints || lds > kGuessLdsBytes) {  // by the slice's longest fit
        hipLaunchKernelGGL((guess_ragged_wide_kernel<BF_MODEL>),
                           dim3((count + kWideWaves - 1) / kWideWaves), dim3(64 * kWideWaves), 0,
                           stream, fits, order, count, grids, grid_info, x, y, guess);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((guess_ragged_kernel<BF_MODEL>), dim3(count), dim3(64), lds, stream, fits,
                       order, count, grids, grid_info, x, y, guess, m_pad, extra_n);
    return hipGetLastError();
}

}  // namespace bfhip
