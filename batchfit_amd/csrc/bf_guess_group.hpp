// bf_guess_group.hpp — data-only initial guess of the single-peak models
// (Gaussian, Lorentzian, pseudo-Voigt; cpu/initial_guess.hpp:120-197) in two
// stages with several fits per wave: the edge levels (sorting networks in
// 8-lane groups for edge windows up to 128 values, bisection in 64-lane groups
// up to 1024, one wave per fit reading the keys from memory beyond), then the
// peak by groups of 8 lanes, streaming the row. Equal to the FP32 reference
// guess bit for bit, like the one-wave-per-fit kernel of bf_guess.hpp, which
// serves the two-peak model, point-major data and edge windows of 129–230
// values. No LDS.
//
// The bisection medians are exact order statistics with each group's own
// search state in VGPRs, so the groups of a wave search at once and the wave
// runs as many steps as its slowest median.
#pragma once

#include "bf_guess.hpp"

#include <type_traits>

#pragma clang fp contract(off)

namespace bfhip {

// Reductions of 32-bit integers over aligned groups of G = 2 … 64 lanes
// (every lane gets the result): xlane<1>, <2> (quad_perm), <4> (row_half_mirror), <8> (row_mirror),
// <16> (ds_swizzle), <32> (across the halves).
template <int G, typename T, typename Op> __device__ __forceinline__ T group_fold(T v, Op op) {
    static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "G is a power of two up to 64");
    static_assert(sizeof(T) == 4, "32-bit integers");
    auto move = [](auto step, T w) { return T(xlane<decltype(step)::value>(int(w))); };
    v = op(v, move(std::integral_constant<int, 1>{}, v));
    if constexpr (G >= 4)
        v = op(v, move(std::integral_constant<int, 2>{}, v));
    if constexpr (G >= 8)
        v = op(v, move(std::integral_constant<int, 4>{}, v));
    if constexpr (G >= 16)
        v = op(v, move(std::integral_constant<int, 8>{}, v));
    if constexpr (G >= 32)
        v = op(v, move(std::integral_constant<int, 16>{}, v));
    if constexpr (G >= 64)
        v = op(v, move(std::integral_constant<int, 32>{}, v));
    return v;
}
template <int G> __device__ __forceinline__ uint32_t group_min_u(uint32_t v) {
    return group_fold<G>(v, [](uint32_t a, uint32_t b) { return min(a, b); });
}
template <int G> __device__ __forceinline__ uint32_t group_max_u(uint32_t v) {
    return group_fold<G>(v, [](uint32_t a, uint32_t b) { return max(a, b); });
}
template <int G> __device__ __forceinline__ int group_sum_i(int v) {
    return group_fold<G>(v, [](int a, int b) { return a + b; });
}
// Largest value of the group, ties to the smallest index (finite values).
template <int G> __device__ __forceinline__ void group_argmax(float &v, int &idx) {
    static_assert(G >= 8 && G <= 64, "groups of 8 to 64 lanes");
    auto step = [&](float w, int wi) {
        const bool take = w > v || (w == v && wi < idx);
        v = take ? w : v;
        idx = take ? wi : idx;
    };
    step(__int_as_float(xlane<1>(__float_as_int(v))), xlane<1>(idx));
    step(__int_as_float(xlane<2>(__float_as_int(v))), xlane<2>(idx));
    step(__int_as_float(xlane<4>(__float_as_int(v))), xlane<4>(idx));
    if constexpr (G >= 16)
        step(__int_as_float(xlane<8>(__float_as_int(v))), xlane<8>(idx));
    if constexpr (G >= 32)
        step(__int_as_float(xlane<16>(__float_as_int(v))), xlane<16>(idx));
    if constexpr (G >= 64)
        step(__int_as_float(xlane<32>(__float_as_int(v))), xlane<32>(idx));
}

// Number of the group's keys at or below p.
template <int G, int S>
__device__ __forceinline__ int group_count_le(const uint32_t (&key)[S], uint32_t p) {
    int c = 0;
#pragma unroll
    for (int q = 0; q < S; ++q)
        c += key[q] <= p ? 1 : 0;
    return group_sum_i<G>(c);
}

// Result of a median search: the median value, and the middle keys (for an
// even count the lower and upper middle, else the middle key twice).
struct MedianKeys {
    float value;
    uint32_t lower, upper;
};

// A group's keys held in registers: S slots per lane, 0xffffffff in slots
// that hold no value (above every finite key).
template <int G, int S> struct RegKeys {
    const uint32_t (&key)[S];
    __device__ int count_le(uint32_t p) const { return group_count_le<G, S>(key, p); }
    __device__ void range(uint32_t &kmin, uint32_t &kmax) const {
        kmin = 0xffffffffu;
        kmax = 0u;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            kmin = min(kmin, key[q]);
            kmax = max(kmax, key[q] != 0xffffffffu ? key[q] : 0u);
        }
        kmin = group_min_u<G>(kmin);
        kmax = group_max_u<G>(kmax);
    }
    // The largest key in [lo, hi] (0 if none).
    __device__ uint32_t max_in(uint32_t lo, uint32_t hi) const {
        uint32_t found = 0u;
#pragma unroll
        for (int q = 0; q < S; ++q)
            found = (key[q] >= lo && key[q] <= hi && key[q] > found) ? key[q] : found;
        return group_max_u<G>(found);
    }
};

// Keys read from memory at every use by one whole wave: key_at(e) for the
// window's elements e < n (lane l takes l, l + 64, …).
template <typename KeyAt> struct StreamKeys {
    KeyAt key_at;
    int n;
    __device__ int count_le(uint32_t p) const {
        int c = 0;
        for (int e = int(threadIdx.x & 63); e < n; e += 64)
            c += key_at(e) <= p ? 1 : 0;
        return group_sum_i<64>(c);
    }
    __device__ void range(uint32_t &kmin, uint32_t &kmax) const {
        kmin = 0xffffffffu;
        kmax = 0u;
        for (int e = int(threadIdx.x & 63); e < n; e += 64) {
            const uint32_t k = key_at(e);
            kmin = min(kmin, k);
            kmax = max(kmax, k);
        }
        kmin = group_min_u<64>(kmin);
        kmax = group_max_u<64>(kmax);
    }
    __device__ uint32_t max_in(uint32_t lo, uint32_t hi) const {
        uint32_t found = 0u;
        for (int e = int(threadIdx.x & 63); e < n; e += 64) {
            const uint32_t k = key_at(e);
            found = (k >= lo && k <= hi && k > found) ? k : found;
        }
        return group_max_u<64>(found);
    }
};

// Median of the n keys of a key set (initial_guess.hpp:25-30: the middle
// order statistic, and for an even n the midpoint of it and the largest value
// below it). The k-th key (k = n/2) is searched in [lo_key, hi_key], which
// must hold it, with lo keys below lo_key and hi keys at or below hi_key. Each
// step counts the keys at or below a pivot in [lo_key, hi_key) and keeps the
// side that holds the k-th key, so the interval shrinks every step; the
// pivot is the midpoint of the two end values as floats (noise around one
// level is bisected by value, not by sign and exponent bits) for 12 steps,
// then the midpoint of the keys, so the search always ends. It stops when the
// interval holds one key (the k-th) or one value. Each group keeps its own
// state; the wave loops until every group of it has its key.
template <typename Keys>
__device__ MedianKeys median_in(const Keys &keys, int n, uint32_t lo_key, uint32_t hi_key, int lo,
                                int hi) {
    const int k = n / 2;
    int steps = 0;
    bool done = lo_key == hi_key || hi - lo == 1;
    while (__ballot(!done)) {
        if (!done) {
            const uint32_t width = hi_key - lo_key;  // ≥ 1
            uint32_t pivot = float_key(0.5f * key_float(lo_key) + 0.5f * key_float(hi_key));
            if (steps++ >= 12 || pivot < lo_key || pivot >= hi_key)
                pivot = lo_key + (width >> 1);  // in [lo_key, hi_key)
            const int c = keys.count_le(pivot);
            if (c < k + 1) {
                lo_key = pivot + 1u;
                lo = c;
            } else {
                hi_key = pivot;
                hi = c;
            }
            done = hi - lo == 1 || lo_key == hi_key;
        }
    }
    // One key in [lo_key, hi_key] (or all of them one value): the k-th.
    const uint32_t t = lo_key != hi_key ? keys.max_in(lo_key, hi_key) : lo_key;
    if (n & 1)
        return {key_float(t), t, t};
    // lo counts the keys below lo_key, and no key lies in [lo_key, t): so lo
    // keys are below t. If the k smallest are all below t, the lower middle
    // is the largest of them; else it equals t.
    const uint32_t lower = (lo >= k && t > 0u) ? keys.max_in(0u, t - 1u) : t;
    return {std_midpoint(key_float(lower), key_float(t)), lower, t};
}

// The search of median_in started on [a, b] (a ≤ b keys): if the k-th key
// lies there, the search starts there; else on the side of it that holds the
// k-th key. Two counts; any a ≤ b gives the same result.
template <typename Keys>
__device__ MedianKeys median_around(const Keys &keys, int n, uint32_t a, uint32_t b) {
    uint32_t kmin, kmax;
    keys.range(kmin, kmax);
    a = a < kmin ? kmin : a;
    b = b > kmax ? kmax : b;
    const int k = n / 2;
    if (a > b)  // an empty bracket: search everything
        return median_in(keys, n, kmin, kmax, 0, n);
    const int below_a = a > kmin ? keys.count_le(a - 1u) : 0;
    const int upto_b = keys.count_le(b);
    if (below_a > k)  // the k-th key is below a
        return median_in(keys, n, kmin, a - 1u, 0, below_a);
    if (upto_b <= k)  // the k-th key is above b
        return median_in(keys, n, b + 1u, kmax, upto_b, n);
    return median_in(keys, n, a, b, below_a, upto_b);
}

template <int G, int S>
__device__ MedianKeys group_median_around(const uint32_t (&key)[S], int n, uint32_t a, uint32_t b) {
    return median_around(RegKeys<G, S>{key}, n, a, b);
}

// Keys of mean ∓ σ/2 of the group's values (the floats of its valid keys):
// for noise around a level, a bracket that holds the median and about 40 % of
// the values. Only a starting guess: group_median_around checks it.
template <int G, int S>
__device__ void mean_bracket(const float (&val)[S], const bool (&valid)[S], int n, uint32_t &a,
                             uint32_t &b) {
    float sum = 0.0f, sq = 0.0f;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const float v = valid[q] ? val[q] : 0.0f;
        sum += v;
        sq += v * v;
    }
    for (int o = 1; o < G; o <<= 1) {
        sum += __shfl_xor(sum, o, G);
        sq += __shfl_xor(sq, o, G);
    }
    const float inv = 1.0f / float(n);
    const float mean = sum * inv;
    const float var = sq * inv - mean * mean;
    const float half_sd = 0.5f * sqrtf(var > 0.0f ? var : 0.0f);
    a = float_key(mean - half_sd);
    b = float_key(mean + half_sd);
    if (!(a <= b)) {  // NaN or inf: no bracket
        a = 0u;
        b = 0xffffffffu;
    }
}

// DPP moves of one float: ctrl codes of gfx9 (row_shl:1 0x101, row_shr:1
// 0x111, row_ror:1 0x121, row_ror:15 0x12F, wave_shl:1 0x130, wave_rol:1
// 0x134, wave_shr:1 0x138, wave_ror:1 0x13C, row_half_mirror 0x141). A lane
// whose source lies outside its row (or the wave) keeps `old`.
template <int CTRL> __device__ __forceinline__ float dpp_f(float old, float src) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xF, 0xF, false));
}
// A rotation (every lane has a source).
template <int CTRL> __device__ __forceinline__ float dpp_rot(float src) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), CTRL, 0xF, 0xF, false));
}
// The value at point j − 1 for the point of slot k of every lane (cur = slot
// k, prev = slot k − 1): from the lane below, or for the group's first lane
// from the group's last lane one slot earlier. Groups of 8 lanes share a DPP
// row, so their first lane takes row_half_mirror of prev (lane 7 of its
// group) by a select.
template <int G> __device__ __forceinline__ float from_left(float prev, float cur) {
    if constexpr (G == 8) {
        const float wrap = dpp_rot<0x141>(prev), below = dpp_rot<0x111>(cur);
        return (threadIdx.x & 7) == 0 ? wrap : below;
    } else if constexpr (G == 16) {
        return dpp_f<0x111>(dpp_rot<0x121>(prev), cur);
    } else {
        return dpp_f<0x138>(dpp_rot<0x13C>(prev), cur);
    }
}
// The value at point j + 1 (cur = slot k, next = slot k + 1).
template <int G> __device__ __forceinline__ float from_right(float cur, float next) {
    if constexpr (G == 8) {
        const float wrap = dpp_rot<0x141>(next), above = dpp_rot<0x101>(cur);
        return (threadIdx.x & 7) == 7 ? wrap : above;
    } else if constexpr (G == 16) {
        return dpp_f<0x101>(dpp_rot<0x12F>(next), cur);
    } else {
        return dpp_f<0x130>(dpp_rot<0x134>(next), cur);
    }
}

// Stage 1: the edge levels of a fit (initial_guess.hpp:142-169) by a group of
// G lanes, CE edge values per lane (max(1, M/10) ≤ G·CE): the baseline (the
// median of both edges) and the noise scale (1.4826 × the median absolute
// deviation from each edge's own median), written into the fit's first two
// parameter slots for stage 2. Few registers, so many waves per SIMD hide the
// latency of the searches' reduction chains.
template <int G, int CE>
__global__ void __launch_bounds__(256) guess_levels_kernel(GuessArgs a, int param_stride) {
    const int gl = int(threadIdx.x & (G - 1));
    const uint32_t fit = blockIdx.x * (256 / G) + threadIdx.x / G;
    const bool live = fit < a.n;
    const uint32_t f = live ? fit : a.n - 1;  // a group past the end repeats the last fit
    const int m = int(a.m);
    const int count = a.grid->count;
    const float *y = a.y + size_t(f) * size_t(a.ys_fit);
    float left[CE], right[CE];
    bool in[CE];
    uint32_t kl[CE], kr[CE], ku[2 * CE];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = gl + G * i;
        in[i] = e < count;
        left[i] = in[i] ? y[e] : 0.0f;
        right[i] = in[i] ? y[m - count + e] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        kl[i] = in[i] ? float_key(left[i]) : 0xffffffffu;
        kr[i] = in[i] ? float_key(right[i]) : 0xffffffffu;
        ku[i] = kl[i];
        ku[CE + i] = kr[i];
    }
    uint32_t ba, bb;
    mean_bracket<G, CE>(left, in, count, ba, bb);
    const MedianKeys lm = group_median_around<G, CE>(kl, count, ba, bb);
    mean_bracket<G, CE>(right, in, count, ba, bb);
    const MedianKeys rm = group_median_around<G, CE>(kr, count, ba, bb);
    // The median of both edges lies between the two edges' middle values
    // (each edge has more than half its values on either side of its middle
    // pair, so more than half of the union does).
    const MedianKeys um = group_median_around<G, 2 * CE>(
        ku, 2 * count, min(lm.lower, rm.lower), max(lm.upper, rm.upper));
    const float left_level = lm.value, right_level = rm.value, level = um.value;
    float dev[2 * CE];
    bool in2[2 * CE];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        dev[i] = fabsf(left[i] - left_level);
        dev[CE + i] = fabsf(right[i] - right_level);
        in2[i] = in2[CE + i] = in[i];
        ku[i] = in[i] ? float_key(dev[i]) : 0xffffffffu;
        ku[CE + i] = in[i] ? float_key(dev[CE + i]) : 0xffffffffu;
    }
    mean_bracket<G, 2 * CE>(dev, in2, 2 * count, ba, bb);
    const float spread = group_median_around<G, 2 * CE>(ku, 2 * count, ba, bb).value;
    if (live && gl == 0) {
        a.params[size_t(fit) * param_stride] = level;
        a.params[size_t(fit) * param_stride + 1] = 1.4826f * spread;
    }
}

// Edge levels of one fit's row y[0..m) by one whole wave, the window keys
// read from memory at every step of the searches (the two edges stay in the
// L2 cache), so M has no limit: the median of each edge, of both, and of the
// absolute deviations from each edge's own median (initial_guess.hpp:142-165).
__device__ EdgeLevels stream_levels(const float *y, int m, int count, int lane) {
    const float *right = y + (m - count);
    // mean ∓ σ/2 of a window, as mean_bracket.
    auto bracket = [&](auto value_at, int n, uint32_t &lo, uint32_t &hi) {
        float sum = 0.0f, sq = 0.0f;
        for (int e = lane; e < n; e += 64) {
            const float v = value_at(e);
            sum += v;
            sq += v * v;
        }
        for (int o = 1; o < 64; o <<= 1) {
            sum += __shfl_xor(sum, o, 64);
            sq += __shfl_xor(sq, o, 64);
        }
        const float mean = sum / float(n), var = sq / float(n) - mean * mean;
        const float half_sd = 0.5f * sqrtf(var > 0.0f ? var : 0.0f);
        lo = float_key(mean - half_sd);
        hi = float_key(mean + half_sd);
        if (!(lo <= hi)) {
            lo = 0u;
            hi = 0xffffffffu;
        }
    };
    auto left_v = [&](int e) { return y[e]; };
    auto right_v = [&](int e) { return right[e]; };
    auto union_v = [&](int e) { return e < count ? y[e] : right[e - count]; };
    uint32_t ba, bb;
    bracket(left_v, count, ba, bb);
    auto lk = [&](int e) { return float_key(y[e]); };
    const MedianKeys lm = median_around(StreamKeys<decltype(lk)>{lk, count}, count, ba, bb);
    bracket(right_v, count, ba, bb);
    auto rk = [&](int e) { return float_key(right[e]); };
    const MedianKeys rm = median_around(StreamKeys<decltype(rk)>{rk, count}, count, ba, bb);
    auto uk = [&](int e) { return float_key(union_v(e)); };
    const MedianKeys um = median_around(StreamKeys<decltype(uk)>{uk, 2 * count}, 2 * count,
                                        min(lm.lower, rm.lower), max(lm.upper, rm.upper));
    const float left_level = lm.value, right_level = rm.value;
    auto dev_v = [&](int e) {
        return e < count ? fabsf(y[e] - left_level) : fabsf(right[e - count] - right_level);
    };
    bracket(dev_v, 2 * count, ba, bb);
    auto dk = [&](int e) { return float_key(dev_v(e)); };
    const float spread =
        median_around(StreamKeys<decltype(dk)>{dk, 2 * count}, 2 * count, ba, bb).value;
    return {left_level, right_level, um.value, spread};
}

// Stage 1 for edge windows beyond 1024 values (M ≥ 10,250): one wave per fit,
// stream_levels, so M has no limit.
template <int P> __global__ void __launch_bounds__(256) guess_levels_stream_kernel(GuessArgs a) {
    constexpr int param_stride = P;
    const uint32_t fit = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = int(threadIdx.x & 63);
    const bool live = fit < a.n;
    const uint32_t f = live ? fit : a.n - 1;
    const EdgeLevels lv =
        stream_levels(a.y + size_t(f) * size_t(a.ys_fit), int(a.m), a.grid->count, lane);
    if (live && lane == 0) {
        a.params[size_t(fit) * param_stride] = lv.both;
        a.params[size_t(fit) * param_stride + 1] = 1.4826f * lv.spread;
    }
}

// Stage 1 for edge windows of up to 128 values (M < 1,290) by sorting: one
// fit per group of 8 lanes (eight per wave), each edge's keys in 8·S slots,
// element e = (lane in the group)·S + slot. The bitonic networks are in the
// form whose comparators all put the smaller key at the lower position (each
// merge of sorted halves opens with the mirror comparator e ↔ e ^ (2^p − 1)):
// a pair inside a lane is a min and a max, a pair across lanes a DPP move
// (quad_perm xor 1, reverse of a quad, xor 2, row_half_mirror) or a ds_swizzle
// (xor 4) and a min or max. The edge medians are read at fixed positions. For
// the medians of unions the half-cleaner identity serves: for X ascending and
// Y descending, c values each, min(X[i], Y[i]) over i < c are the c smallest
// of X ∪ Y and max(X[i], Y[i]) the c largest. No LDS and no data-dependent
// loops.
// Keys of two slots of one lane in order.
__device__ __forceinline__ void key_cx(uint32_t &a, uint32_t &b) {
    const uint32_t lo = min(a, b);
    b = max(a, b);
    a = lo;
}
// Move of a key from the partner lane: CTRL 1 (xor 1), 2 (xor 2), 3 (reverse
// of the quad: xor 3), 4 (xor 4, ds_swizzle), 7 (row_half_mirror: xor 7).
// Every source lane lies in the row (bound_ctrl never applies).
template <int CTRL> __device__ __forceinline__ uint32_t key_from(uint32_t v) {
    constexpr int ctrl = CTRL == 1 ? 0xB1 : CTRL == 2 ? 0x4E : CTRL == 3 ? 0x1B : 0x141;
    if constexpr (CTRL == 4)
        return uint32_t(__builtin_amdgcn_ds_swizzle(int(v), 0x101F));
    else
        return uint32_t(__builtin_amdgcn_mov_dpp(int(v), ctrl, 0xF, 0xF, true));
}
// Comparators e ↔ e ^ X inside each lane (X < S).
template <int S, int X> __device__ __forceinline__ void lane_stage(uint32_t (&v)[S]) {
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (s < (s ^ X))
            key_cx(v[s], v[s ^ X]);
}
template <int S, int X> __device__ __forceinline__ void lane_xors(uint32_t (&v)[S]) {
    if constexpr (X >= 1) {
        lane_stage<S, X>(v);
        lane_xors<S, X / 2>(v);
    }
}
// Merges of blocks of B = 2, 4, … S slots inside each lane.
template <int S, int B> __device__ __forceinline__ void lane_merges(uint32_t (&v)[S]) {
    if constexpr (B <= S) {
        lane_stage<S, B - 1>(v);
        lane_xors<S, B / 4>(v);
        lane_merges<S, B * 2>(v);
    }
}
// Comparators across lanes: the partner lane by CTRL, the same slot or (MIRROR)
// the reversed slot; the lane whose bit is set keeps the larger key.
template <int S, int CTRL, bool MIRROR>
__device__ __forceinline__ void cross_stage(uint32_t (&v)[S], bool upper) {
    uint32_t lo[S], hi[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t src = v[MIRROR ? S - 1 - s : s];
        const uint32_t o = key_from<CTRL>(src);
        lo[s] = min(o, v[s]);
        hi[s] = max(o, v[s]);
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        v[s] = upper ? hi[s] : lo[s];
}
// Ascending sort of the group's 8·S keys.
template <int S> __device__ __forceinline__ void sort_keys8(uint32_t (&v)[S], int gl) {
    lane_merges<S, 2>(v);
    cross_stage<S, 1, true>(v, gl & 1);  // blocks of 2 lanes
    lane_xors<S, S / 2>(v);
    cross_stage<S, 3, true>(v, gl & 2);  // blocks of 4 lanes
    cross_stage<S, 1, false>(v, gl & 1);
    lane_xors<S, S / 2>(v);
    cross_stage<S, 7, true>(v, gl & 4);  // all 8 lanes
    cross_stage<S, 2, false>(v, gl & 2);
    cross_stage<S, 1, false>(v, gl & 1);
    lane_xors<S, S / 2>(v);
}
// Ascending order of a bitonic sequence of the group's 8·S keys.
template <int S> __device__ __forceinline__ void merge_keys8(uint32_t (&v)[S], int gl) {
    cross_stage<S, 4, false>(v, gl & 4);
    cross_stage<S, 2, false>(v, gl & 2);
    cross_stage<S, 1, false>(v, gl & 1);
    lane_xors<S, S / 2>(v);
}
// The key at position e (the same for every group of the wave).
template <int S> __device__ __forceinline__ uint32_t key_at8(const uint32_t (&v)[S], int e, int gl) {
    // Halving by the slot's bits, high to low (selects, not an indexed read:
    // registers cannot be indexed per group).
    const int slot = e & (S - 1);
    uint32_t t[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        t[s] = v[s];
#pragma unroll
    for (int w = S / 2; w >= 1; w /= 2) {
        const bool upper = (slot & w) != 0;
#pragma unroll
        for (int s = 0; s < w; ++s)
            t[s] = upper ? t[s + w] : t[s];
    }
    return group_max_u<8>(gl == e / S ? t[0] : 0u);
}

// Level and noise of fit f by its 8-lane group (every lane gets them).
template <int S>
__device__ __forceinline__ void sort_levels(const GuessArgs &a, uint32_t f, int gl, float &level_out,
                                            float &noise_out) {
    constexpr int G = 8, N = G * S;
    const int m = int(a.m);
    const int c = a.grid->count;  // 1 … N
    const float *y = a.y + size_t(f) * size_t(a.ys_fit);
    const float *right = y + (m - c);
    // Left edge ascending (pads above every key, at the end); right edge
    // descending, held as complemented keys (pads: key 0, at the end). Which
    // slot a value starts in does not matter: lane gl loads values gl, gl + 8, ….
    // Loads at a clamped index (in the row; every load issued at once), the
    // pads chosen after.
    float lv[S], rv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = min(s * G + gl, c - 1);
        lv[s] = y[i];
        rv[s] = right[i];
    }
    uint32_t lk[S], rk[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const bool in = s * G + gl < c;
        lk[s] = in ? float_key(lv[s]) : 0xffffffffu;
        rk[s] = in ? ~float_key(rv[s]) : 0xffffffffu;
    }
    sort_keys8<S>(lk, gl);
    sort_keys8<S>(rk, gl);
    // Each edge's median (initial_guess.hpp:25-30): ascending positions k − 1
    // and k, k = c / 2 (k alone for odd c); the right edge's ascending position
    // t is descending position c − 1 − t.
    const int k = c / 2;
    const float left_hi = key_float(key_at8<S>(lk, k, gl));
    const float right_hi = key_float(~key_at8<S>(rk, c - 1 - k, gl));
    float left_level = left_hi, right_level = right_hi;
    if (!(c & 1)) {
        left_level = std_midpoint(key_float(key_at8<S>(lk, k - 1, gl)), left_hi);
        right_level = std_midpoint(key_float(~key_at8<S>(rk, c - k, gl)), right_hi);
    }
    // Median of both edges: positions c − 1 and c of the 2c values.
    uint32_t below = 0u, above = 0xffffffffu;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t r = ~rk[s];
        below = max(below, min(lk[s], r));
        above = min(above, max(lk[s], r));
    }
    below = group_max_u<G>(below);
    above = group_min_u<G>(above);
    const float level = std_midpoint(key_float(below), key_float(above));
    // Absolute deviations from each edge's own median: along a sorted edge they
    // fall to the median and rise after it, and the pads (above every key)
    // follow, so each edge is a bitonic sequence, ordered by one merge. The
    // right edge is ordered descending (complemented keys), so the two
    // together (left ascending, then right descending) are bitonic again: the
    // half-cleaner min(left[i], right[i]) holds the N smallest, merged once
    // more; the pads are above every deviation, so positions c − 1 and c are
    // the median pair of the 2c deviations.
    uint32_t ld[S], rd[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int e = gl * S + s;
        const float dl = fabsf(key_float(lk[s]) - left_level);
        const float dr = fabsf(key_float(~rk[s]) - right_level);
        ld[s] = e < c ? (__float_as_uint(dl) | 0x80000000u) : 0xffffffffu;
        rd[s] = e < c ? ~(__float_as_uint(dr) | 0x80000000u) : 0u;
    }
    merge_keys8<S>(ld, gl);
    merge_keys8<S>(rd, gl);
    uint32_t top = 0xffffffffu;
    if (c == N) {  // position c lies in the upper half
#pragma unroll
        for (int s = 0; s < S; ++s)
            top = min(top, max(ld[s], ~rd[s]));
        top = group_min_u<G>(top);
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        ld[s] = min(ld[s], ~rd[s]);
    merge_keys8<S>(ld, gl);
    const uint32_t d_lo = key_at8<S>(ld, c - 1, gl);
    const uint32_t d_hi = c == N ? top : key_at8<S>(ld, c, gl);
    const float spread = std_midpoint(key_float(d_lo), key_float(d_hi));
    level_out = level;
    noise_out = 1.4826f * spread;
}

// Stage 2 of the guess, one fit per group of G = 8 or 16 lanes (eight or four
// per wave), streaming the row (a sliding window, so the registers do not grow
// with M). The smoothed-maximum scan is described at the scan. The
// half-height searches of peak() walk outwards from the peak G points per step
// (a few steps: a half width is tens of points), and the centroid and the
// crossings read the few points they need in one batch. Level and noise come
// from stage 1 through the fit's parameter slots, which this stage then
// overwrites.
template <int MODEL, int G>
__device__ __forceinline__ void peak_stage(const GuessArgs &a, uint32_t fit, float level,
                                           float noise) {
    static_assert(MODEL != TWO_LORENTZIAN, "the two-peak guess runs one wave per fit");
    static_assert(G == 8 || G == 16, "groups of 8 or 16 lanes");
    constexpr int W = 4 * G;  // points per step of the scan
    constexpr int P = Params<MODEL>::P;
    const int lane = int(threadIdx.x & 63);
    const int gl = lane & (G - 1);
    const bool live = fit < a.n;
    const uint32_t f = live ? fit : a.n - 1;
    const int m = int(a.m);
    const float *y = a.y + size_t(f) * size_t(a.ys_fit);
    const float *x = a.x;
    const GridInfo grid = *a.grid;
    float out[P];
#pragma unroll
    for (int q = 0; q < P; ++q)
        out[q] = qnan_f();
    float baseline = qnan_f();
    int32_t status = GS_INVALID;
    auto write = [&]() {
        if (live && gl == 0) {
#pragma unroll
            for (int q = 0; q < P; ++q)
                a.params[size_t(fit) * P + q] = out[q];
            if (a.baseline)
                a.baseline[fit] = baseline;
            if (a.gstatus)
                a.gstatus[fit] = status;
        }
    };
    if (m < 5 || !grid.guess_grid_ok) {  // the same for every fit of the launch
        write();
        return;
    }
    // The signal y − level. The single-peak models have slope 0, so
    // level + 0·(x − left_x) is level unless level is −0; then the grid is
    // read (the sum's sign depends on x).
    const bool neg_zero = __float_as_uint(level) == 0x80000000u;
    auto sig = [&](float yj, float xj) {
        return neg_zero ? yj - (level + 0.0f * (xj - grid.left_x)) : yj - level;
    };
    // Strongest smoothed local maximum (initial_guess.hpp:170-182). The first
    // largest smoothed value over 1 ≤ j ≤ M − 2 is that maximum (nothing
    // before it is as large, nothing after it larger), unless it sits at j = 1
    // or M − 2 and the row's end point beats it: then, and only then, the
    // reference's scan is replayed point by point (replay_scan). The fast pass
    // streams the row 4·G points per step, four consecutive points per lane
    // (one 16-byte load when the rows are 16-byte aligned), so a point's
    // neighbours sit in its own lane except at the lane's two ends (DPP).
    // Each lane keeps its largest smoothed value and the first step that
    // reached it; the lanes holding the group's maximum recompute that step's
    // four points to find its first index.
    float best = 0.0f;
    int first = 0x7fffffff;
    int sbad = 0;
    {
        constexpr int CS = 4;  // steps per chunk (loads in flight)
        const int steps = (m + W - 1) / W;
        float lane_best = -__builtin_inff();
        int lane_step = -1;
        float nan_acc = 0.0f;  // Σ s·0: NaN once a signal value is not finite
        // Quarter and half of the four points of the step being smoothed, and
        // the quarter of the last point of the step before it.
        float pq3 = 0.0f, cq[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ch[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        // FULL: the step lies inside the row; VEC: 16-byte rows (one load).
        auto load4 = [&](int k, float(&s)[4], auto full_tag, auto vec_tag, auto nz_tag) {
            constexpr bool FULL = decltype(full_tag)::value, VEC = decltype(vec_tag)::value;
            constexpr bool NZ = decltype(nz_tag)::value;
            const int p = W * k + 4 * gl;
            if constexpr (VEC) {  // M is a multiple of 4: a load is all in or all out
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                const bool in = FULL || p < m;
                if (in)
                    v = *reinterpret_cast<const float4 *>(y + p);
                s[0] = in ? v.x - level : 0.0f;
                s[1] = in ? v.y - level : 0.0f;
                s[2] = in ? v.z - level : 0.0f;
                s[3] = in ? v.w - level : 0.0f;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int j = p + q;
                    const bool in = FULL || j < m;
                    float yj = 0.0f, xj = grid.left_x;
                    if (in) {
                        yj = y[j];
                        if (NZ)
                            xj = x[j];
                    }
                    const float v = NZ ? yj - (level + 0.0f * (xj - grid.left_x)) : yj - level;
                    s[q] = in ? v : 0.0f;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                nan_acc = __builtin_fmaf(s[q], 0.0f, nan_acc);
        };
        // Smooths step kk (nq0: quarter of the next step's first point of this
        // lane) and keeps the lane's largest value over 1 ≤ j ≤ M − 2.
        auto track = [&](int kk, float nq0, auto inner_tag) {
            constexpr bool INNER = decltype(inner_tag)::value;
            const float lq = from_left<G>(pq3, cq[3]);  // quarter of point p − 1
            const float rq = from_right<G>(cq[0], nq0);  // quarter of point p + 4
            float sm[4];
            sm[0] = (lq + ch[0]) + cq[1];
            sm[1] = (cq[0] + ch[1]) + cq[2];
            sm[2] = (cq[1] + ch[2]) + cq[3];
            sm[3] = (cq[2] + ch[3]) + rq;
            if (!INNER) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int j = W * kk + 4 * gl + q;
                    sm[q] = (j >= 1 && j + 1 < m) ? sm[q] : -__builtin_inff();
                }
            }
            const float mx = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
            if (mx > lane_best) {
                lane_best = mx;
                lane_step = kk;
            }
        };
        auto shift_in = [&](const float(&s)[4]) {
            pq3 = cq[3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cq[q] = 0.25f * s[q];
                ch[q] = 0.5f * s[q];
            }
        };
        // Steps k0 … k0 + CS − 1 are loaded and k0 − 1 … k0 + CS − 2 smoothed.
        // INNER: every step loaded lies in the row and every point smoothed
        // in 1 … M − 2 (k0 ≥ 2 and W·(k0 + CS) ≤ M).
        auto chunk4 = [&](int k0, auto inner_tag, auto vec_tag, auto nz_tag) {
            float s[CS][4];
#pragma unroll
            for (int c = 0; c < CS; ++c)
                load4(k0 + c, s[c], inner_tag, vec_tag, nz_tag);
#pragma unroll
            for (int c = 0; c < CS; ++c) {
                if (k0 + c >= 1)
                    track(k0 + c - 1, 0.25f * s[c][0], inner_tag);
                shift_in(s[c]);
            }
        };
        using yes = std::true_type;
        using no = std::false_type;
        const bool vec = !neg_zero && (reinterpret_cast<uintptr_t>(y) & 15u) == 0 && (m & 3) == 0;
        int k0 = 0;
        for (; k0 < steps; k0 += CS) {
            const bool inner = k0 >= 2 && W * (k0 + CS) <= m;
            if (vec)
                inner ? chunk4(k0, yes{}, yes{}, no{}) : chunk4(k0, no{}, yes{}, no{});
            else if (neg_zero)
                chunk4(k0, no{}, no{}, yes{});
            else
                inner ? chunk4(k0, yes{}, no{}, no{}) : chunk4(k0, no{}, no{}, no{});
        }
        track(k0 - 1, 0.0f, no{});  // the last step loaded: nothing beyond it
        sbad = group_or<G>(!finite_f(nan_acc));
        const float top = group_max_f<G>(lane_best);
        uint32_t cand = 0xffffffffu;
        if (lane_step >= 0 && lane_best == top) {
            // This lane's four points of that step, from its six signal values
            // (the same operations as the pass), the first equal to the top.
            const int p = W * lane_step + 4 * gl;
            float sv[6];
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                const int j = p - 1 + d;
                sv[d] = (j >= 0 && j < m) ? sig(y[j], neg_zero ? x[j] : 0.0f) : 0.0f;
            }
#pragma unroll
            for (int q = 3; q >= 0; --q) {
                const float smq = (0.25f * sv[q] + 0.5f * sv[q + 1]) + 0.25f * sv[q + 2];
                const int j = p + q;
                if (j >= 1 && j + 1 < m && smq == top)
                    cand = uint32_t(j);
            }
        }
        cand = group_min_u<G>(cand);
        if (top > 0.0f && cand != 0xffffffffu) {
            best = top;
            first = int(cand);
        }
    }
    // A first maximum at either end of 1 … M − 2 is the reference's only when
    // it beats the row's end point (smooth() returns end points as they are).
    bool replay = false;
    if (!sbad && first == 1)
        replay = !(best > sig(y[0], neg_zero ? x[0] : 0.0f));
    else if (!sbad && first == m - 2)
        replay = !(best >= sig(y[m - 1], neg_zero ? x[m - 1] : 0.0f));
    if (replay) {
        // The reference's scan (initial_guess.hpp:170-182) point by point:
        // 16 points per slot, one per lane; the slot being tested lags the
        // slot being smoothed by one, which lags the newest loaded by one.
        best = 0.0f;
        first = 0x7fffffff;
        const int slots = (m + G - 1) / G;
        float v_prev = 0.0f, v_cur = 0.0f, sm_prev = 0.0f, sm_cur = 0.0f;
        auto slot_step = [&](int k, float vin) {  // vin: slot k's value
            if (k >= 1) {  // smooth slot k - 1
                const float l0 = from_left<G>(v_prev, v_cur), r0 = from_right<G>(v_cur, vin);
                float sm_new = 0.25f * l0 + 0.5f * v_cur + 0.25f * r0;
                const int jn = gl + G * (k - 1);
                sm_new = (jn >= 1 && jn + 1 < m) ? sm_new : v_cur;
                if (k >= 2) {  // test slot k - 2
                    const float l = from_left<G>(sm_prev, sm_cur), r = from_right<G>(sm_cur, sm_new);
                    const int j = gl + G * (k - 2);
                    if (j >= 1 && j + 1 < m && sm_cur > (l > best ? l : best) && sm_cur >= r) {
                        best = sm_cur;
                        first = j;
                    }
                }
                sm_prev = sm_cur;
                sm_cur = sm_new;
            }
            v_prev = v_cur;
            v_cur = vin;
        };
        for (int k = 0; k < slots + 2; ++k) {  // two slots of zeros drain the lag
            const int j = gl + G * k;
            slot_step(k, (k < slots && j < m) ? sig(y[j], neg_zero ? x[j] : 0.0f) : 0.0f);
        }
        group_argmax<G>(best, first);
    }
    int ybad = 0;
    if (sbad) {  // which kind of non-finite value: an observation's or the signal's
        for (int j = gl; j < m; j += G)
            ybad |= !finite_f(y[j]);
        ybad = group_or<G>(ybad);
    }
    if (ybad) {  // non-finite observation: no baseline either (initial_guess.hpp:131-139)
        write();
        return;
    }
    baseline = level;
    if (sbad || !finite_f(noise)) {
        write();
        return;
    }
    status = GS_NO_PEAK;
    if (first == 0x7fffffff || best <= 3.0f * noise) {
        write();
        return;
    }
    // The seven points first − 3 … first + 3 (clamped into the row), read in
    // one batch: the raw maximum and the centroid use no others.
    float yw[7], xw[7], sw[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        const int j = min(max(first - 3 + d, 0), m - 1);
        yw[d] = y[j];
        xw[d] = x[j];
    }
#pragma unroll
    for (int d = 0; d < 7; ++d)
        sw[d] = sig(yw[d], xw[d]);
    // raw_maximum (initial_guess.hpp:111-115): first largest of the three.
    int r = 2;
    float amplitude = sw[2], x_peak = xw[2];
    if (amplitude < sw[3]) {
        r = 3;
        amplitude = sw[3];
        x_peak = xw[3];
    }
    if (amplitude < sw[4]) {
        r = 4;
        amplitude = sw[4];
        x_peak = xw[4];
    }
    const int base = first - 3;  // row index of window slot 0
    first = base + r;
    // peak() (initial_guess.hpp:61-109) over [0, M-1], no mirroring.
    if (!(amplitude > 0.0f)) {
        write();
        return;
    }
    const float half = amplitude / 2.0f;
    // The nearest point at or below half on each side: G points per step
    // outwards from the peak, ending at the first step that holds one (or at
    // the row's end: then the end point, which is above half).
    int lft = -1, rgt = -1;
    const int shift = lane & ~(G - 1);
    constexpr uint32_t kGroupMask = (1u << G) - 1u;
    for (int t = 0;; ++t) {
        const bool want_l = lft < 0, want_r = rgt < 0;
        if (!__ballot(want_l || want_r))
            break;
        const int jl = first - 1 - t * G - gl, jr = first + 1 + t * G + gl;
        bool cl = false, cr = false;
        if (want_l && jl >= 0)
            cl = sig(y[jl], neg_zero ? x[jl] : 0.0f) <= half;
        if (want_r && jr < m)
            cr = sig(y[jr], neg_zero ? x[jr] : 0.0f) <= half;
        const uint32_t ml = uint32_t(__ballot(cl) >> shift) & kGroupMask;
        const uint32_t mr = uint32_t(__ballot(cr) >> shift) & kGroupMask;
        if (want_l) {
            if (ml)
                lft = first - 1 - t * G - __builtin_ctz(ml);
            else if (first - 1 - t * G - (G - 1) <= 0)
                lft = 0;
        }
        if (want_r) {
            if (mr)
                rgt = first + 1 + t * G + __builtin_ctz(mr);
            else if (first + 1 + t * G + (G - 1) >= m - 1)
                rgt = m - 1;
        }
    }
    // The four crossing points, read in one batch (lft < first ≤ M-1, rgt ≥ 1).
    const float y_l0 = y[lft], y_l1 = y[lft + 1], y_r0 = y[rgt - 1], y_r1 = y[rgt];
    const float x_l0 = x[lft], x_l1 = x[lft + 1], x_r0 = x[rgt - 1], x_r1 = x[rgt];
    const float s_l0 = sig(y_l0, x_l0), s_l1 = sig(y_l1, x_l1);
    const float s_r0 = sig(y_r0, x_r0), s_r1 = sig(y_r1, x_r1);
    const bool has_left = s_l0 <= half, has_right = s_r1 <= half;
    if (!has_left || !has_right) {
        write();
        return;
    }
    int low = first > 2 ? first - 2 : 0;
    int high = min(first + 2, m - 1);
    low = max(low, lft + 1);
    high = min(high, rgt - 1);
    float total = 0.0f;
    float center = x_peak;
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        const int j = base + d;
        if (j >= low && j <= high) {
            const float sj = sw[d];
            const float w = (0.0f < sj ? sj : 0.0f) / amplitude;
            if (w > 0.0f) {
                total += w;
                center = std_lerp(center, xw[d], w / total);
            }
        }
    }
    auto crossing = [&](float xp, float sp, float xq, float sq) {
        const float den = sq - sp;
        if (!finite_f(den) || den == 0.0f)
            return qnan_f();
        return std_lerp(xp, xq, (half - sp) / den);
    };
    const float left_radius = center - crossing(x_l0, s_l0, x_l1, s_l1);
    const float right_radius = crossing(x_r0, s_r0, x_r1, s_r1) - center;
    const float width = left_radius + right_radius;
    if (!(finite_f(center) && finite_f(width) && left_radius > 0.0f && right_radius > 0.0f)) {
        write();
        return;
    }
    out[0] = amplitude;
    out[1] = center;
    if constexpr (MODEL == GAUSSIAN)
        out[2] = width / 2.354820045030949382f;
    else if constexpr (MODEL == LORENTZIAN)
        out[2] = width / 2.0f;
    else {
        out[2] = width;
        out[3] = 0.5f;
    }
    status = GS_SUCCESS;
#pragma unroll
    for (int q = 0; q < P; ++q)
        if (!finite_f(out[q]))
            status = GS_INVALID;
    if (!(out[2] > 0.0f))
        status = GS_INVALID;  // the computed values stay, as in the reference
    write();
}

template <int MODEL, int G> __global__ void __launch_bounds__(256) guess_peak_kernel(GuessArgs a) {
    constexpr int P = Params<MODEL>::P;
    const uint32_t fit = blockIdx.x * (256 / G) + threadIdx.x / G;
    const uint32_t f = fit < a.n ? fit : a.n - 1;
    peak_stage<MODEL, G>(a, fit, a.params[size_t(f) * P], a.params[size_t(f) * P + 1]);
}

// Both stages in one launch for edge windows up to 8·S values: a wave sorts
// the edges of its eight fits (VALU work) and streams their rows (memory
// work), so the waves of a SIMD overlap one with the other.
template <int MODEL, int S> __global__ void __launch_bounds__(256) guess_single_kernel(GuessArgs a) {
    const uint32_t fit = blockIdx.x * 32 + threadIdx.x / 8;
    float level, noise;
    sort_levels<S>(a, fit < a.n ? fit : a.n - 1, int(threadIdx.x & 7), level, noise);
    peak_stage<MODEL, 8>(a, fit, level, noise);
}

template <int G, int CE> hipError_t launch_guess_levels(const GuessArgs &a, int stride,
                                                        hipStream_t stream) {
    const uint32_t per_block = 256 / G;
    hipLaunchKernelGGL((guess_levels_kernel<G, CE>), dim3((a.n + per_block - 1) / per_block),
                       dim3(256), 0, stream, a, stride);
    return hipGetLastError();
}

// Whether the two-stage guess serves a launch: single-peak model, fit-major y
// (any M: stage 2 streams the row, stage 1 reads long edges from memory).
__host__ __device__ constexpr bool guess_group_serves(int model, long long ys_pt) {
    return model != TWO_LORENTZIAN && ys_pt == 1;
}

// Launches the two-stage guess when it serves and returns true; false leaves
// the launch to the wave kernel.
template <int MODEL>
bool launch_guess_group(const GuessArgs &a, hipStream_t stream, hipError_t &e) {
    if constexpr (MODEL == TWO_LORENTZIAN) {
        return false;
    } else {
        constexpr int P = Params<MODEL>::P;
        const uint32_t count = a.m / 10 > 1 ? a.m / 10 : 1;
        if (!guess_group_serves(MODEL, a.ys_pt) || a.n == 0)
            return false;
        const uint32_t blocks8 = (a.n + 31) / 32;  // 8 lanes per fit
        if (count <= 128) {  // both stages in one launch
            if (count <= 32)
                hipLaunchKernelGGL((guess_single_kernel<MODEL, 4>), dim3(blocks8), dim3(256), 0,
                                   stream, a);
            else if (count <= 64)
                hipLaunchKernelGGL((guess_single_kernel<MODEL, 8>), dim3(blocks8), dim3(256), 0,
                                   stream, a);
            else
                hipLaunchKernelGGL((guess_single_kernel<MODEL, 16>), dim3(blocks8), dim3(256), 0,
                                   stream, a);
            e = hipGetLastError();
            return true;
        }
        // Windows of 129–230 values: the one-wave kernel (signal in LDS) is
        // faster than the two stages here (0.117 against 0.155 ms per 16,384
        // fits at M = 2048; they cross near M = 2,300).
        if (count <= 230)
            return false;
        if (count <= 448)
            e = launch_guess_levels<64, 7>(a, P, stream);
        else if (count <= 1024)
            e = launch_guess_levels<64, 16>(a, P, stream);
        else {
            hipLaunchKernelGGL(guess_levels_stream_kernel<P>, dim3((a.n + 3) / 4), dim3(256), 0,
                               stream, a);
            e = hipGetLastError();
        }
        if (e != hipSuccess)
            return true;
        hipLaunchKernelGGL((guess_peak_kernel<MODEL, 8>), dim3(blocks8), dim3(256), 0, stream, a);
        e = hipGetLastError();
        return true;
    }
}

}  // namespace bfhip

// The code that follows this header keeps hipcc's default contraction.
#pragma clang fp contract(fast)
