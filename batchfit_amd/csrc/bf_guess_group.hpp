// bf_guess_group.hpp — data-only initial guess of the single-peak models
// (Gaussian, Lorentzian, pseudo-Voigt; cpu/initial_guess.hpp:120-197) in two
// stages with several fits per wave: the edge levels by groups of 16 or 64
// lanes, then the peak by groups of 16 lanes, streaming the row. Equal to the
// FP32 reference guess bit for bit, like the one-wave-per-fit kernel of
// bf_guess.hpp, which still serves the two-peak model, edge windows over 1024
// points and point-major data. No LDS; few registers per lane.
//
// The edge medians are exact order statistics found by bisection with each
// group's own search state in VGPRs, so the groups of a wave search at once
// and the wave runs as many steps as its slowest median.
#pragma once

#include "bf_guess.hpp"

#include <type_traits>

#pragma clang fp contract(off)

namespace bfhip {

template <int G> __device__ __forceinline__ uint32_t group_min_u(uint32_t v) {
    v = min(v, uint32_t(xlane<1>(int(v))));
    v = min(v, uint32_t(xlane<2>(int(v))));
    v = min(v, uint32_t(xlane<4>(int(v))));
    v = min(v, uint32_t(xlane<8>(int(v))));
    if constexpr (G >= 32)
        v = min(v, uint32_t(xlane<16>(int(v))));
    if constexpr (G >= 64)
        v = min(v, uint32_t(xlane<32>(int(v))));
    return v;
}
template <int G> __device__ __forceinline__ uint32_t group_max_u(uint32_t v) {
    v = max(v, uint32_t(xlane<1>(int(v))));
    v = max(v, uint32_t(xlane<2>(int(v))));
    v = max(v, uint32_t(xlane<4>(int(v))));
    v = max(v, uint32_t(xlane<8>(int(v))));
    if constexpr (G >= 32)
        v = max(v, uint32_t(xlane<16>(int(v))));
    if constexpr (G >= 64)
        v = max(v, uint32_t(xlane<32>(int(v))));
    return v;
}
template <int G> __device__ __forceinline__ int group_sum_i(int v) {
    v += xlane<1>(v);
    v += xlane<2>(v);
    v += xlane<4>(v);
    v += xlane<8>(v);
    if constexpr (G >= 32)
        v += xlane<16>(v);
    if constexpr (G >= 64)
        v += xlane<32>(v);
    return v;
}
// Largest value of the group, ties to the smallest index (finite values).
template <int G> __device__ __forceinline__ void group_argmax(float &v, int &idx) {
    auto step = [&](float w, int wi) {
        const bool take = w > v || (w == v && wi < idx);
        v = take ? w : v;
        idx = take ? wi : idx;
    };
    step(__int_as_float(xlane<1>(__float_as_int(v))), xlane<1>(idx));
    step(__int_as_float(xlane<2>(__float_as_int(v))), xlane<2>(idx));
    step(__int_as_float(xlane<4>(__float_as_int(v))), xlane<4>(idx));
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
// 0x134, wave_shr:1 0x138, wave_ror:1 0x13C). A lane whose source lies
// outside its row (or the wave) keeps `old`.
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
// from the group's last lane one slot earlier.
template <int G> __device__ __forceinline__ float from_left(float prev, float cur) {
    if constexpr (G == 16)
        return dpp_f<0x111>(dpp_rot<0x121>(prev), cur);
    else
        return dpp_f<0x138>(dpp_rot<0x13C>(prev), cur);
}
// The value at point j + 1 (cur = slot k, next = slot k + 1).
template <int G> __device__ __forceinline__ float from_right(float cur, float next) {
    if constexpr (G == 16)
        return dpp_f<0x101>(dpp_rot<0x12F>(next), cur);
    else
        return dpp_f<0x130>(dpp_rot<0x134>(next), cur);
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

// Stage 1 for edge windows beyond 1024 values (M ≥ 10,250): one wave per fit,
// the same searches with the window's keys read from memory at every step
// (the two edges stay in the L2 cache), so M has no limit.
__global__ void __launch_bounds__(256) guess_levels_stream_kernel(GuessArgs a, int param_stride) {
    const uint32_t fit = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = int(threadIdx.x & 63);
    const bool live = fit < a.n;
    const uint32_t f = live ? fit : a.n - 1;
    const int m = int(a.m);
    const int count = a.grid->count;
    const float *y = a.y + size_t(f) * size_t(a.ys_fit);
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
    if (live && lane == 0) {
        a.params[size_t(fit) * param_stride] = um.value;
        a.params[size_t(fit) * param_stride + 1] = 1.4826f * spread;
    }
}

// Stage 2 of the guess, one fit per group of 16 lanes (four per wave): lane gl
// takes the points gl, gl + 16, … as it streams the row in chunks of 16·CH
// points (coalesced; a sliding window of slots, so the registers do not grow
// with M). The smoothed-maximum scan takes each point's neighbours from the
// adjacent lanes with DPP (row shifts; the group's edge lanes from the
// neighbouring slot by a row rotate). The half-height searches of peak() walk
// outwards from the peak 16 points per step (a few steps: a half width is
// tens of points), and the centroid and the crossings read the few points
// they need in one batch. Level and noise come from stage 1 through the
// fit's parameter slots, which this stage then overwrites.
template <int MODEL>
__global__ void __launch_bounds__(256) guess_peak_kernel(GuessArgs a) {
    static_assert(MODEL != TWO_LORENTZIAN, "the two-peak guess runs one wave per fit");
    constexpr int G = 16, CH = 8;
    constexpr int P = Params<MODEL>::P;
    const int lane = int(threadIdx.x & 63);
    const int gl = lane & (G - 1);
    const uint32_t fit = blockIdx.x * (256 / G) + threadIdx.x / G;
    const bool live = fit < a.n;
    const uint32_t f = live ? fit : a.n - 1;
    const int m = int(a.m);
    const float *y = a.y + size_t(f) * size_t(a.ys_fit);
    const float *x = a.x;
    const GridInfo grid = *a.grid;
    const float level = a.params[size_t(f) * P];
    const float noise = a.params[size_t(f) * P + 1];
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
    const int slots = (m + G - 1) / G;
    // Strongest smoothed local maximum (initial_guess.hpp:170-182), each lane
    // in increasing index keeping its first best; the slot being tested lags
    // the slot being smoothed by one, which lags the newest loaded by one.
    int ybad = 0, sbad = 0;
    float best = 0.0f;
    int first = 0x7fffffff;
    float v_prev = 0.0f, v_cur = 0.0f, sm_prev = 0.0f, sm_cur = 0.0f;
    // One chunk of CH slots from slot k0. INNER: every point the chunk loads,
    // smooths or tests lies inside the row (1 ≤ j ≤ M − 2), so no lane needs
    // a range test; NZ: the level is −0 (the grid is read).
    auto chunk = [&](int k0, auto inner_tag, auto nz_tag) {
        constexpr bool INNER = decltype(inner_tag)::value, NZ = decltype(nz_tag)::value;
        float vin[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = gl + G * (k0 + c);
            float yj = 0.0f, xj = grid.left_x;
            if (INNER || j < m) {
                yj = y[j];
                if (NZ)
                    xj = x[j];
            }
            const float s = NZ ? yj - (level + 0.0f * (xj - grid.left_x)) : yj - level;
            ybad |= !finite_f(yj);
            sbad |= !finite_f(s);
            vin[c] = (INNER || j < m) ? s : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int k = k0 + c;  // newest loaded slot
            if (k >= 1) {
                // smooth slot k - 1
                const float l0 = from_left<G>(v_prev, v_cur), r0 = from_right<G>(v_cur, vin[c]);
                float sm_new = 0.25f * l0 + 0.5f * v_cur + 0.25f * r0;
                if (!INNER) {
                    const int j = gl + G * (k - 1);
                    sm_new = (j >= 1 && j + 1 < m) ? sm_new : v_cur;
                }
                if (k >= 2) {  // test slot k - 2
                    const float l = from_left<G>(sm_prev, sm_cur), r = from_right<G>(sm_cur, sm_new);
                    bool in = true;
                    if (!INNER) {
                        const int j = gl + G * (k - 2);
                        in = j >= 1 && j + 1 < m;
                    }
                    if (in && sm_cur > (l > best ? l : best) && sm_cur >= r) {
                        best = sm_cur;
                        first = gl + G * (k - 2);
                    }
                }
                sm_prev = sm_cur;
                sm_cur = sm_new;
            }
            v_prev = v_cur;
            v_cur = vin[c];
        }
    };
    using yes = std::true_type;
    using no = std::false_type;
    // A chunk is inner when the slots it smooths and tests (k0 - 2 … k0 + CH
    // - 2) hold only points 1 … M - 2 and the slots it loads lie in the row.
    for (int k0 = 0; k0 < slots; k0 += CH) {
        const bool inner = k0 >= 3 && G * (k0 + CH) <= m - 1;
        if (neg_zero)
            inner ? chunk(k0, yes{}, yes{}) : chunk(k0, no{}, yes{});
        else
            inner ? chunk(k0, yes{}, no{}) : chunk(k0, no{}, no{});
    }
    {
        // Drain: the last slot loaded is smoothed with zeros beyond, and the
        // last two are tested. Slots past the row are zeros and never tested.
        const int k = slots + ((slots % CH) ? CH - slots % CH : 0);  // slots loaded
        const float l0 = from_left<G>(v_prev, v_cur), r0 = from_right<G>(v_cur, 0.0f);
        float sm_new = 0.25f * l0 + 0.5f * v_cur + 0.25f * r0;
        {
            const int j = gl + G * (k - 1);
            sm_new = (j >= 1 && j + 1 < m) ? sm_new : v_cur;
        }
        auto test = [&](int kt, float smp, float smc, float smn) {
            const float l = from_left<G>(smp, smc), r = from_right<G>(smc, smn);
            const int j = gl + G * kt;
            if (j >= 1 && j + 1 < m && smc > (l > best ? l : best) && smc >= r) {
                best = smc;
                first = j;
            }
        };
        test(k - 2, sm_prev, sm_cur, sm_new);
        test(k - 1, sm_cur, sm_new, 0.0f);
    }
    group_argmax<G>(best, first);
    ybad = group_or<G>(ybad);
    sbad = group_or<G>(sbad);
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
    // The nearest point at or below half on each side: 16 points per step
    // outwards from the peak, ending at the first step that holds one (or at
    // the row's end: then the end point, which is above half).
    int lft = -1, rgt = -1;
    const int shift = lane & ~(G - 1);
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
        const uint32_t ml = uint32_t(__ballot(cl) >> shift) & 0xffffu;
        const uint32_t mr = uint32_t(__ballot(cr) >> shift) & 0xffffu;
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
        if (count <= 112)
            e = launch_guess_levels<16, 7>(a, P, stream);
        else if (count <= 448)
            e = launch_guess_levels<64, 7>(a, P, stream);
        else if (count <= 1024)
            e = launch_guess_levels<64, 16>(a, P, stream);
        else {
            hipLaunchKernelGGL(guess_levels_stream_kernel, dim3((a.n + 3) / 4), dim3(256), 0,
                               stream, a, P);
            e = hipGetLastError();
        }
        if (e != hipSuccess)
            return true;
        hipLaunchKernelGGL((guess_peak_kernel<MODEL>), dim3((a.n + 15) / 16), dim3(256), 0,
                           stream, a);
        e = hipGetLastError();
        return true;
    }
}

}  // namespace bfhip

// The code that follows this header keeps hipcc's default contraction.
#pragma clang fp contract(fast)
