// bf_guess.hpp — data-only initial guess on the device, one 64-lane wave per
// fit (cpu/initial_guess.hpp:120-249).
//
// The fit's signal is staged in LDS; the edge medians are exact order
// statistics selected in registers (searched by value over the staged
// signal for windows beyond 1024 values); every scan of the reference (smoothed-maximum search, half-height
// walks, prominence walks, valley search) becomes a wave-wide reduction with
// the reference's tie rule (first index wins). The O(1) arithmetic between
// scans runs redundantly on all lanes with FP contraction disabled, so the
// FP32 result equals the FP32 reference guess bit for bit.
#pragma once

#include "bf_device.hpp"

namespace bfhip {

struct GuessArgs {
    const float *x;
    const float *y;
    long long ys_fit, ys_pt;
    const GridInfo *grid;
    uint32_t n, m;
    uint32_t m_pad;   // LDS floats reserved for the signal
    uint32_t extra_n;  // LDS floats reserved beyond the signal (guess_extra_floats)
    float *params;    // [N,P]
    float *baseline;  // [N] or null
    int32_t *gstatus; // [N] or null
};

// LDS the one-wave guess may take per fit (the signal and, for the two-peak
// model, its step-minimum tables): M up to 16,384 points (15,886 for the
// two-peak model). Longer spectra take the wide guess (bf_guess_wide.hpp),
// which stages nothing.
constexpr uint32_t kGuessLdsBytes = 65536;
// Largest M whose edge windows (M / 10 values each) the sorting-network guess
// of the single-peak models holds: 128 values, eight lanes of 16 slots.
constexpr uint32_t kGuessSortedMaxPoints = 1289;
// Largest M for which the automatic choice takes the one-wave kernel where
// the wide kernel also applies (fit-major rows). Beyond it the wide kernel
// is faster: from 5,130 points both search the edge levels by value, and
// without LDS the wide kernel runs more waves per CU (8,192 fits at M = 8192:
// two-peak 1.9 against 2.6 ms, Lorentzian 0.70 against 1.13 ms; at M = 4096
// the one-wave kernel is ahead, 0.56 against 0.89 and 0.16 against 0.30 ms).
constexpr uint32_t kGuessWaveMaxPoints = 5120;

// LDS floats the guess of `model` needs beyond the signal of M points: for
// the two-peak model, two minima per 64-point step of the second-peak search.
__host__ __device__ inline uint32_t guess_extra_floats(int model, uint32_t m) {
    return model == TWO_LORENTZIAN ? 2 * ((m + 63) / 64) : 0;
}

// Whether the one-wave guess of `model` holds M points in its LDS.
__host__ __device__ inline bool guess_fits_lds(int model, uint32_t m) {
    return (uint64_t(m) + guess_extra_floats(model, m)) * sizeof(float) <= kGuessLdsBytes;
}

// The largest M for which guess_fits_lds(model, M) holds.
__host__ __device__ inline uint32_t guess_lds_max_points(int model) {
    uint32_t m = kGuessLdsBytes / sizeof(float);
    while (m > 0 && !guess_fits_lds(model, m))
        --m;
    return m;
}

// Which guess kernel serves a launch: the fastest that applies (AUTO), or,
// for tests, the one-wave kernel with the signal in LDS (WAVE) or the wide
// kernel that reads memory (WIDE).
enum GuessForm : uint32_t { GUESS_AUTO = 0, GUESS_WAVE = 1, GUESS_WIDE = 2 };

enum : int32_t { GS_SUCCESS = 0, GS_INVALID = 1, GS_NO_PEAK = 2, GS_INSUFFICIENT = 3 };

#pragma clang fp contract(off)

__device__ __forceinline__ float std_midpoint(float a, float b) {
    const float small = 1.17549435082228750797e-38f * 2.0f;
    const float big = 3.40282346638528859812e+38f / 2.0f;
    const float aa = a < 0 ? -a : a, ab = b < 0 ? -b : b;
    if (aa <= big && ab <= big)
        return (a + b) / 2.0f;
    if (aa < small)
        return a + b / 2.0f;
    if (ab < small)
        return a / 2.0f + b;
    return a / 2.0f + b / 2.0f;
}
__device__ __forceinline__ float std_lerp(float a, float b, float t) {
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0))
        return t * b + (1 - t) * a;
    if (t == 1)
        return b;
    const float v = a + t * (b - a);
    return (t > 1) == (b > a) ? (b < v ? v : b) : (b > v ? v : b);
}

// Hand-over of LDS data between the lanes of one wave (a wave owns its LDS
// slice; the LDS unit runs one wave's instructions in order, and the fences
// keep the compiler from moving accesses across). Correct whether the block
// holds one wave or several, each with its own slice.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_or_i(int v) {
    for (int o = 32; o > 0; o >>= 1)
        v |= __shfl_xor(v, o, 64);
    return v;
}
// Wave-wide minimum and maximum of unsigned values as scalars: the
// all-reduce steps of xlane within each 32-lane half, then lanes 0 and 32.
__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) {
    v = min(v, uint32_t(xlane<1>(int(v))));
    v = min(v, uint32_t(xlane<2>(int(v))));
    v = min(v, uint32_t(xlane<4>(int(v))));
    v = min(v, uint32_t(xlane<8>(int(v))));
    v = min(v, uint32_t(xlane<16>(int(v))));
    return min(uint32_t(__builtin_amdgcn_readlane(int(v), 0)),
               uint32_t(__builtin_amdgcn_readlane(int(v), 32)));
}
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
    v = max(v, uint32_t(xlane<1>(int(v))));
    v = max(v, uint32_t(xlane<2>(int(v))));
    v = max(v, uint32_t(xlane<4>(int(v))));
    v = max(v, uint32_t(xlane<8>(int(v))));
    v = max(v, uint32_t(xlane<16>(int(v))));
    return max(uint32_t(__builtin_amdgcn_readlane(int(v), 0)),
               uint32_t(__builtin_amdgcn_readlane(int(v), 32)));
}
// Exact wave minimum of floats with no NaN, as a wave-uniform value: the
// all-reduce steps of xlane within each 32-lane half (DPP within rows, one
// swizzle), then lanes 0 and 32.
__device__ __forceinline__ float wave_min_fx(float v) {
    v = fminf(v, __int_as_float(xlane<1>(__float_as_int(v))));
    v = fminf(v, __int_as_float(xlane<2>(__float_as_int(v))));
    v = fminf(v, __int_as_float(xlane<4>(__float_as_int(v))));
    v = fminf(v, __int_as_float(xlane<8>(__float_as_int(v))));
    v = fminf(v, __int_as_float(xlane<16>(__float_as_int(v))));
    return fminf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)));
}
__device__ __forceinline__ float wave_min_f(float v) {
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = (w < v) ? w : v;
    }
    return v;
}
// Argmax by value, ties to the smallest index; idx == INT_MAX means none.
__device__ __forceinline__ void wave_argmax(float &v, int &idx) {
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        const int wi = __shfl_xor(idx, o, 64);
        if (w > v || (w == v && wi < idx)) {
            v = w;
            idx = wi;
        }
    }
}
__device__ __forceinline__ void wave_argmin(float &v, int &idx) {
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        const int wi = __shfl_xor(idx, o, 64);
        if (w < v || (w == v && wi < idx)) {
            v = w;
            idx = wi;
        }
    }
}

// Order-preserving 32-bit key of a finite float (-0 sorts just below +0) and
// its inverse.
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
// Number of lanes of the wave for which `p` holds (a scalar).
__device__ __forceinline__ int wave_count(bool p) { return __popcll(__ballot(p)); }

// Median of count ≤ 64·SLOTS finite values get(0..count) by exact selection
// (initial_guess.hpp:25-30: the middle order statistic, and for an even count
// the midpoint of it and the largest value below it). The keys stay in
// registers; the k-th smallest key is found bit by bit, each step counting
// the keys at or below a bound with one ballot per slot. The search keeps
// the interval [t, t + 2^b) that holds the k-th key with the count of keys
// below it (lo) and an upper bound on the count at or below its top (hi,
// exact once the search has narrowed from above). It starts below the bits
// that the smallest and largest keys share, and once the interval holds a
// single key, that key is the answer and the remaining bits are read from it
// instead of searched.
template <int SLOTS, typename F>
__device__ float select_median(int count, int lane, F &&get) {
    uint32_t key[SLOTS];
    uint32_t kmin = 0xffffffffu, kmax = 0;
#pragma unroll
    for (int q = 0; q < SLOTS; ++q) {
        const int e = q * 64 + lane;
        key[q] = e < count ? float_key(get(e)) : 0xffffffffu;  // above every finite key
        kmin = min(kmin, key[q]);
        kmax = max(kmax, e < count ? key[q] : 0u);
    }
    kmin = wave_min_u(kmin);
    kmax = wave_max_u(kmax);
    const uint32_t diff = kmin ^ kmax;
    const int top = diff ? 31 - __builtin_clz(diff) : -1;  // highest bit that differs
    const int k = count / 2;
    uint32_t t = top >= 0 ? kmin & ~((2u << top) - 1u) : kmin;  // 2u << 31 is 0
    int lo = 0, hi = 64 * SLOTS;  // lo ≤ k < exact count at or below the top
    for (int b = top; b >= 0; --b) {
        const uint32_t span = (1u << b) - 1u;
        int c = 0;
#pragma unroll
        for (int q = 0; q < SLOTS; ++q)
            c += wave_count(key[q] <= (t | span));
        if (c < k + 1) {
            t |= 1u << b;
            lo = c;
        } else {
            hi = c;
        }
        if (hi - lo == 1) {  // one key in [t, t + span]: lo = k, so it is the k-th
            uint32_t found = t;
#pragma unroll
            for (int q = 0; q < SLOTS; ++q) {
                const uint64_t in = __ballot(key[q] - t <= span);
                if (in)
                    found = uint32_t(__builtin_amdgcn_readlane(int(key[q]), __builtin_ctzll(in)));
            }
            t = found;
            break;
        }
    }
    if (count % 2)
        return key_float(t);
    // lo counts the keys below t whether the search ended early or not.
    uint32_t under = 0;
#pragma unroll
    for (int q = 0; q < SLOTS; ++q)
        under = (key[q] < t && key[q] > under) ? key[q] : under;
    uint32_t lower = t;
    if (lo >= k)  // the k smallest are all below t: the largest of them
        lower = wave_max_u(under);
    return std_midpoint(key_float(lower), key_float(t));
}

// Median of get(0..count), count ≤ 1024 (initial_guess.hpp:25-30): exact
// selection in registers.
template <typename F> __device__ float wave_median(int count, int lane, F &&get) {
    if (count <= 64)
        return select_median<1>(count, lane, get);
    if (count <= 128)
        return select_median<2>(count, lane, get);
    if (count <= 256)
        return select_median<4>(count, lane, get);
    if (count <= 512)
        return select_median<8>(count, lane, get);
    return select_median<16>(count, lane, get);
}

// The four edge levels of a row (initial_guess.hpp:142-165): the median of
// each edge window, of both, and of the absolute deviations from each edge's
// own median. stream_levels (bf_guess_group.hpp) reads them from y[0..m) by
// one whole wave at every step of its searches, for windows of any length.
struct EdgeLevels {
    float left, right, both, spread;
};
__device__ EdgeLevels stream_levels(const float *y, int m, int count, int lane);

struct GuessPeak {
    float amplitude, center, width, left_radius, right_radius;
    bool valid;
};

template <typename S> __device__ __forceinline__ float smooth_at(const S &s, int m, int i) {
    if (i == 0 || i + 1 == m)
        return s[i];
    return 0.25f * s[i - 1] + 0.5f * s[i] + 0.25f * s[i + 1];
}

// Smoothed value at 1 ≤ j < m as smooth_at (the raw value at the last
// point); j ≥ m gives 0, a value no caller uses.
template <typename S> __device__ __forceinline__ float smooth_in(const S &s, int m, int j) {
    if (j + 1 < m)
        return 0.25f * s[j - 1] + 0.5f * s[j] + 0.25f * s[j + 1];
    return j < m ? s[j] : 0.0f;
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
// The value of lane − 1, with `first` in lane 0 (DPP wave_shr:1), and of
// lane + 1, with `last` in lane 63 (DPP wave_shl:1). All 64 lanes active.
__device__ __forceinline__ float from_lane_below(float v, float first) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, first),
                                                                 __builtin_bit_cast(int, v), 0x138,
                                                                 0xF, 0xF, false));
}
__device__ __forceinline__ float from_lane_above(float v, float last) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, last),
                                                                 __builtin_bit_cast(int, v), 0x130,
                                                                 0xF, 0xF, false));
}

// initial_guess.hpp:61-109 with both half-height walks as wave reductions.
template <typename S>
__device__ GuessPeak measure_peak_dev(const float *x, const S &s, int m, int index, int begin,
                                      int end, bool mirror, int lane) {
    GuessPeak r{0, 0, 0, 0, 0, false};
    r.amplitude = s[index];
    if (!(r.amplitude > 0.0f))
        return r;
    const float half = r.amplitude / 2.0f;
    // The nearest point at or below half on each side: 64-point steps
    // outwards from the peak, ending at the first step that holds one.
    int left_c = -1, right_c = 0x7fffffff;
    for (int top = index; top > begin; top -= 64) {  // the step [top - 64, top)
        const int j = top - 64 + lane;
        const uint64_t below = __ballot(j >= begin && s[j] <= half);
        if (below) {
            left_c = top - 1 - __builtin_clzll(below);
            break;
        }
    }
    for (int bottom = index + 1; bottom <= end; bottom += 64) {  // [bottom, bottom + 64)
        const int j = bottom + lane;
        const uint64_t below = __ballot(j <= end && s[j] <= half);
        if (below) {
            right_c = bottom + __builtin_ctzll(below);
            break;
        }
    }
    const int left = left_c >= 0 ? left_c : begin;
    const int right = right_c != 0x7fffffff ? right_c : end;
    const bool has_left = s[left] <= half, has_right = s[right] <= half;
    if ((!has_left && (!mirror || begin == 0)) || (!has_right && (!mirror || end + 1 == m)) ||
        (!has_left && !has_right))
        return r;
    int low = index > 2 ? index - 2 : 0;
    int high = min(index + 2, end);
    low = max(low, has_left ? left + 1 : begin);
    high = min(high, has_right ? right - 1 : end);
    float total = 0.0f;
    r.center = x[index];
    for (int j = low; j <= high; ++j) {
        const float sj = s[j];
        const float w = (0.0f < sj ? sj : 0.0f) / r.amplitude;
        if (w > 0.0f) {
            total += w;
            r.center = std_lerp(r.center, x[j], w / total);
        }
    }
    auto crossing = [&](int a, int b) {
        const float den = s[b] - s[a];
        if (!finite_f(den) || den == 0.0f)
            return qnan_f();
        return std_lerp(x[a], x[b], (half - s[a]) / den);
    };
    if (has_left)
        r.left_radius = r.center - crossing(left, left + 1);
    if (has_right)
        r.right_radius = crossing(right - 1, right) - r.center;
    if (!has_left)
        r.left_radius = r.right_radius;
    if (!has_right)
        r.right_radius = r.left_radius;
    r.width = r.left_radius + r.right_radius;
    r.valid = finite_f(r.center) && finite_f(r.width) && r.left_radius > 0.0f &&
              r.right_radius > 0.0f;
    return r;
}

template <typename S> __device__ __forceinline__ int raw_peak_dev(const S &s, int i) {
    int best = i - 1;
    for (int j = i; j <= i + 1; ++j)
        if (s[best] < s[j])
            best = j;
    return best;
}

// Length of the prefix of [0, m) on which pred holds (pred true up to some
// index and false from there on), found by the whole wave from index 0 ≤
// from < m outwards, 64 indices per step: the boundary lies near `from`.
template <typename Pred>
__device__ __forceinline__ int prefix_length(int m, int from, Pred pred, int lane) {
    if (pred(from)) {
        for (int b = from + 1; b < m; b += 64) {
            const int i = b + lane;
            const uint64_t no = __ballot(i < m && !pred(i));
            if (no)
                return b + __builtin_ctzll(no);
        }
        return m;
    }
    for (int b = from - 1; b >= 0; b -= 64) {
        const int i = b - lane;
        const uint64_t yes = __ballot(i >= 0 && pred(i));
        if (yes)
            return b - __builtin_ctzll(yes) + 1;
    }
    return 0;
}

// The raw values at j − 1, j, j + 1 where smooth_in reads them (zeros
// elsewhere) and the smoothed value v = smooth_in(s, m, j).
struct Raw3 {
    float a, b, c, v;
};
template <typename S> __device__ __forceinline__ Raw3 raw3_in(const S &s, int m, int j) {
    Raw3 r{0.0f, 0.0f, 0.0f, 0.0f};
    if (j + 1 < m) {
        r.a = s[j - 1];
        r.b = s[j];
        r.c = s[j + 1];
        r.v = 0.25f * r.a + 0.5f * r.b + 0.25f * r.c;
    } else if (j < m) {
        r.b = s[j];
        r.v = r.b;
    }
    return r;
}
// raw_peak_dev(s, j) from the raw values of raw3_in(s, m, j), j + 1 < m.
__device__ __forceinline__ int raw_peak3(const Raw3 &r, int j) {
    int best = j - 1;
    float bv = r.a;
    if (bv < r.b) {
        best = j;
        bv = r.b;
    }
    if (bv < r.c)
        best = j + 1;
    return best;
}

// Whether the smoothed maximum at j of height h has prominence > t
// (initial_guess.hpp:38-54 and the `<= 3·noise` test of :212), decided by
// the whole wave for one candidate (j, h, t wave-uniform). The reference
// walks each side to the first higher point, takes the side minima L and R
// and compares h − max(L, R) with t. Rounding of h − v is monotone in v, so
// the test holds iff h − L > t and h − R > t, and h − L > t iff some point
// of the walk has h − v > t: a side passes iff, going outwards from j, a
// point with h − v > t comes before the first point above h. The wave looks
// at 64 points of a side per step (lane i at distance 1 + i from the step's
// start); a step that holds such a point decides the side by the nearer of
// the two kinds. t ≥ 0, so no point is of both kinds, and the points beyond
// an edge (taken as v = h) are of neither.
template <typename S>
__device__ bool prominent_wave(const S &s, int m, int j, float h, float t, int lane) {
    bool left = false;
    for (int k0 = j - 1; k0 >= 0; k0 -= 64) {
        const int k = k0 - lane;
        const float v = k >= 0 ? smooth_at(s, m, k) : h;
        const uint64_t higher = __ballot(v > h), deep = __ballot(h - v > t);
        const uint64_t ends = higher | deep;
        if (ends) {
            left = (deep & (ends & (~ends + 1))) != 0;
            break;
        }
    }
    if (!left)
        return false;
    for (int k0 = j + 1; k0 < m; k0 += 64) {
        const int k = k0 + lane;
        const float v = k < m ? smooth_at(s, m, k) : h;
        const uint64_t higher = __ballot(v > h), deep = __ballot(h - v > t);
        const uint64_t ends = higher | deep;
        if (ends)
            return (deep & (ends & (~ends + 1))) != 0;
    }
    return false;
}

// The prominence test of prominent_wave for the candidate at 1 ≤ j < m − 1
// of height h, by its own lane, over the nearest kLocalWalk points of both
// sides (the raw values they need are loaded together: one memory round trip
// instead of one per point): 1 if each side reaches a point with h − v > t
// before a point above h or the edge (the candidate is prominent), 2 if a
// side reaches a point above h or the edge first (it is not), 0 if that is
// not settled within reach.
constexpr int kLocalWalk = 8;
template <typename S>
__device__ __forceinline__ int local_sides(const S &s, int m, int j, float h, float t) {
    constexpr int W = kLocalWalk + 2;
    // wl[e] = s[j − e] and wr[e] = s[j + e]: the raw values around the points
    // j − 1 … j − 8 and j + 1 … j + 8. Indices beyond an edge are clamped;
    // such values are never used, because the side is settled at the edge.
    float wl[W], wr[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        wl[e] = s[max(j - e, 0)];
        wr[e] = s[min(j + e, m - 1)];
    }
    // Points in walk order; the first that settles a side decides it. Left
    // points lie below j ≤ m − 2 and right ones above j ≥ 1, so only k = 0
    // (left) and k = m − 1 (right) take the raw value.
    int ls = 0, rs = 0;  // per side: 0 open, 1 passes, 2 fails
#pragma unroll
    for (int d = 1; d <= kLocalWalk; ++d) {
        const int kl = j - d, kr = j + d;
        const float vl = kl == 0 ? wl[d] : 0.25f * wl[d + 1] + 0.5f * wl[d] + 0.25f * wl[d - 1];
        const float vr = kr == m - 1 ? wr[d] : 0.25f * wr[d - 1] + 0.5f * wr[d] + 0.25f * wr[d + 1];
        const int sl = (kl < 0 || vl > h) ? 2 : (h - vl > t) ? 1 : 0;
        const int sr = (kr >= m || vr > h) ? 2 : (h - vr > t) ? 1 : 0;
        ls = ls ? ls : sl;
        rs = rs ? rs : sr;
    }
    return (ls == 2 || rs == 2) ? 2 : (ls == 1 && rs == 1) ? 1 : 0;
}

// In-place suffix minima of t[0..n) by one wave: t[e] = min(t[e..n)).
__device__ __forceinline__ void wave_suffix_min(float *t, int n, int lane) {
    wave_sync();
    float carry = __uint_as_float(0x7f800000u);
    for (int b = (n - 1) & ~63; b >= 0; b -= 64) {
        const int e = b + lane;
        float v = e < n ? t[e] : __uint_as_float(0x7f800000u);
        for (int d = 1; d < 64; d <<= 1)
            v = fminf(v, __shfl_down(v, d, 64));
        v = fminf(v, carry);
        if (e < n)
            t[e] = v;
        carry = lane_value(v, 0);
    }
    wave_sync();
}

// Result of one fit's guess, identical in every lane of the wave.
struct GuessResult {
    float p[8];
    float baseline;
    int32_t status;
};

// The guess from its edge levels on (initial_guess.hpp:170-249), by one whole
// wave: s[j] is the baseline-removed signal (finite at every point), sfx has
// two floats per run of 2^sfx_shift 64-point steps (the two-peak model; runs
// = ((m − 2) / 64 >> sfx_shift) + 1 runs), res holds the baseline and
// GS_INVALID on entry.
template <int MODEL, typename S>
__device__ GuessResult guess_tail(const S &s, float *sfx, int sfx_shift, const float *x, int m,
                                  const GridInfo &grid, float left_level, float slope, float noise,
                                  int lane, GuessResult res) {
    constexpr int P = Params<MODEL>::P;
    constexpr bool paired = MODEL == TWO_LORENTZIAN;
    float *const out = res.p;
    int32_t &gstatus = res.status;
    auto finish = [&]() { return res; };
    const float left_x = grid.left_x;
    // Strongest smoothed local maximum (initial_guess.hpp:170-182). Each
    // smoothed value is formed once, in the lane of its point; a point's
    // neighbours come from the adjacent lanes, and across a step's edges
    // from the previous and next steps.
    // The two-peak model's scan also records, per run of 2^sfx_shift steps
    // (the last step holds the point m − 1 alone when (m − 2) % 64 = 0), the
    // minimum of the smoothed signal within the run (sfx[run], made suffix
    // minima below) and from the start through the run (pfx[run]).
    const int runs = paired ? ((m - 2) / 64 >> sfx_shift) + 1 : 0;
    float *const pfx = sfx + runs;
    float best = 0.0f;
    int first = 0x7fffffff;
    {
        const float inf = __uint_as_float(0x7f800000u);
        const int run_mask = (1 << sfx_shift) - 1;
        float cur = smooth_in(s, m, 1 + lane);
        float before = s[0];  // the smoothed value at 0 is the raw one
        float prefix = s[0], run = inf;
        for (int base = 1; base < (paired ? m : m - 1); base += 64) {
            const int j = base + lane;
            const float next = smooth_in(s, m, j + 64);
            const float left = from_lane_below(cur, before);
            const float right = from_lane_above(cur, lane_value(next, 0));
            if (j + 1 < m && cur > best && cur > left && cur >= right) {
                best = cur;
                first = j;
            }
            if constexpr (paired) {
                const float step_min = wave_min_fx(j < m ? cur : inf);
                prefix = fminf(prefix, step_min);
                run = fminf(run, step_min);
                const int c = (base - 1) / 64;
                if (((c + 1) & run_mask) == 0 || base + 64 >= m) {
                    if (lane == 0) {
                        sfx[c >> sfx_shift] = run;
                        pfx[c >> sfx_shift] = prefix;
                    }
                    run = inf;
                }
            }
            before = lane_value(cur, 63);
            cur = next;
        }
    }
    wave_argmax(best, first);
    gstatus = GS_NO_PEAK;
    if (first == 0x7fffffff || best <= 3.0f * noise)
        return finish();
    first = raw_peak_dev(s, first);
    GuessPeak primary = measure_peak_dev(x, s, m, first, 0, m - 1, false, lane);
    if (!primary.valid)
        return finish();
    if constexpr (!paired) {
        out[0] = primary.amplitude;
        out[1] = primary.center;
        if constexpr (MODEL == GAUSSIAN)
            out[2] = primary.width / 2.354820045030949382f;
        else if constexpr (MODEL == LORENTZIAN)
            out[2] = primary.width / 2.0f;
        else {
            out[2] = primary.width;
            out[3] = 0.5f;
        }
    } else {
        gstatus = GS_INSUFFICIENT;
        const float lr = primary.left_radius, rr = primary.right_radius;
        const float a3 = 3.0f * grid.guess_spacing, a2 = 2.0f * ((rr < lr) ? rr : lr);
        const float separation = (a3 < a2) ? a2 : a3;
        // The prominent local maximum of largest smoothed value, earliest on
        // ties (initial_guess.hpp:203-217): candidates are taken in
        // increasing index, so a later one must be strictly larger to win.
        const float t = 3.0f * noise;
        // x is strictly increasing (guess_grid_ok), so x[i] − c is
        // non-decreasing in i and |x[i] − c| < separation holds exactly on
        // the indices [sep_lo, sep_hi): the prefixes of the points with
        // x[i] − c ≤ −separation and with x[i] − c < separation (a NaN
        // centre or separation gives sep_lo = m: no index).
        const float c = primary.center;
        const int sep_lo =
            prefix_length(m, first, [&](int i) { return !(x[i] - c > -separation); }, lane);
        const int sep_hi = prefix_length(m, first, [&](int i) { return x[i] - c < separation; }, lane);
        // A side can pass only if it holds a point with h − v > t, so only if
        // h − (minimum over any range that contains the side) > t: the minima
        // from the start of a candidate's run to the end (sfx[run]) and from
        // the start through its run (pfx[run]) reject most candidates
        // without a walk.
        wave_suffix_min(sfx, runs, lane);
        float v_best = 0.0f;
        int j_best = 0x7fffffff;
        // The queue: lane q < queued holds the candidate at qj of height qh.
        int queued = 0, qj = 0;
        float qh = 0.0f;
        auto drain_queue = [&]() {
            const bool mine = lane < queued;
            const int state = mine ? local_sides(s, m, qj, qh, t) : 0;
            const uint64_t settled_yes = __ballot(mine && state == 1);
            const uint64_t settled_no = __ballot(mine && state == 2);
            const uint64_t all = queued == 64 ? ~0ull : ((1ull << queued) - 1);
            for (uint64_t mask = all & ~settled_no; mask; mask &= mask - 1) {
                const int l = __builtin_ctzll(mask);
                const float h = __builtin_bit_cast(
                    float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qh), l));
                if (!(h > v_best))
                    continue;
                const int jc = __builtin_amdgcn_readlane(qj, l);
                if (((settled_yes >> l) & 1) || prominent_wave(s, m, jc, h, t, lane)) {
                    v_best = h;
                    j_best = jc;
                }
            }
            queued = 0;
        };
        Raw3 cur = raw3_in(s, m, 1 + lane);
        float before = s[0];
        for (int base = 1; base + 1 < m; base += 64) {
            const int j = base + lane;
            const Raw3 here = cur;
            const float v = here.v;  // the smoothed value wherever j + 1 < m
            cur = raw3_in(s, m, j + 64);
            const float left = from_lane_below(v, before);
            const float right = from_lane_above(v, lane_value(cur.v, 0));
            before = lane_value(v, 63);
            bool cand = false;
            const int r = ((base - 1) / 64) >> sfx_shift;
            const float left_min = pfx[r], right_min = sfx[r];
            if (j + 1 < m) {
                cand = v > v_best && v - left_min > t && v - right_min > t && v > left &&
                       v >= right;
                if (cand) {
                    const int idx = raw_peak3(here, j);
                    const int bins = idx > first ? idx - first : first - idx;
                    cand = !(bins < 3 || (idx >= sep_lo && idx < sep_hi));
                }
            }
            // Candidates join a queue, one per lane in index order, which is
            // tested when full and at the end: each queued candidate first
            // reads the nearest kLocalWalk points of both sides on its own
            // lane (most are settled there), the rest are walked by the
            // whole wave. Queued candidates passed `v > v_best` against the
            // best when they joined; they are taken in index order and must
            // beat the best at their turn, so the outcome is that of testing
            // each step's candidates at once, while the per-lane tests run
            // with every lane busy instead of a few.
            const uint64_t cm = __ballot(cand);
            if (cm) {
                const int nc = __popcll(cm);
                if (queued + nc > 64) {
                    drain_queue();
                }
                // One forward permute moves the step's candidates, in lane
                // order, to lanes queued … queued + nc − 1 (the other lanes
                // fill the remaining places, so every lane has one sender).
                const int k = lane - queued;
                const int below = __builtin_amdgcn_mbcnt_hi(
                    uint32_t(cm >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(cm), 0));
                const int dest = cand ? queued + below : (queued + nc + lane - below) & 63;
                const int sj = __builtin_amdgcn_ds_permute(dest * 4, j);
                const float sh =
                    __int_as_float(__builtin_amdgcn_ds_permute(dest * 4, __float_as_int(v)));
                if (k >= 0 && k < nc) {
                    qj = sj;
                    qh = sh;
                }
                queued += nc;
            }
        }
        drain_queue();
        int second = j_best == 0x7fffffff ? -1 : raw_peak_dev(s, j_best);
        if (second < 0)
            return finish();
        if (first > second) {
            const int t = first;
            first = second;
            second = t;
        }
        float vmin = __uint_as_float(0x7f800000u);
        int valley = 0x7fffffff;
        for (int j = first + 1 + lane; j < second; j += 64)
            if (s[j] < vmin || (s[j] == vmin && j < valley)) {
                vmin = s[j];
                valley = j;
            }
        wave_argmin(vmin, valley);
        if (valley == 0x7fffffff)
            valley = second;
        primary = measure_peak_dev(x, s, m, first, 0, valley, true, lane);
        const GuessPeak secondary = measure_peak_dev(x, s, m, second, valley, m - 1, true, lane);
        if (!primary.valid || !secondary.valid || !(primary.center < secondary.center))
            return finish();
        out[0] = primary.amplitude;
        out[1] = primary.center;
        out[2] = primary.width / 2.0f;
        out[3] = secondary.amplitude;
        out[4] = secondary.center;
        out[5] = secondary.width / 2.0f;
        out[6] = left_level - slope * left_x;
        out[7] = slope;
    }
    for (int k = 0; k < P; ++k)
        if (!finite_f(out[k])) {
            gstatus = GS_INVALID;
            return finish();
        }
    if (!(out[2] > 0.0f) || (paired && !(out[5] > 0.0f))) {
        gstatus = GS_INVALID;
        return finish();
    }
    gstatus = GS_SUCCESS;
    return finish();
}

// The guess of one fit by one whole wave (lane = lane in the wave), on the
// wave's own LDS: s[0..m) holds the fit's raw observations on entry (it is
// turned into the baseline-removed signal), buf has the floats the host
// reserved beyond it (guess_extra_floats: the two-peak model's step-minimum
// tables), y_bad says whether any observation is non-finite
// (initial_guess.hpp:131-139).
template <int MODEL>
__device__ GuessResult guess_core(float *s, float *buf, const float *x, int m, GridInfo grid,
                                  bool y_bad, int lane) {
    constexpr bool paired = MODEL == TWO_LORENTZIAN;
    GuessResult res;
    for (int k = 0; k < 8; ++k)
        res.p[k] = qnan_f();
    float &baseline = res.baseline;
    baseline = qnan_f();
    res.status = GS_INVALID;
    auto finish = [&]() { return res; };
    if (m < 5 || !grid.guess_grid_ok || y_bad)
        return finish();
    const int count = grid.count;
    // Edge levels, baseline and noise scale (initial_guess.hpp:142-169):
    // selected in registers while both windows together hold at most 1024
    // values (M < 5,130), else searched by value over the staged signal.
    float left_level, right_level, spread;
    if (2 * count <= 1024) {
        left_level = wave_median(count, lane, [&](int i) { return s[i]; });
        right_level = wave_median(count, lane, [&](int i) { return s[m - count + i]; });
        baseline = wave_median(2 * count, lane, [&](int i) {
            return i < count ? s[i] : s[m - count + (i - count)];
        });
        spread = wave_median(2 * count, lane, [&](int i) {
            return i < count ? fabsf(s[i] - left_level)
                             : fabsf(s[m - count + (i - count)] - right_level);
        });
    } else {
        const EdgeLevels lv = stream_levels(s, m, count, lane);
        left_level = lv.left;
        right_level = lv.right;
        baseline = lv.both;
        spread = lv.spread;
    }
    const float left_x = grid.left_x, right_x = grid.right_x;
    const float slope = paired ? (right_level - left_level) / (right_x - left_x) : 0.0f;
    const float level = paired ? left_level : baseline;
    if (!finite_f(slope))
        return finish();
    const float noise = 1.4826f * spread;
    int bad = 0;
    if (paired || __float_as_uint(level) == 0x80000000u) {
#pragma unroll 8
        for (int j = lane; j < m; j += 64) {
            const float v = s[j] - (level + slope * (x[j] - left_x));
            bad |= !finite_f(v);
            s[j] = v;
        }
    } else {
        // slope is +0 and x finite, so level + slope·(x − left_x) is level
        // itself unless level is −0: the grid is not read.
#pragma unroll 8
        for (int j = lane; j < m; j += 64) {
            const float v = s[j] - level;
            bad |= !finite_f(v);
            s[j] = v;
        }
    }
    bad = wave_or_i(bad);
    wave_sync();
    if (bad || !finite_f(noise))
        return finish();
    return guess_tail<MODEL>(static_cast<const float *>(s), buf, 0, x, m, grid, left_level, slope,
                             noise, lane, res);
}

// The guess of fit `fit` of `a`, run by one whole wave that owns the dynamic
// LDS of its block (a.m_pad + a.extra_n floats).
template <int MODEL> __device__ void guess_fit(const GuessArgs &a, uint32_t fit) {
    constexpr int P = Params<MODEL>::P;
    extern __shared__ float lds[];
    const int lane = int(threadIdx.x & 63);
    const int m = int(a.m);
    const float *y = a.y + fit * a.ys_fit;
    int bad = 0;
    // 16 loads per lane in flight at a time, then the remainder one by one.
    constexpr int U = 16;
    int j = lane;
    for (; j + 64 * (U - 1) < m; j += 64 * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[u] = y[(j + 64 * u) * a.ys_pt];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            bad |= !finite_f(v[u]);
            lds[j + 64 * u] = v[u];
        }
    }
#pragma unroll 4
    for (; j < m; j += 64) {
        const float v = y[j * a.ys_pt];
        bad |= !finite_f(v);
        lds[j] = v;
    }
    bad = wave_or_i(bad);
    wave_sync();
    const GuessResult g = guess_core<MODEL>(lds, lds + a.m_pad, a.x, m, *a.grid, bad != 0, lane);
    if (lane == 0) {
        for (int k = 0; k < P; ++k)
            a.params[size_t(fit) * P + k] = g.p[k];
        if (a.baseline)
            a.baseline[fit] = g.baseline;
        if (a.gstatus)
            a.gstatus[fit] = g.status;
    }
}

template <int MODEL> __global__ void __launch_bounds__(64) guess_kernel(GuessArgs a) {
    if (blockIdx.x < a.n)
        guess_fit<MODEL>(a, blockIdx.x);
}

// The code that follows this header keeps hipcc's default contraction.
#pragma clang fp contract(fast)

}  // namespace bfhip

// stream_levels (declared above, used by guess_core) is defined there.
#include "bf_guess_group.hpp"
