// bf_device.hpp — CDNA4 device helpers for the batched LM fitter: finiteness
// tests on bit patterns, the four peak models with fused value + analytic
// Jacobian in FP32, and cross-lane sums for groups of 4/8/16/32/64 lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bfhip {

enum ModelId : int { GAUSSIAN = 0, LORENTZIAN = 1, PSEUDO_VOIGT = 2, TWO_LORENTZIAN = 3 };
template <int MODEL> struct Params {
    static constexpr int P = MODEL == PSEUDO_VOIGT ? 4 : MODEL == TWO_LORENTZIAN ? 8 : 3;
    static constexpr int TRI = P * (P + 1) / 2;  // packed lower triangle of JᵀJ
    static constexpr int TERMS = TRI + P + 1;    // JᵀJ, Jᵀr, Σr²
};

// Status codes: cpu/solver.hpp:16-23 and uncertainty_types.hpp:3-10.
enum : int32_t {
    ST_CONVERGED = 0,
    ST_MAX_ITERATIONS = 1,
    ST_SINGULAR = 2,
    ST_INVALID = 3,
    ST_DEGENERATE = 4,
    ST_UNRESOLVED = 5
};
enum : int32_t {
    US_AVAILABLE = 0,
    US_NOT_CONVERGED = 1,
    US_DOF = 2,
    US_SINGULAR = 3,
    US_ACTIVE_BOUND = 4,
    US_INVALID = 5
};

// Finiteness from the exponent bits: immune to any fast-math assumption.
__device__ __forceinline__ bool finite_f(float v) {
    return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u;
}
__device__ __forceinline__ bool nan_f(float v) {
    return (__float_as_uint(v) & 0x7fffffffu) > 0x7f800000u;
}
__device__ __forceinline__ bool finite_d(double v) {
    return (uint32_t(__double_as_longlong(v) >> 32) & 0x7ff00000u) != 0x7ff00000u;
}
__device__ __forceinline__ float qnan_f() { return __uint_as_float(0x7fc00000u); }

// ---------------------------------------------------------------------------
// Peak models (models.hpp:78-185), FP32, value and Jacobian from one set of
// transcendentals. Per-pass constants (reciprocal width, mixing complement …)
// are hoisted into ModelConsts once per parameter vector.
//   Gaussian     exp(-z²/2) = 2^(-z²·log2(e)/2)
//   pseudo-Voigt exp(-4 ln2 z²) = 2^(-4 z²)   (exact rewrite of the exponent)
// Contraction only within one source expression (a·b + c written together):
// the code a model inlines into must not decide which products fuse, or the
// same fit would round differently in the tile and the streamed passes.
#pragma clang fp contract(on)

template <int MODEL> struct ModelConsts;

// Factor that turns the eval() Jacobian entry k into the model's derivative
// (1 unless a model evaluates a column at an exact power-of-two scale).
template <int MODEL> __host__ __device__ constexpr float jac_scale(int k) {
    if constexpr (MODEL == PSEUDO_VOIGT)
        return k == 2 ? 0.5f : 1.0f;
    else
        return 1.0f;
}

template <> struct ModelConsts<GAUSSIAN> {
    float a, c, iw, aiw;
    __device__ __forceinline__ void set(const float *p) {
        a = p[0];
        c = p[1];
        iw = 1.0f / p[2];
        aiw = a * iw;
    }
    // returns f; jac[0..2]
    __device__ __forceinline__ float eval(float x, float *jac) const {
        const float z = (x - c) * iw;
        const float e = __builtin_amdgcn_exp2f(z * z * -0.72134752044448170368f);
        const float j1 = aiw * e * z;
        jac[0] = e;
        jac[1] = j1;
        jac[2] = j1 * z;
        return a * e;
    }
    __device__ __forceinline__ float value(float x) const {
        const float z = (x - c) * iw;
        return a * __builtin_amdgcn_exp2f(z * z * -0.72134752044448170368f);
    }
};

template <> struct ModelConsts<LORENTZIAN> {
    float a, c, iw, k;
    __device__ __forceinline__ void set(const float *p) {
        a = p[0];
        c = p[1];
        iw = 1.0f / p[2];
        k = 2.0f * a * iw;
    }
    __device__ __forceinline__ float eval(float x, float *jac) const {
        const float z = (x - c) * iw;
        const float s = __builtin_amdgcn_rcpf(__builtin_fmaf(z, z, 1.0f));
        const float j1 = k * s * s * z;
        jac[0] = s;
        jac[1] = j1;
        jac[2] = j1 * z;
        return a * s;
    }
    __device__ __forceinline__ float value(float x) const {
        const float z = (x - c) * iw;
        return a * __builtin_amdgcn_rcpf(__builtin_fmaf(z, z, 1.0f));
    }
};

// In terms of u = 2z = 2(x − c)/w: exp(−4 ln2 z²) = exp2(−u²), the sign a
// source modifier of v_exp_f32, and 1 + 4z² = 1 + u². Scaling by 2 is exact,
// so u, u², the value and the first, second and fourth Jacobian entries have
// the bits of the same expressions in z (with 8a/w halved to 4a/w), and the
// width entry j1·u is exactly twice j1·z: jac_scale undoes that after the
// sums (fused_pass, information_pass) and in model_eval_kernel.
template <> struct ModelConsts<PSEUDO_VOIGT> {
    float a, c, iw2, eta, gmix, kh;
    __device__ __forceinline__ void set(const float *p) {
        a = p[0];
        c = p[1];
        const float iw = 1.0f / p[2];
        iw2 = 2.0f * iw;
        eta = p[3];
        gmix = 1.0f - p[3];
        kh = 4.0f * a * iw;
    }
    __device__ __forceinline__ float eval(float x, float *jac) const {
        const float u = (x - c) * iw2;
        const float u2 = u * u;
        const float g = __builtin_amdgcn_exp2f(-u2);
        const float l = __builtin_amdgcn_rcpf(1.0f + u2);
        const float shape = __builtin_fmaf(gmix, g, eta * l);
        const float scale = kh * __builtin_fmaf(gmix * 0.69314718055994530942f, g, eta * l * l);
        const float j1 = scale * u;
        jac[0] = shape;
        jac[1] = j1;
        jac[2] = j1 * u;
        jac[3] = a * (l - g);
        return a * shape;
    }
    __device__ __forceinline__ float value(float x) const {
        const float u = (x - c) * iw2;
        const float u2 = u * u;
        const float g = __builtin_amdgcn_exp2f(-u2);
        const float l = __builtin_amdgcn_rcpf(1.0f + u2);
        return a * __builtin_fmaf(gmix, g, eta * l);
    }
};

template <> struct ModelConsts<TWO_LORENTZIAN> {
    ModelConsts<LORENTZIAN> first, second;
    float b0, b1;
    __device__ __forceinline__ void set(const float *p) {
        first.set(p);
        second.set(p + 3);
        b0 = p[6];
        b1 = p[7];
    }
    __device__ __forceinline__ float eval(float x, float *jac) const {
        const float f1 = first.eval(x, jac);
        const float f2 = second.eval(x, jac + 3);
        jac[6] = 1.0f;
        jac[7] = x;
        return f1 + f2 + __builtin_fmaf(b1, x, b0);
    }
    __device__ __forceinline__ float value(float x) const {
        return first.value(x) + second.value(x) + __builtin_fmaf(b1, x, b0);
    }
};

// ---------------------------------------------------------------------------
// Cross-lane all-reduce within aligned groups of G lanes (G = 4, 8, 16, 32, 64).
// Steps: quad_perm xor1, quad_perm xor2, row_half_mirror (8), row_mirror (16),
// ds_swizzle xor16 (32), permlane32_swap (64). Mirror partners of a lane lie
// in the other half of its 8/16-lane block, so after each step every lane
// holds the sum over the block.
__device__ __forceinline__ int dpp_quad_xor1(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_quad_xor2(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_half_mirror(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_mirror(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true);
}
__device__ __forceinline__ int swz_xor16(int v) { return __builtin_amdgcn_ds_swizzle(v, 0x401F); }
// Partner across the two 32-lane halves, read through SGPRs: lanes 0-31 take
// lane 32's value and lanes 32-63 take lane 0's. Only used after the 32-lane
// steps, when every lane of a half already holds that half's result.
__device__ __forceinline__ int perm_xor32(int v) {
    const int upper = __builtin_amdgcn_readlane(v, 32);
    const int lower = __builtin_amdgcn_readlane(v, 0);
    return (threadIdx.x & 32) ? lower : upper;
}

template <int STEP> __device__ __forceinline__ int xlane(int v) {
    if constexpr (STEP == 1)
        return dpp_quad_xor1(v);
    else if constexpr (STEP == 2)
        return dpp_quad_xor2(v);
    else if constexpr (STEP == 4)
        return dpp_half_mirror(v);
    else if constexpr (STEP == 8)
        return dpp_mirror(v);
    else if constexpr (STEP == 16)
        return swz_xor16(v);
    else
        return perm_xor32(v);
}

// Exact xor-D partner of a lane (D = 1 … 32): DPP quad_perm for 1 and 2,
// ds_swizzle (bit mode, within 32 lanes) for 4, 8 and 16, ds_bpermute for 32.
template <int D> __device__ __forceinline__ int xor_lane(int v) {
    if constexpr (D == 1)
        return dpp_quad_xor1(v);
    else if constexpr (D == 2)
        return dpp_quad_xor2(v);
    else if constexpr (D == 4)
        return __builtin_amdgcn_ds_swizzle(v, 0x101F);
    else if constexpr (D == 8)
        return __builtin_amdgcn_ds_swizzle(v, 0x201F);
    else if constexpr (D == 16)
        return __builtin_amdgcn_ds_swizzle(v, 0x401F);
    else
        return __shfl_xor(v, D, 64);
}
template <int D> __device__ __forceinline__ double xor_lane_d(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = xor_lane<D>(int(bits & 0xffffffffll));
    const int hi = xor_lane<D>(int(bits >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// All-reduce of T doubles over an aligned group of G lanes by halving
// (reduce-scatter): at distance D = 1, 2, 4, … each lane keeps one half of its
// current terms — the lower half if its D bit is clear — and adds the
// partner's copy of that half, so the data exchanged halves at every step
// (for 15 terms over 64 lanes: 17 doubles move instead of 90). When one term
// is left the remaining distances sum it whole. Every total then sits in a
// known lane and is broadcast from there, so all lanes of the group hold the
// same bits.
namespace detail {
struct ScatterPlan {  // where term t ends up: lane bits and slot
    int lane, slot;
};
constexpr ScatterPlan scatter_place(int T, int G, int t) {
    int count = T, slot = t, lane = 0;
    for (int d = 1; d < G && count > 1; d <<= 1) {
        const int keep = (count + 1) / 2;  // split point of every lane at this step
        if (slot >= keep) {
            lane |= d;
            slot -= keep;
        }
        count = keep;
    }
    return {lane, slot};
}
template <int G, int D, int C, int T>
__device__ __forceinline__ void scatter_steps(double (&v)[T], int gl) {
    if constexpr (D < G) {
        if constexpr (C > 1) {
            constexpr int K = (C + 1) / 2;
            const bool upper = (gl & D) != 0;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const double hi_half = (K + i < C) ? v[K + i] : 0.0;
                const double keep = upper ? hi_half : v[i];
                const double send = upper ? v[i] : hi_half;
                v[i] = keep + xor_lane_d<D>(send);
            }
            scatter_steps<G, D * 2, K, T>(v, gl);
        } else {
            v[0] += xor_lane_d<D>(v[0]);
            scatter_steps<G, D * 2, 1, T>(v, gl);
        }
    }
    // D == G: every slot a lane still holds is a full group total.
}
template <int G, int T> constexpr int scatter_slots() {
    int count = T;
    for (int d = 1; d < G && count > 1; d <<= 1)
        count = (count + 1) / 2;
    return count;
}
}  // namespace detail

// Groups of G = 64·W lanes (W = 2 or 4 waves) own one fit per workgroup of
// G threads (the threadgroup-per-fit form, src/metal/cooperative.metal:16-32,
// 338-349). A reduction runs within each wave first; the waves' results,
// identical in every lane of a wave, are then combined through LDS in wave
// order, so every lane of the workgroup holds the same bits. Each call is one
// barrier pair; the whole workgroup takes part in every call.
template <int G, typename V, typename Op>
__device__ __forceinline__ V across_waves(V v, Op op) {
    static_assert(G == 128 || G == 256, "workgroup groups are 2 or 4 waves");
    __shared__ V slot[G / 64];
    if ((threadIdx.x & 63) == 0)
        slot[threadIdx.x >> 6] = v;
    __syncthreads();
    V r = slot[0];
#pragma unroll
    for (int w = 1; w < G / 64; ++w)
        r = op(r, slot[w]);
    __syncthreads();
    return r;
}
template <int G, int T> __device__ __forceinline__ void across_waves_sum(double (&v)[T]) {
    static_assert(G == 128 || G == 256, "workgroup groups are 2 or 4 waves");
    __shared__ double slot[G / 64][T];
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int t = 0; t < T; ++t)
            slot[threadIdx.x >> 6][t] = v[t];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < T; ++t) {
        double r = slot[0][t];
#pragma unroll
        for (int w = 1; w < G / 64; ++w)
            r += slot[w][t];
        v[t] = r;
    }
    __syncthreads();
}

template <int G, int T>
__device__ __forceinline__ void group_allreduce_d(double (&v)[T], int gl) {
    if constexpr (G == 1)
        return;
    if constexpr (G > 64) {
        group_allreduce_d<64, T>(v, gl & 63);
        across_waves_sum<G, T>(v);
        return;
    }
    detail::scatter_steps<G, 1, T, T>(v, gl);
    constexpr int S = detail::scatter_slots<G, T>();
    double part[S];
#pragma unroll
    for (int i = 0; i < S; ++i)
        part[i] = v[i];
    const int base = int(threadIdx.x & 63) & ~(G - 1);
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const detail::ScatterPlan pl = detail::scatter_place(T, G, t);
        const long long bits = __double_as_longlong(part[pl.slot]);
        int lo, hi;
        if constexpr (G == 64) {
            lo = __builtin_amdgcn_readlane(int(bits & 0xffffffffll), pl.lane);
            hi = __builtin_amdgcn_readlane(int(bits >> 32), pl.lane);
        } else {
            lo = __shfl(int(bits & 0xffffffffll), base + pl.lane, 64);
            hi = __shfl(int(bits >> 32), base + pl.lane, 64);
        }
        v[t] = __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                                (unsigned long long)(unsigned)lo));
    }
}

template <int STEP> __device__ __forceinline__ float xlane_f(float v) {
    return __int_as_float(xlane<STEP>(__float_as_int(v)));
}
template <int STEP> __device__ __forceinline__ double xlane_d(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = xlane<STEP>(int(bits & 0xffffffffll));
    const int hi = xlane<STEP>(int(bits >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// Sum over the G-lane group in FP32 (plain form).
template <int G> __device__ __forceinline__ float group_sum_f(float v) {
    if constexpr (G > 64)
        return across_waves<G>(group_sum_f<64>(v), [](float a, float b) { return a + b; });
    v += xlane_f<1>(v);
    v += xlane_f<2>(v);
    if constexpr (G >= 8)
        v += xlane_f<4>(v);
    if constexpr (G >= 16)
        v += xlane_f<8>(v);
    if constexpr (G >= 32)
        v += xlane_f<16>(v);
    if constexpr (G >= 64)
        v += xlane_f<32>(v);
    return v;
}
// Sum over the G-lane group in FP64 (combination of FP32 compensated partials).
template <int G> __device__ __forceinline__ double group_sum_d(double v) {
    if constexpr (G > 64)
        return across_waves<G>(group_sum_d<64>(v), [](double a, double b) { return a + b; });
    v += xlane_d<1>(v);
    v += xlane_d<2>(v);
    if constexpr (G >= 8)
        v += xlane_d<4>(v);
    if constexpr (G >= 16)
        v += xlane_d<8>(v);
    if constexpr (G >= 32)
        v += xlane_d<16>(v);
    if constexpr (G >= 64)
        v += xlane_d<32>(v);
    return v;
}
// Bitwise OR / integer max / min over the group (validation flags, indices).
template <int G> __device__ __forceinline__ int group_or(int v) {
    if constexpr (G > 64)
        return across_waves<G>(group_or<64>(v), [](int a, int b) { return a | b; });
    v |= xlane<1>(v);
    v |= xlane<2>(v);
    if constexpr (G >= 8)
        v |= xlane<4>(v);
    if constexpr (G >= 16)
        v |= xlane<8>(v);
    if constexpr (G >= 32)
        v |= xlane<16>(v);
    if constexpr (G >= 64)
        v |= xlane<32>(v);
    return v;
}
// Group maximum of floats (NaN ignored where one operand is a number, as
// fmaxf); group_min_f below is its minimum counterpart.
template <int G> __device__ __forceinline__ float group_max_f(float v) {
    if constexpr (G > 64)
        return across_waves<G>(group_max_f<64>(v), [](float a, float b) { return fmaxf(a, b); });
    v = fmaxf(v, xlane_f<1>(v));
    v = fmaxf(v, xlane_f<2>(v));
    if constexpr (G >= 8)
        v = fmaxf(v, xlane_f<4>(v));
    if constexpr (G >= 16)
        v = fmaxf(v, xlane_f<8>(v));
    if constexpr (G >= 32)
        v = fmaxf(v, xlane_f<16>(v));
    if constexpr (G >= 64)
        v = fmaxf(v, xlane_f<32>(v));
    return v;
}
template <int G> __device__ __forceinline__ int group_max_i(int v) {
    if constexpr (G > 64)
        return across_waves<G>(group_max_i<64>(v), [](int a, int b) { return max(a, b); });
    v = max(v, xlane<1>(v));
    v = max(v, xlane<2>(v));
    if constexpr (G >= 8)
        v = max(v, xlane<4>(v));
    if constexpr (G >= 16)
        v = max(v, xlane<8>(v));
    if constexpr (G >= 32)
        v = max(v, xlane<16>(v));
    if constexpr (G >= 64)
        v = max(v, xlane<32>(v));
    return v;
}
template <int G> __device__ __forceinline__ int group_min_i(int v) {
    if constexpr (G > 64)
        return across_waves<G>(group_min_i<64>(v), [](int a, int b) { return min(a, b); });
    v = min(v, xlane<1>(v));
    v = min(v, xlane<2>(v));
    if constexpr (G >= 8)
        v = min(v, xlane<4>(v));
    if constexpr (G >= 16)
        v = min(v, xlane<8>(v));
    if constexpr (G >= 32)
        v = min(v, xlane<16>(v));
    if constexpr (G >= 64)
        v = min(v, xlane<32>(v));
    return v;
}
template <int G> __device__ __forceinline__ float group_min_f(float v) {
    if constexpr (G > 64)
        return across_waves<G>(group_min_f<64>(v), [](float a, float b) { return fminf(a, b); });
    v = fminf(v, xlane_f<1>(v));
    v = fminf(v, xlane_f<2>(v));
    if constexpr (G >= 8)
        v = fminf(v, xlane_f<4>(v));
    if constexpr (G >= 16)
        v = fminf(v, xlane_f<8>(v));
    if constexpr (G >= 32)
        v = fminf(v, xlane_f<16>(v));
    if constexpr (G >= 64)
        v = fminf(v, xlane_f<32>(v));
    return v;
}

// Grid facts shared by all fits, derived on the device from x
// (cpu/solver.hpp:58-69 grid_spacing; batch.hpp:236-240 varied_x;
//  initial_guess.hpp:131-153 guess preconditions and edge abscissae).
struct GridInfo {
    float min_spacing;      // grid_spacing(): min |Δx| if x is monotone, else 0
    int32_t varied_x;       // any x[j] != x[0]
    int32_t x_finite;       // all x finite
    int32_t guess_grid_ok;  // strictly increasing, finite, finite span
    float guess_spacing;    // min (x[j] - x[j-1])
    float left_x, right_x;  // midpoints of the two edge windows
    int32_t count;          // edge window length max(1, M/10)
    float max_abs_x;        // largest |x[j]| (+inf if some x is not finite)
};

// Grids of |x| up to this bound take the fixed-baseline form of the two-peak
// LM kernels: there the baseline columns' sums (Σ J_k, Σ J_k·x, Σ r, Σ r·x,
// Σ 1, Σ x, Σ x²) are finite whenever the free columns' sums and χ² are
// (Cauchy–Schwarz), so leaving them out changes no finiteness test.
constexpr float kFixedBaselineMaxX = 1.0e6f;

}  // namespace bfhip
