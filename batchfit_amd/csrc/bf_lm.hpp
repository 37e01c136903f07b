// bf_lm.hpp — bounded Levenberg–Marquardt kernels for batches of independent
// spectra on CDNA4 (gfx950).
//
// Mapping: a group of G lanes owns one fit (G = 1 lane-per-fit, or 4/8/16/32/64
// lanes of a wave, or a workgroup of 128/256 lanes; 4-lane groups only stream). With R > 0 the group holds its fit's points in registers, lane l of
// the group owning points l, l+G, …, l+G(R-1); y is read from HBM once per fit.
// With R = 0 every pass streams x and y from memory (lane-per-fit form).
//
// Algorithm: the reference's bounded LM (cpu/solver.hpp:196-342; GPU forms
// src/metal/lm.metal:108-272 and cooperative.metal:232-388) with one change of
// schedule that does not change any result in exact arithmetic: each LM
// iteration makes ONE pass over the data at the trial point, producing χ², JᵀJ
// and Jᵀr together. If the step is accepted, that linearisation is the next
// iteration's; if it is rejected, the previous one is kept. The reference
// instead recomputes the linearisation at the start of every iteration and
// evaluates the trial cost in a separate pass.
//
// Numerics (bf_config.numerics):
//   bit0 (1) compensated per-lane partials (Kahan, numerics.hpp:12-18)
//   bit1 (2) equilibrated Cholesky (cpu/solver.hpp:101-113)
//   bit2 (4) mixed precision: lane partials are combined across the group in
//            FP64, and the P×P solve, accept/reject test, λ schedule and
//            covariance inverse run in FP64. Without bit2 everything is FP32,
//            as in the reference GPU kernels.
// Model value, residual and Jacobian are always FP32.
#pragma once

#include "bf_guess.hpp"

#include <type_traits>

// Contraction only within one source expression, so that a fit's rounding
// does not depend on the code around the point loop (register tile or
// streamed, uniform or ragged launch).
#pragma clang fp contract(on)

namespace bfhip {

struct FitArgs {
    const float *x;
    const float *y;
    long long ys_fit, ys_pt;  // y[fit*ys_fit + j*ys_pt]
    const float *p0;          // [N,P] initial parameters
    float *params;
    int32_t *status;
    uint32_t *iterations;
    float *chi2;
    float *cov;  // [N,P,P] or null
    float *stderr_;
    int32_t *ustatus;
    const GridInfo *grid;
    uint32_t n, m;
    uint32_t max_iterations;
    float initial_lambda, gradient_tolerance, step_tolerance, cost_tolerance;
    float lower[8], upper[8];
    uint32_t numerics;
    uint32_t *pool;      // work counter of the retiring streamed kernel (zeroed), or null
    int32_t options_ok;  // λ and tolerances valid, bounds ordered/non-NaN/non-empty-side
    int32_t free_shape;  // cpu/solver.hpp:70-82
    int32_t flat_check;  // some fit may have flat observations (its p0 is not a device guess)
    float width_factor;  // cpu/solver.hpp:86-88
    int32_t wide_grid_only;  // 1: a full two-peak pool kernel that runs only where the
                             // fixed-baseline form does not (grids beyond kFixedBaselineMaxX)
};

__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    return (v < lo) ? lo : (hi < v) ? hi : v;  // std::clamp
}
template <typename S> __device__ __forceinline__ S smax(S a, S b) { return (a < b) ? b : a; }
template <typename S> __device__ __forceinline__ S smin(S a, S b) { return (b < a) ? b : a; }
template <typename S> __device__ __forceinline__ bool finite_s(S v) {
    if constexpr (sizeof(S) == 8)
        return finite_d(v);
    else
        return finite_f(v);
}
template <typename S> __device__ __forceinline__ S sabs(S v) { return v < S(0) ? -v : v; }

// Packed lower-triangle index of (k, l), l <= k.
__host__ __device__ constexpr int tri(int k, int l) { return k * (k + 1) / 2 + l; }

// Factor that turns pass term t (packed JᵀJ, Jᵀr, χ²) summed from eval()
// Jacobians into the model's own term (jac_scale of its entries).
template <int MODEL> __host__ __device__ constexpr float term_scale(int t) {
    constexpr int P = Params<MODEL>::P, TRI = Params<MODEL>::TRI;
    if (t < TRI) {
        int k = 0;
        while ((k + 1) * (k + 2) / 2 <= t)
            ++k;
        return jac_scale<MODEL>(k) * jac_scale<MODEL>(t - k * (k + 1) / 2);
    }
    return t < TRI + P ? jac_scale<MODEL>(t - TRI) : 1.0f;
}

// Checked, unpivoted Cholesky solve on a packed lower triangle
// (cpu/solver.hpp:99-143). Returns false on a non-positive pivot, a
// non-finite entry or a non-finite solution.
template <typename S, int P>
__device__ __forceinline__ bool chol_solve(S *a, S *b, bool scaled) {
    S d[P], dinv[P];
    if (scaled) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            if (!finite_s(a[tri(i, i)]) || a[tri(i, i)] <= S(0))
                return false;
            d[i] = sqrt(a[tri(i, i)]);
            dinv[i] = S(1) / d[i];
        }
#pragma unroll
        for (int i = 0; i < P; ++i) {
            b[i] *= dinv[i];
#pragma unroll
            for (int j = 0; j <= i; ++j)
                a[tri(i, j)] = (a[tri(i, j)] * dinv[i]) * dinv[j];
        }
    }
    // Factor and forward substitution; one reciprocal per pivot.
    bool ok = true;
    S inv[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            S v = a[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k)
                v -= a[tri(i, k)] * a[tri(j, k)];
            if (i == j) {
                ok = ok && finite_s(v) && v > S(0);
                a[tri(i, i)] = sqrt(v);
                inv[i] = S(1) / a[tri(i, i)];
            } else {
                ok = ok && finite_s(v);
                a[tri(i, j)] = v * inv[j];
            }
        }
        S v = b[i];
#pragma unroll
        for (int j = 0; j < i; ++j)
            v -= a[tri(i, j)] * b[j];
        b[i] = v * inv[i];
    }
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
        S v = b[i];
#pragma unroll
        for (int j = i + 1; j < P; ++j)
            v -= a[tri(j, i)] * b[j];
        b[i] = v * inv[i];
        ok = ok && finite_s(b[i]);
    }
    if (scaled)
#pragma unroll
        for (int i = 0; i < P; ++i) {
            b[i] *= dinv[i];
            ok = ok && finite_s(b[i]);
        }
    return ok;
}

// Per-lane partial accumulation. Plain: one FMA per term and point.
// Compensated: FMA into a block partial over 32 points of the lane, then the
// block partial is added into a Kahan pair (sum, carry) (numerics.hpp:12-18),
// so the compensation costs 4 VALU per term per block instead of per point.
// A lane's partial is (sum − carry).
//
// Flush kinds: FlushClear adds the block into the pair and clears the block
// partial; FlushKeep adds it and leaves the partial for the next block's
// first product to overwrite (add_prod with std::true_type: b = x·y, the same
// value as fma(x, y, +0) except that a −0 product stays −0, which a later
// product or the pair's additions cannot tell from +0, since the pair starts
// at +0); ClearOnly clears the partial.
struct FlushClear {};
struct FlushKeep {};
struct ClearOnly {};
template <bool COMP> struct Acc;
template <> struct Acc<false> {
    float s;
    __device__ __forceinline__ void zero() { s = 0.0f; }
    template <bool FIRST>
    __device__ __forceinline__ void add_prod(float a, float b, std::bool_constant<FIRST> = {}) {
        s = __builtin_fmaf(a, b, s);
    }
    template <typename Kind> __device__ __forceinline__ void flush(Kind) {}
    __device__ __forceinline__ void flush() {}
    __device__ __forceinline__ float hi() const { return s; }
    __device__ __forceinline__ float lo() const { return 0.0f; }
};
template <> struct Acc<true> {
    float s, c, b;
    __device__ __forceinline__ void zero() { s = c = b = 0.0f; }
    template <bool FIRST>
    __device__ __forceinline__ void add_prod(float x, float y, std::bool_constant<FIRST> = {}) {
        if constexpr (FIRST)
            b = x * y;
        else
            b = __builtin_fmaf(x, y, b);
    }
    __device__ __forceinline__ void add_pair() {
#pragma clang fp contract(off)
        const float adjusted = b - c;
        const float next = s + adjusted;
        c = (next - s) - adjusted;
        s = next;
    }
    __device__ __forceinline__ void flush(FlushClear = {}) {
        add_pair();
        b = 0.0f;
    }
    __device__ __forceinline__ void flush(FlushKeep) { add_pair(); }
    __device__ __forceinline__ void flush(ClearOnly) { b = 0.0f; }
    __device__ __forceinline__ float hi() const { return s; }
    __device__ __forceinline__ float lo() const { return c; }
};

// Group total of one term: FP64 in mixed mode, FP32 otherwise.
template <int G, typename S, bool COMP>
__device__ __forceinline__ S reduce_term(const Acc<COMP> &a) {
    if constexpr (sizeof(S) == 8) {
        double v = double(a.hi()) - double(a.lo());
        if constexpr (G > 1)
            v = group_sum_d<G>(v);
        return v;
    } else {
        float v = a.hi();
        if constexpr (G > 1)
            v = group_sum_f<G>(v);
        if constexpr (COMP) {
            float c = a.lo();
            if constexpr (G > 1)
                c = group_sum_f<G>(c);
            v -= c;
        }
        return v;
    }
}

// Points per lane loaded at once when streaming (R == 0): 16, or 8 for the
// two-peak model, whose larger state leaves fewer registers free.
constexpr int stream_chunk(int model) { return model == TWO_LORENTZIAN ? 8 : 16; }

// Points of one fit as seen by one lane: a register tile (R > 0) or a
// strided stream from memory (R == 0), CHUNK points per lane at a time.
// With a register tile, x is read from the block's LDS copy of the shared
// grid (every fit of a launch shares it), which keeps R VGPRs free.
template <int G, int R, int CHUNK = 16> struct Points {
    float yv[R > 0 ? R : 1];
    const float *x, *y;
    long long ys_pt;
    int gl, m;
    // Observation j of the fit. Group mappings read fit-major rows (ys_pt is
    // 1, so a lane's loads take immediate offsets); only the lane-per-fit
    // form may read a point-major copy.
    __device__ __forceinline__ float obs(int j) const {
        if constexpr (G > 1)
            return y[j];
        else
            return y[j * ys_pt];
    }
    __device__ __forceinline__ void load(const float *x_, const float *yrow, long long ys_pt_,
                                         int gl_, int m_) {
        x = x_;
        y = yrow;
        ys_pt = ys_pt_;
        gl = gl_;
        m = m_;
        if constexpr (R > 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = gl + G * r;
                yv[r] = j < m ? obs(j) : 0.0f;
            }
        }
    }
    // f(x, y) or, where f takes it, f(x, y, first): first is
    // std::bool_constant<true> for the first point of a whole block (every
    // lane of the group holds every point of it), after a FlushKeep.
    template <typename F, bool FIRST>
    __device__ __forceinline__ static void point(F &f, float xv, float yv,
                                                 std::bool_constant<FIRST>, bool first) {
        if constexpr (std::is_invocable_v<F &, float, float, std::true_type>) {
            if constexpr (FIRST) {
                if (first)
                    f(xv, yv, std::true_type{});
                else
                    f(xv, yv, std::false_type{});
            } else {
                f(xv, yv, std::false_type{});
            }
        } else {
            f(xv, yv);
        }
    }
    // flush() or, where it takes one, flush(kind); a flush() with no kind
    // runs at FlushKeep (it holds no partials) and not at ClearOnly.
    template <typename Flush, typename Kind>
    __device__ __forceinline__ static void flush_as(Flush &flush, Kind kind) {
        if constexpr (std::is_invocable_v<Flush &, Kind>)
            flush(kind);
        else if constexpr (!std::is_same_v<Kind, ClearOnly>)
            flush();
    }
    // f(x, y) for every point this lane owns, then flush() after every block
    // of kBlock points. Slots below m/G are full for every lane of the group
    // (a scalar branch); only slot m/G is partial.
    static constexpr int kBlock = 32;
    static constexpr int kChunk = CHUNK;
    template <typename F, typename Flush>
    __device__ __forceinline__ void each(F &&f, Flush &&flush) const {
        if constexpr (R > 0) {
            if (m == G * R) {
                // Every slot full: straight-line code with no joins.
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    point(f, x[gl + G * r], yv[r], std::bool_constant<false>{}, false);
                    if (r % kBlock == kBlock - 1 || r == R - 1)
                        flush_as(flush, FlushClear{});
                }
                return;
            }
            // Blocks that hold no point of the group are skipped entirely, so
            // the arithmetic of a fit depends on (G, M) and not on the tile R.
            const int full = m / G;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r < full)
                    point(f, x[gl + G * r], yv[r], std::bool_constant<false>{}, false);
                else if (r == full && gl + G * r < m)
                    point(f, x[gl + G * r], yv[r], std::bool_constant<false>{}, false);
                if ((r % kBlock == kBlock - 1 || r == R - 1) && G * (r - r % kBlock) < m)
                    flush_as(flush, FlushClear{});
            }
        } else {
            int j = gl;
            int k = 0;  // points since the last flush
            if constexpr (G > 1) {
                // Whole chunks of kChunk points per lane: all their loads are
                // issued before the first point is evaluated, so one memory
                // latency is paid per chunk instead of per point. Points and
                // flushes come in the same order as in the per-point loop.
                static_assert(kBlock % kChunk == 0, "a block holds whole chunks");
                // Whole blocks first: the block's first point starts its
                // partials and the flush keeps them, so they are cleared once
                // after the loop instead of at every flush.
                for (; j + G * (kBlock - 1) < m; j += G * kBlock) {
#pragma unroll
                    for (int h = 0; h < kBlock / kChunk; ++h) {
                        float xc[kChunk], yc[kChunk];
#pragma unroll
                        for (int c = 0; c < kChunk; ++c) {
                            xc[c] = x[j + G * (h * kChunk + c)];
                            yc[c] = obs(j + G * (h * kChunk + c));
                        }
                        point(f, xc[0], yc[0], std::bool_constant<true>{}, h == 0);
#pragma unroll
                        for (int c = 1; c < kChunk; ++c)
                            point(f, xc[c], yc[c], std::bool_constant<false>{}, false);
                    }
                    flush_as(flush, FlushKeep{});
                }
                flush_as(flush, ClearOnly{});
                for (; j + G * (kChunk - 1) < m; j += G * kChunk) {
                    float xc[kChunk], yc[kChunk];
#pragma unroll
                    for (int c = 0; c < kChunk; ++c) {
                        xc[c] = x[j + G * c];
                        yc[c] = obs(j + G * c);
                    }
#pragma unroll
                    for (int c = 0; c < kChunk; ++c) {
                        point(f, xc[c], yc[c], std::bool_constant<false>{}, false);
                        if constexpr (kChunk >= kBlock)
                            if (c % kBlock == kBlock - 1)
                                flush_as(flush, FlushClear{});
                    }
                    if constexpr (kBlock > kChunk) {
                        k += kChunk;
                        if (k == kBlock) {
                            flush_as(flush, FlushClear{});
                            k = 0;
                        }
                    }
                }
            }
            for (; j < m; j += G) {
                point(f, x[j], obs(j), std::bool_constant<false>{}, false);
                if (++k == kBlock) {
                    flush_as(flush, FlushClear{});
                    k = 0;
                }
            }
            if (k)
                flush_as(flush, FlushClear{});
            if constexpr (G > 1) {
                // The register tile flushes every block that holds a point of
                // the group, so a lane whose points end one block early
                // flushes once more: a fit's arithmetic depends on (G, M)
                // only, whether its points stream or sit in registers.
                const int mine = (m - gl + G - 1) / G;
                if ((mine + kBlock - 1) / kBlock < (m + kBlock * G - 1) / (kBlock * G))
                    flush_as(flush, FlushClear{});
            }
        }
    }
    template <typename F> __device__ __forceinline__ void each(F &&f) const {
        each(f, [] {});
    }
};

// Index in the full term layout (packed JᵀJ, Jᵀr, χ² of P parameters) of term
// t of the layout of the first PA parameters (the packed triangle of the
// first PA rows is the start of the full one).
template <int P, int PA> __host__ __device__ constexpr int full_term(int t) {
    constexpr int TRI = P * (P + 1) / 2, TRIA = PA * (PA + 1) / 2;
    return t < TRIA ? t : t < TRIA + PA ? TRI + (t - TRIA) : TRI + P;
}

// One pass at parameters p: χ² (terms[TRI+P]), JᵀJ packed (terms[0..TRI)),
// Jᵀr (terms[TRI..TRI+P)), reduced over the group. `lin_ok` is the finiteness
// check of the linearisation (cpu/solver.hpp:182-190). Only the first PA
// parameters' rows are summed (PA < P for the fixed-baseline form of the
// two-peak model); the other rows' terms are written as 0 and take no part in
// `lin_ok` (the caller guarantees they are never used and always finite).
template <int MODEL, int G, int R, bool COMP, typename S, int PA = Params<MODEL>::P>
__device__ __forceinline__ void fused_pass(const Points<G, R, stream_chunk(MODEL)> &pts, const float *p, S *terms,
                                           bool &lin_ok, float *yrange = nullptr) {
    constexpr int P = Params<MODEL>::P, TRI = Params<MODEL>::TRI;
    constexpr int TRIA = PA * (PA + 1) / 2, TA = TRIA + PA + 1;
    ModelConsts<MODEL> mc;
    mc.set(p);
    Acc<COMP> acc[TA];
#pragma unroll
    for (int t = 0; t < TA; ++t)
        acc[t].zero();
    // With yrange: also the group's smallest and largest observation.
    float ylo = __uint_as_float(0x7f800000u), yhi = -__uint_as_float(0x7f800000u);
    pts.each([&](float xj, float yj, auto first) {
        if (yrange) {
            ylo = fminf(ylo, yj);
            yhi = fmaxf(yhi, yj);
        }
        float jac[P];
        const float f = mc.eval(xj, jac);
        const float r = yj - f;
#pragma unroll
        for (int k = 0; k < PA; ++k) {
#pragma unroll
            for (int l = 0; l <= k; ++l)
                acc[tri(k, l)].add_prod(jac[k], jac[l], first);
            acc[TRIA + k].add_prod(jac[k], r, first);
        }
        acc[TRIA + PA].add_prod(r, r, first);
    }, [&](auto kind) {
#pragma unroll
        for (int t = 0; t < TA; ++t)
            acc[t].flush(kind);
    });
    if (yrange) {
        if constexpr (G > 1) {
            ylo = group_min_f<G>(ylo);
            yhi = group_max_f<G>(yhi);
        }
        yrange[0] = ylo;
        yrange[1] = yhi;
    }
    lin_ok = true;
    if constexpr (PA < P) {
#pragma unroll
        for (int t = TRIA; t < TRI; ++t)
            terms[t] = S(0);
#pragma unroll
        for (int k = PA; k < P; ++k)
            terms[TRI + k] = S(0);
    }
    if constexpr (sizeof(S) == 8) {
        // Each lane's (sum, carry) pair as one double, then one halving
        // all-reduce over the group for all terms together.
        double v[TA];
#pragma unroll
        for (int t = 0; t < TA; ++t)
            v[t] = double(acc[t].hi()) - double(acc[t].lo());
        group_allreduce_d<G, TA>(v, pts.gl);
#pragma unroll
        for (int t = 0; t < TA; ++t) {
            const int u = full_term<P, PA>(t);
            terms[u] = v[t] * double(term_scale<MODEL>(u));
            if (t < TRIA + PA)
                lin_ok = lin_ok && finite_s(terms[u]);
        }
    } else {
#pragma unroll
        for (int t = 0; t < TA; ++t) {
            const int u = full_term<P, PA>(t);
            terms[u] = reduce_term<G, S, COMP>(acc[t]) * term_scale<MODEL>(u);
            if (t < TRIA + PA)
                lin_ok = lin_ok && finite_s(terms[u]);
        }
    }
}

// JᵀJ only (uncertainty at the accepted point), always compensated per lane
// (cpu/uncertainty.hpp:297-309); `jac_ok` is false if any Jacobian entry is
// non-finite. Only the rows of the first PA parameters are summed (the lower
// triangle's first PA(PA+1)/2 terms); the others are written as zeros, for a
// caller that replaces them (fixed parameters). Each term's sum is formed the
// same way whatever PA is, so the summed terms keep their bits.
template <int MODEL, int G, int R, typename S, int PA = Params<MODEL>::P>
__device__ __forceinline__ void information_pass(const Points<G, R, stream_chunk(MODEL)> &pts, const float *p,
                                                 S *info, bool &jac_ok) {
    constexpr int P = Params<MODEL>::P, TRI = Params<MODEL>::TRI, TA = PA * (PA + 1) / 2;
    ModelConsts<MODEL> mc;
    mc.set(p);
    Acc<true> acc[TA];
#pragma unroll
    for (int t = 0; t < TA; ++t)
        acc[t].zero();
    int bad = 0;
    pts.each([&](float xj, float, auto first) {
        float jac[P];
        (void)mc.eval(xj, jac);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            bad |= !finite_f(jac[k]);
            if (k < PA) {
#pragma unroll
                for (int l = 0; l <= k; ++l)
                    acc[tri(k, l)].add_prod(jac[k], jac[l], first);
            }
        }
    }, [&](auto kind) {
#pragma unroll
        for (int t = 0; t < TA; ++t)
            acc[t].flush(kind);
    });
    if constexpr (G > 1)
        bad = group_or<G>(bad);
    jac_ok = !bad;
    if constexpr (sizeof(S) == 8) {
        double v[TA];
#pragma unroll
        for (int t = 0; t < TA; ++t)
            v[t] = double(acc[t].hi()) - double(acc[t].lo());
        group_allreduce_d<G, TA>(v, pts.gl);
#pragma unroll
        for (int t = 0; t < TA; ++t)
            info[t] = v[t] * double(term_scale<MODEL>(t));
    } else {
#pragma unroll
        for (int t = 0; t < TA; ++t)
            info[t] = reduce_term<G, S, true>(acc[t]) * term_scale<MODEL>(t);
    }
#pragma unroll
    for (int t = TA; t < TRI; ++t)
        info[t] = S(0);
}

template <int MODEL> __device__ __forceinline__ int32_t resolved_status(const float *p,
                                                                        float spacing,
                                                                        float factor) {
    if (spacing > 0.0f && fabsf(p[2]) * factor < spacing)
        return ST_UNRESOLVED;
    if constexpr (Params<MODEL>::P == 8)
        if (spacing > 0.0f && fabsf(p[5]) * factor < spacing)
            return ST_UNRESOLVED;
    return ST_CONVERGED;
}

// Covariance at the accepted point (cpu/uncertainty.hpp:13-110), written for
// one fit by lane 0 of its group. FIXB (the two-peak model with both
// baseline parameters fixed, on a grid with |x| ≤ kFixedBaselineMaxX): the
// information pass sums only the six free parameters' rows; the baseline
// rows are replaced by the identity's below in either form, and on such a
// grid their sums are finite whenever the free rows' are (Cauchy–Schwarz),
// so no finiteness test changes. CLDS (groups of at most 64 lanes): the
// unsymmetrised covariance is kept in cl (P·P values, the group's own, in
// LDS) instead of in every lane's registers (for P = 8 in FP64 that is 128
// registers, which would hold the kernel to one wave per SIMD), and its
// columns are solved by different lanes instead of by every lane in turn.
template <int MODEL, int G, int R, typename S, bool FIXB = false, bool CLDS = false>
__device__ __forceinline__ void covariance(const FitArgs &a, const Points<G, R, stream_chunk(MODEL)> &pts, uint32_t fit,
                                           int gl, const float *p, int32_t status, float chi2,
                                           S *cl = nullptr) {
    constexpr int P = Params<MODEL>::P, TRI = Params<MODEL>::TRI;
    const float nan = qnan_f();
    float *cov = a.cov + size_t(fit) * P * P;
    float *se = a.stderr_ + size_t(fit) * P;
    auto fail = [&](int32_t us) {
        if (gl == 0) {
            for (int k = 0; k < P * P; ++k)
                cov[k] = nan;
            for (int k = 0; k < P; ++k)
                se[k] = nan;
            a.ustatus[fit] = us;
        }
    };
    if (status != ST_CONVERGED)
        return fail(US_NOT_CONVERGED);
    if (!finite_f(chi2) || chi2 < 0.0f)
        return fail(US_INVALID);
    bool fixed[P];
    int free = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        if (!finite_f(p[k]) || nan_f(a.lower[k]) || nan_f(a.upper[k]) || a.lower[k] > a.upper[k] ||
            p[k] < a.lower[k] || p[k] > a.upper[k])
            return fail(US_INVALID);
        fixed[k] = a.lower[k] == a.upper[k];
        free += !fixed[k];
        if (!fixed[k] && (p[k] == a.lower[k] || p[k] == a.upper[k]))
            return fail(US_ACTIVE_BOUND);
    }
    if (int(a.m) <= free)
        return fail(US_DOF);
    if (!a.grid->x_finite)
        return fail(US_INVALID);
    S info[TRI];
    bool jac_ok;
    information_pass<MODEL, G, R, S, (FIXB && MODEL == TWO_LORENTZIAN) ? 6 : P>(pts, p, info, jac_ok);
    if (!jac_ok)
        return fail(US_INVALID);
#pragma unroll
    for (int t = 0; t < TRI; ++t)
        if (!finite_s(info[t]))
            return fail(US_INVALID);
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
        for (int l = 0; l <= i; ++l)
            if (fixed[i] || fixed[l])
                info[tri(i, l)] = i == l ? S(1) : S(0);
    // The variance uses the reported FP32 χ², as the FP32 reference does.
    const S variance = S(chi2) / S(float(a.m - free));
    if (!finite_s(variance))
        return fail(US_INVALID);
    // The factor does not depend on the right-hand side: factor once, then
    // solve each unit column (the reference refactors per column; same values).
    S d[P];
#pragma unroll
    for (int i = 0; i < P; ++i)
        d[i] = sqrt(info[tri(i, i)]);
    S factor[TRI];
#pragma unroll
    for (int t = 0; t < TRI; ++t)
        factor[t] = info[t];
    S unit[P];
#pragma unroll
    for (int k = 0; k < P; ++k)
        unit[k] = S(0);
    // Factor with a zero right-hand side to obtain the scaled factor.
    if (!chol_solve<S, P>(factor, unit, true))
        return fail(US_SINGULAR);
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (!fixed[i] && factor[tri(i, i)] * factor[tri(i, i)] <= S(64) * S(1.1920928955078125e-7))
            return fail(US_SINGULAR);
    // c[i·P + col]: in cl (lane 0's copy) or in registers; its diagonal also
    // in registers, for the sign test.
    S creg[CLDS ? 1 : P * P];
    S cdiag[CLDS ? 1 : P];
    uint32_t fixed_mask = 0;
#pragma unroll
    for (int k = 0; k < P; ++k)
        fixed_mask |= uint32_t(fixed[k]) << k;
    auto cset = [&](int i, int col, S v) {
        if constexpr (CLDS) {
            cl[i * P + col] = v;  // by the lane that solves the column
        } else {
            creg[i * P + col] = v;
            if (i == col)
                cdiag[i] = v;
        }
    };
    auto cget = [&](int i, int j) {
        if constexpr (CLDS)
            return cl[i * P + j];
        else
            return creg[i * P + j];
    };
    // Column col of the covariance: 0, or the uncertainty status to fail
    // with.
    auto column = [&](int col) -> int32_t {
        S b[P];
#pragma unroll
        for (int i = 0; i < P; ++i)
            b[i] = (i == col ? S(1) : S(0)) / d[i];
        // forward and back substitution with the equilibrated factor
#pragma unroll
        for (int i = 0; i < P; ++i) {
            S v = b[i];
#pragma unroll
            for (int j = 0; j < i; ++j)
                v -= factor[tri(i, j)] * b[j];
            b[i] = v / factor[tri(i, i)];
        }
        bool ok = true;
#pragma unroll
        for (int i = P - 1; i >= 0; --i) {
            S v = b[i];
#pragma unroll
            for (int j = i + 1; j < P; ++j)
                v -= factor[tri(j, i)] * b[j];
            b[i] = v / factor[tri(i, i)];
            ok = ok && finite_s(b[i]);
        }
#pragma unroll
        for (int i = 0; i < P; ++i) {
            b[i] /= d[i];
            ok = ok && finite_s(b[i]);
        }
        if ((fixed_mask >> col) & 1) {
#pragma unroll
            for (int i = 0; i < P; ++i)
                cset(i, col, S(0));
            return 0;
        }
        if (!ok)
            return US_SINGULAR;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const S v = fixed[i] ? S(0) : b[i] * variance;
            cset(i, col, v);
            if (!finite_s(v))
                return US_INVALID;
        }
        return 0;
    };
    if constexpr (CLDS) {
        // The columns are independent: lane gl of the group solves columns
        // gl, gl + G, … into cl (the registers hold the factor and one
        // column). The first failing column in index order decides the
        // status, as when the columns are taken in turn; then lane 0 tests
        // and writes the result.
        int first_fail = 0x7fffffff;  // 8·column + status
#pragma unroll 1
        for (int col = gl; col < P; col += G)
            if (const int32_t us = column(col)) {
                first_fail = 8 * col + us;
                break;
            }
        first_fail = group_min_i<G>(first_fail);
        if (first_fail != 0x7fffffff)
            return fail(first_fail & 7);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (gl == 0) {
#pragma unroll
            for (int i = 0; i < P; ++i)
                if (cl[i * P + i] < S(0))
                    return fail(US_INVALID);
#pragma unroll
            for (int i = 0; i < P; ++i) {
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const S sym = i == j ? cl[i * P + i] : cl[i * P + j] / S(2) + cl[j * P + i] / S(2);
                    cov[i * P + j] = float(sym);
                }
                se[i] = float(sqrt(cl[i * P + i]));
            }
            a.ustatus[fit] = US_AVAILABLE;
        }
        return;
    }
#pragma unroll
    for (int col = 0; col < P; ++col)
        if (const int32_t us = column(col))
            return fail(us);
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (cdiag[i] < S(0))
            return fail(US_INVALID);
    if (gl == 0) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const S sym = i == j ? cdiag[i] : cget(i, j) / S(2) + cget(j, i) / S(2);
                cov[i * P + j] = float(sym);
            }
            se[i] = float(sqrt(cdiag[i]));
        }
        a.ustatus[fit] = US_AVAILABLE;
    }
}

// The LM driver for one fit, run redundantly by every lane of its group.
template <int MODEL, int G, int R, bool COMP, typename S>
__device__ __forceinline__ void solve_fit(const FitArgs &a, uint32_t fit, int gl,
                                          const float *xgrid) {
    constexpr int P = Params<MODEL>::P, TRI = Params<MODEL>::TRI;
    const GridInfo grid = *a.grid;
    Points<G, R, stream_chunk(MODEL)> pts;
    pts.load(xgrid, a.y + fit * a.ys_fit, a.ys_pt, gl, int(a.m));

    float init[P], p[P];
#pragma unroll
    for (int k = 0; k < P; ++k)
        init[k] = a.p0[size_t(fit) * P + k];
    bool bad = !a.options_ok || !grid.x_finite;
#pragma unroll
    for (int k = 0; k < P; ++k)
        bad = bad || !finite_f(init[k]);
    // Per-fit observation checks: finiteness and flatness (lm.metal:133-139).
    const float y0 = a.y[fit * a.ys_fit];
    int flags = 0;
    pts.each([&](float, float yj) {
        flags |= finite_f(yj) ? 0 : 1;
        flags |= (yj != y0) ? 2 : 0;
    });
    if constexpr (G > 1)
        flags = group_or<G>(flags);
    bad = bad || (flags & 1);

    float *out_p = a.params + size_t(fit) * P;
    auto write = [&](const float *q, int32_t st, uint32_t it, float c2) {
        if (gl == 0) {
#pragma unroll
            for (int k = 0; k < P; ++k)
                out_p[k] = q[k];
            a.status[fit] = st;
            a.iterations[fit] = it;
            a.chi2[fit] = c2;
        }
    };
    if (bad) {
        write(init, ST_INVALID, 0, qnan_f());
        if (a.cov)
            covariance<MODEL, G, R, S>(a, pts, fit, gl, init, ST_INVALID, qnan_f());
        return;
    }
#pragma unroll
    for (int k = 0; k < P; ++k)
        p[k] = clampf(init[k], a.lower[k], a.upper[k]);

    S terms[Params<MODEL>::TERMS];
    bool lin_ok;
    fused_pass<MODEL, G, R, COMP, S>(pts, p, terms, lin_ok);
    S chi = terms[TRI + P];
    int32_t status = ST_MAX_ITERATIONS;
    uint32_t iterations = 0;
    if (!finite_s(chi)) {
        write(p, ST_INVALID, 0, qnan_f());
        if (a.cov)
            covariance<MODEL, G, R, S>(a, pts, fit, gl, p, ST_INVALID, qnan_f());
        return;
    }
    if (!(flags & 2) && grid.varied_x && a.free_shape) {
        write(p, ST_DEGENERATE, 0, float(chi));
        if (a.cov)
            covariance<MODEL, G, R, S>(a, pts, fit, gl, p, ST_DEGENERATE, float(chi));
        return;
    }
    S A[TRI], g[P];
#pragma unroll
    for (int t = 0; t < TRI; ++t)
        A[t] = terms[t];
#pragma unroll
    for (int k = 0; k < P; ++k)
        g[k] = terms[TRI + k];

    constexpr float eps = 1.1920928955078125e-7f;  // FLT_EPSILON: the options are FP32
    const S lambda0 = S(a.initial_lambda);
    const S lambda_cap = smax(lambda0, S(1.0f / eps));
    const S gtol = S(a.gradient_tolerance), stol = S(a.step_tolerance),
            ctol = S(a.cost_tolerance);
    const bool scaled = (a.numerics & 2u) != 0;
    S lambda = lambda0;
    for (uint32_t it = 0; it < a.max_iterations; ++it) {
        if (!lin_ok) {
            status = ST_INVALID;
            break;
        }
        bool pinned[P];
        S gnorm = S(0);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            pinned[k] = a.lower[k] == a.upper[k] || (p[k] == a.lower[k] && g[k] < S(0)) ||
                        (p[k] == a.upper[k] && g[k] > S(0));
            if (!pinned[k])
                gnorm = smax(gnorm, sabs(g[k]));
        }
        if (gnorm <= gtol) {
            status = resolved_status<MODEL>(p, grid.min_spacing, a.width_factor);
            break;
        }
        S damped[TRI], step[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            step[k] = pinned[k] ? S(0) : g[k];
#pragma unroll
            for (int l = 0; l <= k; ++l)
                damped[tri(k, l)] = (pinned[k] || pinned[l]) ? (k == l ? S(1) : S(0)) : A[tri(k, l)];
            if (!pinned[k])
                damped[tri(k, k)] += lambda * A[tri(k, k)];
        }
        iterations = it + 1;
        if (!chol_solve<S, P>(damped, step, scaled)) {
            status = ST_SINGULAR;
            break;
        }
        float trial[P];
        S step_norm = S(0), param_norm = S(0);
        bool finite_trial = true;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            step_norm = smax(step_norm, sabs(step[k]));
            trial[k] = clampf(float(S(p[k]) + step[k]), a.lower[k], a.upper[k]);
            finite_trial = finite_trial && finite_f(trial[k]);
            step[k] = S(trial[k]) - S(p[k]);
            param_norm = smax(param_norm, S(fabsf(p[k])));
        }
        if (step_norm <= stol * (param_norm + stol) / smax(S(1), lambda)) {
            status = resolved_status<MODEL>(p, grid.min_spacing, a.width_factor);
            break;
        }
        S predicted = S(0);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            predicted += S(2) * g[k] * step[k];
#pragma unroll
            for (int l = 0; l < P; ++l)
                predicted -= step[k] * A[k >= l ? tri(k, l) : tri(l, k)] * step[l];
        }
        S trial_chi = S(qnan_f());
        bool trial_ok = false;
        if (finite_trial) {
            fused_pass<MODEL, G, R, COMP, S>(pts, trial, terms, trial_ok);
            trial_chi = terms[TRI + P];
        }
        const S reduction = chi - trial_chi;
        const S threshold = ctol * chi;
        const bool small_change = finite_s(trial_chi) && finite_s(predicted) &&
                                  predicted >= S(0) && sabs(reduction) <= threshold &&
                                  predicted <= threshold / smax(S(1), lambda);
        if (finite_s(trial_chi) && reduction > S(0) && finite_s(predicted) && predicted > S(0)) {
#pragma unroll
            for (int k = 0; k < P; ++k)
                p[k] = trial[k];
            chi = trial_chi;
#pragma unroll
            for (int t = 0; t < TRI; ++t)
                A[t] = terms[t];
#pragma unroll
            for (int k = 0; k < P; ++k)
                g[k] = terms[TRI + k];
            lin_ok = trial_ok;
            const S rho = reduction / predicted;
            const S q = S(2) * rho - S(1);
            const S factor = rho >= S(1) ? S(1) / S(3) : smax(S(1) / S(3), S(1) - q * q * q);
            const S next = lambda * factor;
            lambda = next < S(eps) ? S(eps) : (lambda_cap < next ? lambda_cap : next);
        } else {
            lambda = smin(lambda_cap, lambda * S(10));
        }
        if (small_change) {
            status = resolved_status<MODEL>(p, grid.min_spacing, a.width_factor);
            break;
        }
    }
    const float chi_out = float(chi);
    write(p, status, iterations, chi_out);
    if (a.cov)
        covariance<MODEL, G, R, S>(a, pts, fit, gl, p, status, chi_out);
}

// Streamed groups that retire fits one by one. A wave owns a pool of fits
// [next, end) (a slice of the batch, or tickets of a ragged run) and each of
// its 64/G groups takes the next fit of the pool as soon as its own is done,
// so no group idles while another fit of the wave still iterates. Every trip
// of the loop makes one fused pass for every group that needs one (the others
// pass over zero points), so the cross-lane reductions run in all lanes
// together and the loop exit is wave-uniform. A fit goes through exactly the
// steps of solve_fit (same passes, same order), so its bytes are the same.
// The source of fits `src` provides: take(want) — called by every lane with
// a request flag uniform within each group, returns a fit of the pool or
// kNoFit; begin(fit, pts) — makes `fit` the group's current fit and sets its
// points; args() / grid() — the current fit's arguments and grid facts; row()
// — the current fit's output row in args().
constexpr uint32_t kNoFit = 0xffffffffu;

// FLAT: test the degenerate case (flat observations, cpu/solver.hpp:237-245)
// from each first pass's observation range. A fit that starts from a
// successful device guess cannot have flat observations (the guess needs a
// smoothed peak above three noise scales), so launches whose every initial
// point is a guess leave the range out of their passes.
//
// FIXB (two-peak model, both baseline parameters fixed by equal bounds — the
// caller guarantees it — on a grid within kFixedBaselineMaxX): the baseline
// parameters are pinned at compile time, so their rows of JᵀJ and their
// gradient, which a pinned parameter never reads (identity rows in the damped
// matrix, zero steps), are not summed, and the predicted gain leaves out
// their terms, which are exactly zero (a fixed parameter's step is 0).
template <int MODEL, int G, bool COMP, bool FLAT, typename S, bool FIXB = false, typename Source>
__device__ __forceinline__ void solve_pool(int gl, Source &src) {
    constexpr int P = Params<MODEL>::P, TRI = Params<MODEL>::TRI;
    constexpr int PA = (FIXB && MODEL == TWO_LORENTZIAN) ? 6 : P;  // parameters summed
    constexpr float eps = 1.1920928955078125e-7f;
    // A group that never gets a fit still takes part in every pass, over zero
    // points: its lane index must be set, since the pass loops start there.
    Points<G, 0, stream_chunk(MODEL)> pts;
    pts.load(nullptr, nullptr, 0, gl, 0);
    uint32_t fit = kNoFit;
    bool iterating = false;  // false: the fit's first pass is next
    float p[P], trial[P];
    S A[TRI], g[P];
    S chi = S(0), lambda = S(0), predicted = S(0);
    bool lin_ok = false;
    uint32_t it = 0;
#pragma unroll
    for (int k = 0; k < P; ++k)
        p[k] = trial[k] = 0.0f;
#pragma unroll
    for (int t = 0; t < TRI; ++t)
        A[t] = S(0);
#pragma unroll
    for (int k = 0; k < P; ++k)
        g[k] = S(0);

    auto write = [&](const float *q, int32_t st, uint32_t iters, float c2) {
        const FitArgs &a = src.args();
        const size_t row = src.row();
        if (gl == 0) {
#pragma unroll
            for (int k = 0; k < P; ++k)
                a.params[row * P + k] = q[k];
            a.status[row] = st;
            a.iterations[row] = iters;
            a.chi2[row] = c2;
        }
        fit = kNoFit;
    };
    auto finish = [&](int32_t st) { write(p, st, it, float(chi)); };
    auto resolved = [&]() {
        return resolved_status<MODEL>(p, src.grid().min_spacing, src.args().width_factor);
    };

    for (;;) {
        // Every trip starts at this wave-wide test, a convergent operation the
        // optimiser cannot duplicate into the paths that end the previous
        // trip; so all lanes begin each trip together and the queue below is
        // always read by the whole wave. The loop ends when no group holds a
        // fit and the queue is empty.
        if (!__ballot(fit != kNoFit) && src.dry)
            break;
        const FitArgs &a = src.args();
        // 1. The top of an iteration (cpu/solver.hpp:254-312) for the groups
        //    that iterate: stops, damped solve, trial point, predicted gain.
        bool pass = false;
        if (fit != kNoFit && iterating) {
            if (it >= a.max_iterations) {
                finish(ST_MAX_ITERATIONS);
            } else if (!lin_ok) {
                finish(ST_INVALID);
            } else {
                bool pinned[P];
                S gnorm = S(0);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    pinned[k] = k >= PA || a.lower[k] == a.upper[k] ||
                                (p[k] == a.lower[k] && g[k] < S(0)) ||
                                (p[k] == a.upper[k] && g[k] > S(0));
                    if (!pinned[k])
                        gnorm = smax(gnorm, sabs(g[k]));
                }
                if (gnorm <= S(a.gradient_tolerance)) {
                    finish(resolved());
                } else {
                    S damped[TRI], step[P];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        step[k] = pinned[k] ? S(0) : g[k];
#pragma unroll
                        for (int l = 0; l <= k; ++l)
                            damped[tri(k, l)] = (pinned[k] || pinned[l]) ? (k == l ? S(1) : S(0))
                                                                         : A[tri(k, l)];
                        if (!pinned[k])
                            damped[tri(k, k)] += lambda * A[tri(k, k)];
                    }
                    ++it;
                    if (!chol_solve<S, P>(damped, step, (a.numerics & 2u) != 0)) {
                        finish(ST_SINGULAR);
                    } else {
                        S step_norm = S(0), param_norm = S(0);
                        bool finite_trial = true;
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            step_norm = smax(step_norm, sabs(step[k]));
                            trial[k] = clampf(float(S(p[k]) + step[k]), a.lower[k], a.upper[k]);
                            finite_trial = finite_trial && finite_f(trial[k]);
                            step[k] = S(trial[k]) - S(p[k]);
                            param_norm = smax(param_norm, S(fabsf(p[k])));
                        }
                        const S stol = S(a.step_tolerance);
                        if (step_norm <= stol * (param_norm + stol) / smax(S(1), lambda)) {
                            finish(resolved());
                        } else {
                            predicted = S(0);
#pragma unroll
                            for (int k = 0; k < PA; ++k) {
                                predicted += S(2) * g[k] * step[k];
#pragma unroll
                                for (int l = 0; l < PA; ++l)
                                    predicted -= step[k] * A[k >= l ? tri(k, l) : tri(l, k)] * step[l];
                            }
                            pass = finite_trial;
                        }
                    }
                }
            }
        }
        // 2. Groups without a fit take the next one and check its inputs
        //    (cpu/solver.hpp:204-229); a valid one makes its first pass now.
        const uint32_t next = src.take(fit == kNoFit);
        if (fit == kNoFit && next != kNoFit) {
            fit = next;
            iterating = false;
            src.begin(fit, pts);
            const FitArgs &b = src.args();
            bool bad = !b.options_ok || !src.grid().x_finite;
#pragma unroll
            for (int k = 0; k < P; ++k) {
                p[k] = b.p0[src.row() * P + k];
                bad = bad || !finite_f(p[k]);
            }
            if (bad) {
                write(p, ST_INVALID, 0, qnan_f());
            } else {
#pragma unroll
                for (int k = 0; k < P; ++k)
                    trial[k] = clampf(p[k], b.lower[k], b.upper[k]);
                pass = true;
            }
        }
        // 3. One pass for every group, over zero points where none is due.
        Points<G, 0, stream_chunk(MODEL)> due = pts;
        due.m = pass ? pts.m : 0;
        S terms[Params<MODEL>::TERMS];
        bool pass_ok;
        float yrange[2];
        fused_pass<MODEL, G, 0, COMP, S, PA>(due, trial, terms, pass_ok, FLAT ? yrange : nullptr);
        // 4. After the pass (no early exits: every path reaches the next trip's
        //    test the same way).
        if (fit != kNoFit) {
            const FitArgs &c = src.args();
            const S pass_chi = pass ? terms[TRI + P] : S(qnan_f());
            if (!iterating) {
                // First pass at the projected initial point (cpu/solver.hpp:230-245).
                if (!finite_s(pass_chi)) {
                    // Invalid either way; a non-finite observation reports the
                    // initial parameters as given, as solve_fit does.
                    int flags = 0;
                    pts.each([&](float, float yj) { flags |= finite_f(yj) ? 0 : 1; });
                    flags = group_or<G>(flags);
#pragma unroll
                    for (int k = 0; k < P; ++k)
                        p[k] = flags ? c.p0[src.row() * P + k] : trial[k];
                    write(p, ST_INVALID, 0, qnan_f());
                } else {
#pragma unroll
                    for (int k = 0; k < P; ++k)
                        p[k] = trial[k];
                    chi = pass_chi;
                    it = 0;
                    if (FLAT && yrange[0] == yrange[1] && src.grid().varied_x && c.free_shape) {
                        finish(ST_DEGENERATE);
                    } else {
#pragma unroll
                        for (int t = 0; t < TRI; ++t)
                            A[t] = terms[t];
#pragma unroll
                        for (int k = 0; k < P; ++k)
                            g[k] = terms[TRI + k];
                        lin_ok = pass_ok;
                        lambda = S(c.initial_lambda);
                        iterating = true;
                    }
                }
            } else {
                // Accept or reject the trial (cpu/solver.hpp:313-339).
                const S lambda_cap = smax(S(c.initial_lambda), S(1.0f / eps));
                const S reduction = chi - pass_chi;
                const S threshold = S(c.cost_tolerance) * chi;
                const bool small_change = finite_s(pass_chi) && finite_s(predicted) &&
                                          predicted >= S(0) && sabs(reduction) <= threshold &&
                                          predicted <= threshold / smax(S(1), lambda);
                if (finite_s(pass_chi) && reduction > S(0) && finite_s(predicted) &&
                    predicted > S(0)) {
#pragma unroll
                    for (int k = 0; k < P; ++k)
                        p[k] = trial[k];
                    chi = pass_chi;
#pragma unroll
                    for (int t = 0; t < TRI; ++t)
                        A[t] = terms[t];
#pragma unroll
                    for (int k = 0; k < P; ++k)
                        g[k] = terms[TRI + k];
                    lin_ok = pass_ok;
                    const S rho = reduction / predicted;
                    const S q = S(2) * rho - S(1);
                    const S factor =
                        rho >= S(1) ? S(1) / S(3) : smax(S(1) / S(3), S(1) - q * q * q);
                    const S nl = lambda * factor;
                    lambda = nl < S(eps) ? S(eps) : (lambda_cap < nl ? lambda_cap : nl);
                } else {
                    lambda = smin(lambda_cap, lambda * S(10));
                }
                if (small_change)
                    finish(resolved());
            }
        }
    }
}

// The work queue of a retiring kernel: a wave refills its pool with
// pool_refill(G) consecutive task indices from a counter when it runs dry (one
// atomic per refill, by lane 0) and serves its groups in lane order. Tasks
// run from 0 to total.
constexpr uint32_t pool_refill(int g) { return 2u * uint32_t(64 / g) < 8u ? 8u : 2u * uint32_t(64 / g); }

template <int G> struct PoolQueue {
    static constexpr uint32_t kRefill = pool_refill(G);
    uint32_t *counter;
    uint32_t total;
    uint32_t next = 0, end = 0;
    bool dry = false;
    int lane, gl;
    __device__ PoolQueue(uint32_t *counter_, uint32_t total_, int lane_, int gl_)
        : counter(counter_), total(total_), lane(lane_), gl(gl_) {}
    __device__ uint32_t take(bool want) {
        const uint64_t leaders = __ballot(want && gl == 0);
        if (!leaders)
            return kNoFit;
        const uint32_t rank = uint32_t(__popcll(leaders & ((1ull << (lane - gl)) - 1ull)));
        const uint32_t asked = uint32_t(__popcll(leaders));
        // What is left of the pool goes to the first requests, a refill
        // (kRefill ≥ 64/G, so one is enough) to the rest.
        const uint32_t left_begin = next, left = end - next;
        uint32_t fresh = 0;
        if (asked > left && !dry) {
            uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(counter, kRefill);
            base = __builtin_amdgcn_readfirstlane(base);
            dry = base >= total;
            next = dry ? total : base;
            end = dry ? total : (total - base < kRefill ? total : base + kRefill);
            fresh = end - next;
        } else {
            next += asked < left ? asked : left;
        }
        uint32_t got = kNoFit;
        if (rank < left)
            got = left_begin + rank;
        else if (rank - left < fresh)
            got = next + (rank - left);
        if (fresh)
            next += asked - left < fresh ? asked - left : fresh;
        return want ? got : kNoFit;
    }
};

// The fits of a uniform batch as a source for solve_pool.
template <int G> struct BatchPool : PoolQueue<G> {
    const FitArgs &a;
    GridInfo info;
    uint32_t cur = kNoFit;
    __device__ BatchPool(const FitArgs &args, int lane_, int gl_)
        : PoolQueue<G>(args.pool, args.n, lane_, gl_), a(args), info(*args.grid) {}
    template <typename Pts> __device__ void begin(uint32_t fit, Pts &pts) {
        cur = fit;
        pts.load(a.x, a.y + fit * a.ys_fit, a.ys_pt, this->gl, int(a.m));
    }
    __device__ const FitArgs &args() const { return a; }
    __device__ const GridInfo &grid() const { return info; }
    __device__ size_t row() const { return cur; }
};

// Occupancy floor per instantiation: three waves per SIMD (≤ 168 VGPRs) for
// 64-lane tiles of up to 16 points and for the streamed passes of groups of
// the single-peak models; otherwise unconstrained (the two-peak model's state
// takes a wave per SIMD).
constexpr int lm_min_waves(int model, int g, int r) {
    if (model == TWO_LORENTZIAN)
        return 1;
    if (g == 64 && r > 0 && r <= 16)
        return 3;
    return (g > 1 && r == 0) ? 3 : 1;
}

// Streamed groups with retiring fits (no uncertainties): persistent waves over
// the batch counter a.pool. A two-peak batch with a fixed baseline launches
// the FIXB form and then the full form with wide_grid_only set on the same
// counter: each reads the grid facts first, and exactly one of them runs
// (the FIXB form on grids within kFixedBaselineMaxX).
template <int MODEL, int G, bool COMP, bool MIXED, bool FLAT, bool FIXB = false>
__global__ void __launch_bounds__(256, lm_min_waves(MODEL, G, 0)) lm_pool_kernel(FitArgs a) {
    using S = typename std::conditional<MIXED, double, float>::type;
    const bool narrow = a.grid->max_abs_x <= kFixedBaselineMaxX;
    if (FIXB ? !narrow : (a.wide_grid_only && narrow))
        return;  // the whole launch: the other form serves this grid
    const int lane = int(threadIdx.x & 63);
    BatchPool<G> src(a, lane, lane % G);
    solve_pool<MODEL, G, COMP, FLAT, S, FIXB>(lane % G, src);
}

// Uncertainties after a retiring launch: one group per fit, from the fit's
// written parameters, status and χ² (the values solve_fit hands to
// covariance), over the streamed points of the same group size.
// FIXB: the fixed-baseline form, launched before the full one with
// wide_grid_only set; exactly one of them runs, by the grid's largest |x|.
template <int MODEL, int G, bool MIXED, bool FIXB = false>
__global__ void __launch_bounds__(256) lm_cov_kernel(FitArgs a) {
    using S = typename std::conditional<MIXED, double, float>::type;
    constexpr int P = Params<MODEL>::P;
    const uint32_t fit = blockIdx.x * (256 / G) + threadIdx.x / G;
    const int gl = int(threadIdx.x % G);
    const bool narrow = a.grid->max_abs_x <= kFixedBaselineMaxX;
    if (fit >= a.n || (FIXB ? !narrow : (a.wide_grid_only && narrow)))
        return;
    float p[P];
#pragma unroll
    for (int k = 0; k < P; ++k)
        p[k] = a.params[size_t(fit) * P + k];
    Points<G, 0, stream_chunk(MODEL)> pts;
    pts.load(a.x, a.y + fit * a.ys_fit, a.ys_pt, gl, int(a.m));
    __shared__ S cl[256 / G][P * P];
    covariance<MODEL, G, 0, S, FIXB, true>(a, pts, fit, gl, p, a.status[fit], a.chi2[fit],
                                           cl[threadIdx.x / G]);
}

// Workgroup per fit (G = 128 or 256 lanes, 2 or 4 waves): every workgroup
// solves its fits one after another (a grid-stride loop over the batch, so
// the grid never exceeds the launch limits), with a register tile (x staged
// once in LDS) or streamed points. All lanes of the workgroup follow the
// same control path, so the barriers of the cross-wave reductions are
// reached together.
template <int MODEL, int G, int R, bool COMP, bool MIXED>
__global__ void __launch_bounds__(G) lm_block_kernel(FitArgs a) {
    static_assert(G == 128 || G == 256, "workgroup mappings are 2 or 4 waves");
    using S = typename std::conditional<MIXED, double, float>::type;
    if constexpr (R > 0) {
        __shared__ float xs[G * R];
        for (int j = threadIdx.x; j < G * R; j += G)
            xs[j] = j < int(a.m) ? a.x[j] : 0.0f;
        __syncthreads();
        for (uint32_t fit = blockIdx.x; fit < a.n; fit += gridDim.x)
            solve_fit<MODEL, G, R, COMP, S>(a, fit, int(threadIdx.x), xs);
    } else {
        for (uint32_t fit = blockIdx.x; fit < a.n; fit += gridDim.x)
            solve_fit<MODEL, G, R, COMP, S>(a, fit, int(threadIdx.x), a.x);
    }
}

template <int MODEL, int G, int R, bool COMP, bool MIXED>
__global__ void __launch_bounds__(256, lm_min_waves(MODEL, G, R)) lm_kernel(FitArgs a) {
    constexpr int per_block = 256 / G;
    using S = typename std::conditional<MIXED, double, float>::type;
    const uint32_t fit = blockIdx.x * per_block + threadIdx.x / G;
    if constexpr (R > 0) {
        __shared__ float xs[G * R];
        for (int j = threadIdx.x; j < G * R; j += 256)
            xs[j] = j < int(a.m) ? a.x[j] : 0.0f;
        __syncthreads();
        if (fit < a.n)
            solve_fit<MODEL, G, R, COMP, S>(a, fit, int(threadIdx.x % G), xs);
    } else {
        if (fit < a.n)
            solve_fit<MODEL, G, R, COMP, S>(a, fit, int(threadIdx.x % G), a.x);
    }
}

}  // namespace bfhip
