// bf_synth.hip — seeded synthetic spectra generated directly in HBM.
//
// Random numbers come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): a counter-based
// generator, so every (seed, fit, point) has its own stream and the result
// does not depend on the launch shape. Normals use the Box–Muller transform.
// Parameters of fit i are drawn from uniform ranges per parameter; the
// "doublet" layout draws one HWHM for both Lorentzians and places the second
// centre at c1 + U[sep_lo, sep_hi]·HWHM.
//
// Noise modes: a fixed σ, σ relative to the amplitude, and the two modes of
// the reference's generator (src/cpu/synthetic.cpp:83-108): Gaussian noise at
// a signal-to-noise ratio (σ = RMS of the clean row / snr) and Poisson counts
// (gain = snr²·Σf/Σf², y = Poisson(gain·f) / gain). Poisson counts use
// inversion by sequential search below a mean of 10 and, from 10, the
// transformed rejection method PTRS of W. Hörmann, "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12 (1993) 39-45.
#include "bf_host.hpp"

#include <cmath>

namespace bfhip {
namespace {

struct Philox {
    uint32_t c[4];
};

__device__ __forceinline__ void philox_round(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
    const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
    const uint32_t hi0 = __umulhi(m0, c[0]), lo0 = m0 * c[0];
    const uint32_t hi1 = __umulhi(m1, c[2]), lo1 = m1 * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

// Four 32-bit random words for counter (a, b, c, d) under a 64-bit key.
__device__ __forceinline__ void philox4x32_10(uint32_t a, uint32_t b, uint32_t cc, uint32_t d,
                                              uint64_t seed, uint32_t (&out)[4]) {
    uint32_t c[4] = {a, b, cc, d};
    uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    for (int i = 0; i < 4; ++i)
        out[i] = c[i];
}

// Uniform in (0, 1]: 24 high bits, offset by half a step (the top step rounds
// to exactly 1 in FP32; never 0, so the logarithm of Box–Muller is finite).
__device__ __forceinline__ float unit(uint32_t w) { return (float(w >> 8) + 0.5f) * (1.0f / 16777216.0f); }

struct SynthArgs {
    uint32_t model, n, m;
    const float *x;     // [m]
    float lo[8], hi[8]; // parameter ranges
    uint32_t layout;    // 0 independent, 1 doublet (two_lorentzian)
    float sep_lo, sep_hi;
    uint32_t noise_mode;      // bf_noise: 0 σ = noise, 1 σ = noise·amplitude, 2 SNR Gaussian, 3 SNR Poisson
    float noise;              // σ, relative σ, or the signal-to-noise ratio (modes 2, 3)
    uint64_t seed;
    uint64_t fit_offset;      // counter offset, so batches can be generated in pieces
    float *y;                 // [n, m]
    float *truth;             // [n, 8] or null
};

__device__ __forceinline__ float model_value_synth(uint32_t model, float x, const float *p) {
    switch (model) {
    case GAUSSIAN: {
        ModelConsts<GAUSSIAN> c;
        c.set(p);
        return c.value(x);
    }
    case LORENTZIAN: {
        ModelConsts<LORENTZIAN> c;
        c.set(p);
        return c.value(x);
    }
    case PSEUDO_VOIGT: {
        ModelConsts<PSEUDO_VOIGT> c;
        c.set(p);
        return c.value(x);
    }
    default: {
        ModelConsts<TWO_LORENTZIAN> c;
        c.set(p);
        return c.value(x);
    }
    }
}

// Uniform in (0, 1) with 53 random bits from two words.
__device__ __forceinline__ double unit53(uint32_t hi, uint32_t lo) {
    return (double((uint64_t(hi >> 5) << 26) | (lo >> 6)) + 0.5) * (1.0 / 9007199254740992.0);
}

// A Poisson count of mean mu ≥ 0 from the Philox stream (fit, point): each
// counter (…, point, 8 + t) gives the two uniforms of try t.
__device__ double poisson_count(double mu, uint64_t fit, uint32_t j, uint64_t seed) {
    if (!(mu > 0.0))
        return 0.0;
    if (mu < 10.0) {  // inversion: the smallest k with F(k) ≥ u
        uint32_t w[4];
        philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), j, 8u, seed, w);
        const double u = unit53(w[0], w[1]);
        double p = exp(-mu), f = p, k = 0.0;
        while (u > f && k < 1000.0) {
            k += 1.0;
            p *= mu / k;
            f += p;
        }
        return k;
    }
    // PTRS (Hörmann 1993): a transformed-rejection proposal from one uniform,
    // accepted in a quick region or by the exact density ratio.
    const double smu = sqrt(mu), lmu = log(mu);
    const double b = 0.931 + 2.53 * smu, a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t t = 0; t < 256; ++t) {
        uint32_t w[4];
        philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), j, 8u + t, seed, w);
        const double u = unit53(w[0], w[1]) - 0.5, v = unit53(w[2], w[3]);
        const double us = 0.5 - fabs(u);
        const double k = floor((2.0 * a / us + b) * u + mu + 0.43);
        if (us >= 0.07 && v <= vr)
            return k;
        if (k < 0.0 || (us < 0.013 && v > us))
            continue;
        if (log(v) + log(inv_alpha) - log(a / (us * us) + b) <= -mu + k * lmu - lgamma(k + 1.0))
            return k;
    }
    return mu;  // not reached in practice (acceptance ≈ 0.9 per try)
}

// Block-wide sum of a double over 256 threads (4 waves of 64).
__device__ double block_sum(double v, double *shared) {
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    const int wave = int(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0)
        shared[wave] = v;
    __syncthreads();
    double total = 0.0;
    for (int w = 0; w < 4; ++w)
        total += shared[w];
    __syncthreads();
    return total;
}

// One 256-thread block per fit.
__global__ void __launch_bounds__(256) synth_kernel(SynthArgs a) {
    const uint32_t i = blockIdx.x;
    if (i >= a.n)
        return;
    const uint64_t fit = a.fit_offset + i;
    uint32_t r0[4], r1[4], r2[4];
    philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), 0xFFFFFFFFu, 0u, a.seed, r0);
    philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), 0xFFFFFFFFu, 1u, a.seed, r1);
    philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), 0xFFFFFFFFu, 2u, a.seed, r2);
    const uint32_t words[12] = {r0[0], r0[1], r0[2], r0[3], r1[0], r1[1],
                                r1[2], r1[3], r2[0], r2[1], r2[2], r2[3]};
    float p[8];
    for (int k = 0; k < 8; ++k)
        p[k] = a.lo[k] + (a.hi[k] - a.lo[k]) * unit(words[k]);
    if (a.layout == 1) {
        p[5] = p[2];
        p[4] = p[1] + (a.sep_lo + (a.sep_hi - a.sep_lo) * unit(words[8])) * p[2];
    }
    if (a.truth && threadIdx.x < 8)
        a.truth[size_t(i) * 8 + threadIdx.x] = p[threadIdx.x];
    float *row = a.y + size_t(i) * a.m;
    if (a.noise_mode >= 2) {
        // Sum and sum of squares of the clean row in FP64 first; the second
        // pass evaluates each clean value again in the thread that writes it.
        __shared__ double partial[4];
        double sum = 0.0, squares = 0.0;
        bool negative = false;
        for (uint32_t j = threadIdx.x; j < a.m; j += 256) {
            const float f = model_value_synth(a.model, a.x[j], p);
            sum += f;
            squares += double(f) * f;
            negative = negative || f < 0.0f;
        }
        sum = block_sum(sum, partial);
        squares = block_sum(squares, partial);
        const bool any_negative = __syncthreads_or(negative);
        const double snr = double(a.noise);
        if (a.noise_mode == 2) {
            const double sigma = sqrt(squares / double(a.m)) / snr;
            for (uint32_t j = 2 * threadIdx.x; j < a.m; j += 512) {
                uint32_t w[4];
                philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), j, 7u, a.seed, w);
                const float u1 = unit(w[0]), u2 = unit(w[1]);
                const float rad = sqrtf(-2.0f * logf(u1));
                const float z0 = rad * cosf(6.283185307179586f * u2);
                const float z1 = rad * sinf(6.283185307179586f * u2);
                row[j] = float(double(model_value_synth(a.model, a.x[j], p)) + sigma * double(z0));
                if (j + 1 < a.m)
                    row[j + 1] = float(double(model_value_synth(a.model, a.x[j + 1], p)) +
                                       sigma * double(z1));
            }
        } else if (!(squares > 0.0)) {  // an all-zero signal has zero counts
            for (uint32_t j = threadIdx.x; j < a.m; j += 256)
                row[j] = 0.0f;
        } else {
            // A negative clean value has no Poisson count (the reference
            // raises); such a row is written as NaN, as is a point whose
            // expected count exceeds 1e12.
            const double gain = snr * snr * sum / squares;
            for (uint32_t j = threadIdx.x; j < a.m; j += 256) {
                const double mean = gain * double(model_value_synth(a.model, a.x[j], p));
                if (any_negative || !(gain > 0.0) || !(mean <= 1e12)) {
                    row[j] = __uint_as_float(0x7fc00000u);
                    continue;
                }
                row[j] = float(poisson_count(mean, fit, j, a.seed) / gain);
            }
        }
        return;
    }
    const float sigma = a.noise_mode == 1 ? a.noise * p[0] : a.noise;
    for (uint32_t j = 2 * threadIdx.x; j < a.m; j += 512) {
        uint32_t w[4];
        philox4x32_10(uint32_t(fit), uint32_t(fit >> 32), j, 7u, a.seed, w);
        // Box–Muller: two normals from two uniforms.
        const float u1 = unit(w[0]), u2 = unit(w[1]);
        const float rad = sqrtf(-2.0f * logf(u1));
        const float z0 = rad * cosf(6.283185307179586f * u2);
        const float z1 = rad * sinf(6.283185307179586f * u2);
        row[j] = model_value_synth(a.model, a.x[j], p) + sigma * z0;
        if (j + 1 < a.m)
            row[j + 1] = model_value_synth(a.model, a.x[j + 1], p) + sigma * z1;
    }
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

extern "C" int bf_hip_synthesize(uint32_t model, uint32_t n, uint32_t m, const float *d_x,
                                 const float *lower, const float *upper, uint32_t layout,
                                 float sep_lo, float sep_hi, uint32_t noise_mode, float noise,
                                 uint64_t seed, uint64_t fit_offset, float *d_y, float *d_truth,
                                 void *stream, char *err, size_t errlen) {
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP is unavailable: " + rt.why);
    if (model > 3 || m == 0 || !d_x || !d_y || !lower || !upper || layout > 1 || noise_mode > 3)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "invalid synthesis arguments");
    if (noise_mode >= 2 && !(std::isfinite(noise) && noise > 0.0f))
        return fail(err, errlen, BF_INVALID_ARGUMENT, "snr must be finite and positive");
    if (n == 0)
        return BF_OK;
    SynthArgs a{};
    a.model = model;
    a.n = n;
    a.m = m;
    a.x = d_x;
    for (int k = 0; k < 8; ++k) {
        a.lo[k] = lower[k];
        a.hi[k] = upper[k];
    }
    a.layout = layout;
    a.sep_lo = sep_lo;
    a.sep_hi = sep_hi;
    a.noise_mode = noise_mode;
    a.noise = noise;
    a.seed = seed;
    a.fit_offset = fit_offset;
    a.y = d_y;
    a.truth = d_truth;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : rt.stream;
    hipLaunchKernelGGL(synth_kernel, dim3(n), dim3(256), 0, s, a);
    BF_TRY(hipGetLastError(), "synthesis launch");
    return BF_OK;
}
