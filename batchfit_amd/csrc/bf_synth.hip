// bf_synth.hip — seeded synthetic spectra generated directly in HBM.
//
// Random numbers come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): a counter-based
// generator, so every (seed, fit, point) has its own stream and the result
// does not depend on the launch shape. Normals use the Box–Muller transform.
// Parameters of fit i are drawn from uniform ranges per parameter; the
// "doublet" layout draws one HWHM for both Lorentzians and places the second
// centre at c1 + U[sep_lo, sep_hi]·HWHM.
#include "bf_host.hpp"

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
    uint32_t noise_relative;  // 1: σ = noise · parameter 0 (amplitude); 0: σ = noise
    float noise;
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
    const float sigma = a.noise_relative ? a.noise * p[0] : a.noise;
    if (a.truth && threadIdx.x < 8)
        a.truth[size_t(i) * 8 + threadIdx.x] = p[threadIdx.x];
    float *row = a.y + size_t(i) * a.m;
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
                                 float sep_lo, float sep_hi, uint32_t noise_relative, float noise,
                                 uint64_t seed, uint64_t fit_offset, float *d_y, float *d_truth,
                                 void *stream, char *err, size_t errlen) {
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP is unavailable: " + rt.why);
    if (model > 3 || m == 0 || !d_x || !d_y || !lower || !upper || layout > 1)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "invalid synthesis arguments");
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
    a.noise_relative = noise_relative;
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
