// Compile-and-run check of the C++ façade (include/batchfit/fit.hpp): argument
// errors throw std::invalid_argument before any device work; with a gfx950
// device present, a small Lorentzian batch converges to the generating values,
// and batchfit::submit on a stream of the caller's (include/batchfit/metal/
// batch.hpp:253-264 submit_batch) gives the synchronous fit's bytes once that
// stream is synchronised. Built by tests/test_facade.py with the host compiler
// against the HIP runtime's C API (hipStreamCreate / hipStreamSynchronize).
#include <batchfit/fit.hpp>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

int main(int argc, char **argv) {
    const bool need_device = argc > 1;
    std::vector<float> x(129), y(4 * 129);
    for (int j = 0; j < 129; ++j) {
        x[j] = -4.0f + 8.0f * float(j) / 128.0f;
        for (int i = 0; i < 4; ++i) {
            const float z = (x[j] - 0.2f) / 0.7f;
            y[i * 129 + j] = 2.0f / (1.0f + z * z);
        }
    }
    bool threw = false;
    try {
        batchfit::fit(x, y, "missing");
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    if (!threw)
        return std::puts("unknown model not rejected"), 1;
    threw = false;
    batchfit::Options bad;
    bad.lower[0] = 2;
    bad.upper[0] = 1;
    try {
        batchfit::fit(x, y, batchfit::Model::lorentzian, {}, bad);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    if (!threw)
        return std::puts("invalid bounds not rejected"), 1;
    if (!batchfit::hip_available()) {
        std::puts(need_device ? "no device" : "facade ok (no device)");
        return need_device ? 1 : 0;
    }
    const std::vector<float> p0 = {1.8f, 0.0f, 0.8f};
    std::vector<float> init;
    for (int i = 0; i < 4; ++i)
        init.insert(init.end(), p0.begin(), p0.end());
    const auto r = batchfit::fit(x, y, batchfit::Model::lorentzian, init, {}, true);
    for (int i = 0; i < 4; ++i)
        if (r.status[i] != 0 || std::fabs(r.parameters[i * 3] - 2.0f) > 2e-4f ||
            std::fabs(r.parameters[i * 3 + 1] - 0.2f) > 2e-4f || r.uncertainty_status[i] != 0)
            return std::printf("fit %d: status %d a=%g\n", i, r.status[i], r.parameters[i * 3]), 1;
    // Asynchronous submit on the caller's stream, from the device guess.
    const std::size_t n = 4, m = 129, p = 3;
    void *dx = nullptr, *dy = nullptr, *dout = nullptr;
    char err[512];
    if (bf_hip_malloc(&dx, m * 4, err, sizeof err) || bf_hip_malloc(&dy, n * m * 4, err, sizeof err) ||
        bf_hip_malloc(&dout, n * (p + 3) * 4, err, sizeof err))
        return std::printf("bf_hip_malloc: %s\n", err), 1;
    if (bf_hip_memcpy(dx, x.data(), m * 4, 1, err, sizeof err) ||
        bf_hip_memcpy(dy, y.data(), n * m * 4, 1, err, sizeof err))
        return std::printf("bf_hip_memcpy: %s\n", err), 1;
    hipStream_t stream = nullptr;
    if (hipStreamCreate(&stream) != hipSuccess)
        return std::puts("hipStreamCreate failed"), 1;
    float *out = static_cast<float *>(dout);
    batchfit::DeviceBatch batch{static_cast<const float *>(dx), static_cast<const float *>(dy),
                                nullptr, out, reinterpret_cast<int32_t *>(out + n * p),
                                reinterpret_cast<uint32_t *>(out + n * p + n), out + n * p + 2 * n};
    batch.fits = n;
    batch.points = m;
    batchfit::Options opt;
    opt.launch.stream = stream;
    batchfit::submit(batch, batchfit::Model::lorentzian, opt);
    if (hipStreamSynchronize(stream) != hipSuccess)
        return std::puts("hipStreamSynchronize failed"), 1;
    std::vector<float> got(n * (p + 3));
    if (bf_hip_memcpy(got.data(), dout, got.size() * 4, 2, err, sizeof err))
        return std::printf("bf_hip_memcpy: %s\n", err), 1;
    const auto sync = batchfit::fit(x, y, batchfit::Model::lorentzian);
    if (std::memcmp(got.data(), sync.parameters.data(), n * p * 4) ||
        std::memcmp(got.data() + n * p, sync.status.data(), n * 4) ||
        std::memcmp(got.data() + n * p + n, sync.iterations.data(), n * 4) ||
        std::memcmp(got.data() + n * p + 2 * n, sync.chi_squared.data(), n * 4))
        return std::puts("submit on the caller's stream differs from fit"), 1;
    (void)hipStreamDestroy(stream);
    bf_hip_free(dx);
    bf_hip_free(dy);
    bf_hip_free(dout);
    std::puts("facade ok (device)");
    return 0;
}
