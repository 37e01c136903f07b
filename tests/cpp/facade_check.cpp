// Compile-and-run check of the C++ façade (include/batchfit/fit.hpp): argument
// errors throw std::invalid_argument before any device work; with a gfx950
// device present, a small Lorentzian batch converges to the generating values.
#include <batchfit/fit.hpp>

#include <cmath>
#include <cstdio>
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
    std::puts("facade ok (device)");
    return 0;
}
