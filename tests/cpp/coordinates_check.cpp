// Host checks of batchfit/coordinates.hpp: the reference's known covariance
// values (tests/uncertainty_checks.hpp:76-86 of the reference) and a
// normalise → physical round trip in a frame far from unit scale.
#include <batchfit/coordinates.hpp>

#include <cmath>
#include <cstdio>
#include <vector>

int main() {
    const batchfit::Coordinates frame{5, 2, 3};
    std::vector<double> c(64, 0.0);
    for (int i = 0; i < 8; ++i)
        c[i * 8 + i] = 1;
    c[6 * 8 + 7] = c[7 * 8 + 6] = .4;
    const auto out = frame.physical_covariance(batchfit::Model::two_lorentzian, c);
    if (std::abs(out[6 * 8 + 6] - 47.25) > 1e-12 || std::abs(out[6 * 8 + 7] + 9.45) > 1e-12 ||
        std::abs(out[7 * 8 + 7] - 2.25) > 1e-12)
        return std::printf("covariance %g %g %g\n", out[54], out[55], out[63]), 1;
    std::vector<double> x(129), y(129);
    for (int j = 0; j < 129; ++j) {
        x[j] = 1e6 + .01 * (-1 + j / 64.0);
        y[j] = 1e-6 * std::exp(-j / 40.0);
    }
    const auto f = batchfit::Coordinates::from(x, y);
    if (std::abs(f.x(x.front()) + 1) > 1e-7 || std::abs(f.x(x.back()) - 1) > 1e-7)
        return std::puts("grid endpoints"), 1;
    const std::vector<double> p = {2e-6, 1e6 - 4e-3, 1.2e-3, 1.5e-6, 1e6 + 4e-3, 1.5e-3, -2, 2e-6};
    const auto back = f.physical(batchfit::Model::two_lorentzian,
                                 f.normalize(batchfit::Model::two_lorentzian, p));
    for (int k = 0; k < 8; ++k)
        if (std::abs(back[k] - p[k]) > 1e-12 * std::abs(p[k]) + 1e-14)
            return std::printf("round trip %d: %.17g vs %.17g\n", k, back[k], p[k]), 1;
    std::puts("coordinates ok");
    return 0;
}
