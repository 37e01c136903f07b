// batchfit/coordinates.hpp — float64 coordinate normalisation on the host
// (the reference's include/batchfit/coordinates.hpp): map x to [-1, 1] about
// the grid's midpoint and y by its largest magnitude before converting to
// float32, fit, then map parameters, χ² and covariance back. Bounds and p0 of
// the fit are in the normalised units. The same maps as
// batchfit_amd/coordinates.py, for the model enum of batchfit/fit.hpp.
#pragma once

#include <batchfit/fit.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <span>
#include <stdexcept>
#include <vector>

namespace batchfit {

struct Coordinates {
    double origin = 0, x_scale = 1, y_scale = 1;

    // The frame of one spectrum: origin = midpoint of the x range, x_scale =
    // half of it, y_scale = the largest |y| (1 when y is all zero).
    static Coordinates from(std::span<const double> x, std::span<const double> y) {
        if (x.empty() || x.size() != y.size())
            throw std::invalid_argument("normalization requires matching nonempty x/y");
        double lo = x[0], hi = x[0], peak = 0;
        for (std::size_t j = 0; j < x.size(); ++j) {
            if (!std::isfinite(x[j]) || !std::isfinite(y[j]))
                throw std::invalid_argument("normalization requires finite inputs");
            lo = std::min(lo, x[j]);
            hi = std::max(hi, x[j]);
            peak = std::max(peak, std::abs(y[j]));
        }
        const double half = hi / 2 - lo / 2;
        if (!(half > 0) || !std::isfinite(half))
            throw std::invalid_argument("normalization requires a resolved x range");
        return {std::midpoint(lo, hi), half, peak > 0 ? peak : 1.0};
    }

    double x(double value) const { return (value - origin) / x_scale; }
    double y(double value) const { return value / y_scale; }
    double physical_chi_squared(double value) const { return value * y_scale * y_scale; }

    // Physical → normalised parameters of the same curve (the two-peak
    // model's baseline b + s·x becomes b' + s'·u).
    std::vector<double> normalize(Model model, std::span<const double> p) const {
        const unsigned n = parameter_count(model);
        if (p.size() != n)
            throw std::invalid_argument("normalization parameter count mismatch");
        std::vector<double> q(p.begin(), p.end());
        q[0] = p[0] / y_scale;
        q[1] = x(p[1]);
        q[2] = p[2] / x_scale;
        if (n == 8) {
            q[3] = p[3] / y_scale;
            q[4] = x(p[4]);
            q[5] = p[5] / x_scale;
            q[6] = std::fma(p[7], origin, p[6]) / y_scale;
            q[7] = p[7] * x_scale / y_scale;
        }
        return q;
    }

    // Normalised → physical parameters (the inverse of normalize).
    std::vector<double> physical(Model model, std::span<const double> q) const {
        const unsigned n = parameter_count(model);
        if (q.size() != n)
            throw std::invalid_argument("physical parameter count mismatch");
        std::vector<double> p(q.begin(), q.end());
        p[0] = q[0] * y_scale;
        p[1] = std::fma(q[1], x_scale, origin);
        p[2] = q[2] * x_scale;
        if (n == 8) {
            p[3] = q[3] * y_scale;
            p[4] = std::fma(q[4], x_scale, origin);
            p[5] = q[5] * x_scale;
            p[7] = q[7] * y_scale / x_scale;
            p[6] = std::fma(-p[7], origin, q[6] * y_scale);
        }
        return p;
    }

    // Covariance of the normalised parameters → physical units: T·C·Tᵀ with
    // T = ∂physical/∂normalised (row-major P×P in and out).
    std::vector<double> physical_covariance(Model model, std::span<const double> c) const {
        const unsigned n = parameter_count(model);
        if (c.size() != std::size_t(n) * n)
            throw std::invalid_argument("covariance shape mismatch");
        std::array<double, 64> t{};
        t[0] = y_scale;
        t[1 * n + 1] = t[2 * n + 2] = x_scale;
        if (n == 4)
            t[3 * n + 3] = 1.0;
        if (n == 8) {
            t[3 * n + 3] = t[6 * n + 6] = y_scale;
            t[4 * n + 4] = t[5 * n + 5] = x_scale;
            t[6 * n + 7] = -origin * y_scale / x_scale;
            t[7 * n + 7] = y_scale / x_scale;
        }
        std::vector<double> tc(std::size_t(n) * n, 0.0), out(std::size_t(n) * n, 0.0);
        for (unsigned i = 0; i < n; ++i)
            for (unsigned k = 0; k < n; ++k)
                for (unsigned l = 0; l < n; ++l)
                    tc[i * n + l] += t[i * n + k] * c[k * n + l];
        for (unsigned i = 0; i < n; ++i)
            for (unsigned j = 0; j < n; ++j)
                for (unsigned l = 0; l < n; ++l)
                    out[i * n + j] += tc[i * n + l] * t[j * n + l];
        return out;
    }
};

}  // namespace batchfit
