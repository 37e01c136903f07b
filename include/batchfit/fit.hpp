// batchfit/fit.hpp — C++ façade of the MI355X fitter over the C ABI in
// batchfit_hip.h (header-only; link with -lbatchfit_hip).
//
// batchfit::fit has the semantics of the reference's native entry
// _core.fit(x, y, p0, lower, upper, model, device, uncertainties)
// (python/bindings.cpp:123-198): shared x[M], fit-major y[N*M], optional
// p0[N*P], bounds [P]; it returns the BatchResult of cpu/batch.hpp:11-19 with
// float values, int32 statuses and uint32 iterations. Argument errors throw
// std::invalid_argument and device failures std::runtime_error
// (python/bindings.cpp:88-89, 131-160); per-fit numerical problems are
// statuses. batchfit::submit enqueues on a caller stream and returns at once,
// like batchfit::submit_batch (include/batchfit/metal/batch.hpp:253-264).
#pragma once

#include <batchfit_hip.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace batchfit {

enum class Model : uint32_t {
    gaussian = BF_GAUSSIAN,
    lorentzian = BF_LORENTZIAN,
    pseudo_voigt = BF_PSEUDO_VOIGT,
    two_lorentzian = BF_TWO_LORENTZIAN
};

inline constexpr unsigned parameter_count(Model model) {
    return model == Model::pseudo_voigt ? 4 : model == Model::two_lorentzian ? 8 : 3;
}

inline Model model_from_name(std::string_view name) {
    if (name == "gaussian")
        return Model::gaussian;
    if (name == "lorentzian")
        return Model::lorentzian;
    if (name == "pseudo_voigt")
        return Model::pseudo_voigt;
    if (name == "two_lorentzian")
        return Model::two_lorentzian;
    throw std::invalid_argument("unknown fit model");
}

// cpu/solver.hpp:26-41 (float defaults) plus the HIP numerics bits and mapping.
struct Options {
    unsigned max_iterations = 100;
    float initial_lambda = 1e-3f;
    float gradient_tolerance = 1e-5f, step_tolerance = 1e-6f, cost_tolerance = 1e-6f;
    std::array<float, 8> lower, upper;
    uint32_t numerics = BF_COMPENSATED | BF_MIXED_FP64;
    bf_launch launch{};
    Options() {
        lower.fill(-std::numeric_limits<float>::infinity());
        upper.fill(std::numeric_limits<float>::infinity());
    }
};

// cpu/batch.hpp:11-19.
struct BatchResult {
    std::vector<float> parameters;
    std::vector<int32_t> status;
    std::vector<uint32_t> iterations;
    std::vector<float> chi_squared;
    std::vector<float> covariance{}, standard_errors{};
    std::vector<int32_t> uncertainty_status{};
};

namespace detail {
inline void raise(int rc, const char *err) {
    if (rc == BF_OK)
        return;
    if (rc == BF_INVALID_ARGUMENT)
        throw std::invalid_argument(err);
    throw std::runtime_error(err);
}
inline bf_config config(Model model, std::size_t n, std::size_t m, const Options &o) {
    if (m == 0)
        throw std::invalid_argument("expected x[M] and y[N,M] with M>0");
    if (n > std::numeric_limits<uint32_t>::max() || m > std::numeric_limits<uint32_t>::max())
        throw std::invalid_argument("N and M must fit in 32 bits");
    bf_config c;
    bf_hip_default_config(&c);
    c.fits = uint32_t(n);
    c.points = uint32_t(m);
    c.max_iterations = o.max_iterations;
    c.initial_lambda = o.initial_lambda;
    c.gradient_tolerance = o.gradient_tolerance;
    c.step_tolerance = o.step_tolerance;
    c.cost_tolerance = o.cost_tolerance;
    c.numerics = o.numerics;
    for (unsigned k = 0; k < parameter_count(model); ++k) {
        // python/bindings.cpp:153-161: invalid bounds are argument errors.
        if (std::isnan(o.lower[k]) || std::isnan(o.upper[k]) || o.lower[k] > o.upper[k] ||
            o.lower[k] == std::numeric_limits<float>::infinity() ||
            o.upper[k] == -std::numeric_limits<float>::infinity())
            throw std::invalid_argument("invalid parameter bounds");
        c.lower[k] = o.lower[k];
        c.upper[k] = o.upper[k];
    }
    return c;
}
}  // namespace detail

inline bool hip_available() { return bf_hip_available() != 0; }

// Synchronous fit of host arrays. An empty p0 runs the device's data-only guess.
inline BatchResult fit(std::span<const float> x, std::span<const float> y, Model model,
                       std::span<const float> p0 = {}, const Options &options = {},
                       bool uncertainties = false) {
    const std::size_t m = x.size(), p = parameter_count(model);
    if (m == 0 || y.size() % m != 0)
        throw std::invalid_argument("expected x[M] and y[N,M] with M>0");
    const std::size_t n = y.size() / m;
    if (!p0.empty() && (p0.size() % p != 0 || p0.size() / p != n))
        throw std::invalid_argument("initial parameters must have shape [N,P]");
    const bf_config cfg = detail::config(model, n, m, options);
    BatchResult r{std::vector<float>(n * p), std::vector<int32_t>(n), std::vector<uint32_t>(n),
                  std::vector<float>(n)};
    if (uncertainties) {
        r.covariance.resize(n * p * p);
        r.standard_errors.resize(n * p);
        r.uncertainty_status.resize(n);
    }
    char err[512] = {};
    detail::raise(bf_hip_fit(uint32_t(model), &cfg, &options.launch, x.data(), y.data(),
                             p0.empty() ? nullptr : p0.data(), r.parameters.data(), r.status.data(),
                             r.iterations.data(), r.chi_squared.data(),
                             uncertainties ? r.covariance.data() : nullptr,
                             uncertainties ? r.standard_errors.data() : nullptr,
                             uncertainties ? r.uncertainty_status.data() : nullptr, err, sizeof err),
                  err);
    return r;
}

inline BatchResult fit(std::span<const float> x, std::span<const float> y, std::string_view model,
                       std::span<const float> p0 = {}, const Options &options = {},
                       bool uncertainties = false) {
    return fit(x, y, model_from_name(model), p0, options, uncertainties);
}

// Device-resident arrays; outputs are device pointers. Enqueues on
// options.launch.stream (or the library stream) and returns immediately.
struct DeviceBatch {
    const float *x, *y, *p0 = nullptr;
    float *parameters;
    int32_t *status;
    uint32_t *iterations;
    float *chi_squared;
    float *covariance = nullptr, *standard_errors = nullptr;
    int32_t *uncertainty_status = nullptr;
    std::size_t fits, points;
};

inline void submit(const DeviceBatch &b, Model model, const Options &options = {}) {
    const bf_config cfg = detail::config(model, b.fits, b.points, options);
    char err[512] = {};
    detail::raise(bf_hip_fit_device(uint32_t(model), &cfg, &options.launch, b.x, b.y, b.p0,
                                    b.parameters, b.status, b.iterations, b.chi_squared,
                                    b.covariance, b.standard_errors, b.uncertainty_status, err,
                                    sizeof err),
                  err);
}

}  // namespace batchfit
