// ref_driver.cpp — exports the "ref_" half of oracle_api.h by calling the
// reference's own CPU headers where they lie (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/Makefile into oracle/_ref/libbfref.so with
//   -I <reference>/include  and  <reference>/src/cpu/thread_pool.cpp.
// The batch path is the reference's cpu::fit_batch over its cpu::ThreadPool
// (cpu/batch.hpp:23-65); a null p0 first runs cpu::initial_guess per row and
// fits from its parameters whatever its status (python/bindings.cpp:69-80).
// Nothing here re-implements an algorithm: it converts between flat C arrays
// and the reference's std::span / std::array types.
#include "oracle_api.h"

#include <batchfit/cpu/batch.hpp>
#include <batchfit/cpu/initial_guess.hpp>
#include <batchfit/cpu/solver.hpp>
#include <batchfit/cpu/thread_pool.hpp>
#include <batchfit/cpu/uncertainty.hpp>
#include <batchfit/models.hpp>

#include <cmath>
#include <memory>
#include <span>
#include <vector>

namespace {
namespace cpu = batchfit::cpu;
namespace models = batchfit::models;

template <typename T, std::size_t P>
cpu::SolverOptions<T, P> options_of(const bf_oracle_options *in, const T *lower, const T *upper) {
    cpu::SolverOptions<T, P> o;
    if (in) {
        o.max_iterations = in->max_iterations;
        if (!(in->defaults & 1u))
            o.initial_lambda = T(in->initial_lambda);
        if (!(in->defaults & 2u))
            o.gradient_tolerance = T(in->gradient_tolerance);
        if (!(in->defaults & 4u))
            o.step_tolerance = T(in->step_tolerance);
        if (!(in->defaults & 8u))
            o.cost_tolerance = T(in->cost_tolerance);
        o.compensated_sum = in->compensated != 0;
        o.scaled_cholesky = in->scaled != 0;
    }
    for (std::size_t k = 0; k < P; ++k) {
        if (lower)
            o.lower[k] = lower[k];
        if (upper)
            o.upper[k] = upper[k];
    }
    return o;
}

// One pool per worker count, kept for the life of the process like the
// binding's static pool (python/bindings.cpp:14-17).
cpu::ThreadPool &pool_for(int threads) {
    static std::vector<std::unique_ptr<cpu::ThreadPool>> pools(1025);
    const int t = threads < 1 ? 1 : threads > 1024 ? 1024 : threads;
    if (!pools[t])
        pools[t] = std::make_unique<cpu::ThreadPool>(std::size_t(t));
    return *pools[t];
}

template <typename T, template <typename> class Model>
int fit_model(const T *x, std::size_t m, const T *y, std::size_t n, const T *p0, const T *lower,
              const T *upper, const bf_oracle_options *opt, int unc, int threads, T *params,
              int32_t *status, uint32_t *iterations, T *chi2, T *cov, T *se, int32_t *us) {
    constexpr std::size_t P = Model<T>::parameter_count;
    auto &pool = pool_for(threads);
    const std::span<const T> xs(x, m), ys(y, n * m);
    std::vector<T> generated;
    std::span<const T> initial;
    if (p0) {
        initial = std::span<const T>(p0, n * P);
    } else {
        generated.resize(n * P);
        pool.parallel_for(n, [&](std::size_t lo, std::size_t hi) {
            for (std::size_t i = lo; i < hi; ++i) {
                const auto g = cpu::initial_guess<T, Model>(xs, ys.subspan(i * m, m));
                std::copy(g.parameters.begin(), g.parameters.end(), generated.data() + i * P);
            }
        });
        initial = generated;
    }
    const auto o = options_of<T, P>(opt, lower, upper);
    const auto r = cpu::fit_batch<T, Model>(xs, ys, initial, pool, o, unc != 0);
    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t k = 0; k < P; ++k)
            params[i * P + k] = r.parameters[i * P + k];
        status[i] = int32_t(r.status[i]);
        iterations[i] = r.iterations[i];
        chi2[i] = r.chi_squared[i];
        if (unc) {
            for (std::size_t k = 0; k < P * P; ++k)
                cov[i * P * P + k] = r.covariance[i * P * P + k];
            for (std::size_t k = 0; k < P; ++k)
                se[i * P + k] = r.standard_errors[i * P + k];
            us[i] = int32_t(r.uncertainty_status[i]);
        }
    }
    return 0;
}

template <typename T>
int fit_any(int model, const T *x, std::size_t m, const T *y, std::size_t n, const T *p0,
            const T *lower, const T *upper, const bf_oracle_options *opt, int unc, int threads,
            T *params, int32_t *status, uint32_t *iterations, T *chi2, T *cov, T *se,
            int32_t *us) {
    switch (model) {
    case 0:
        return fit_model<T, models::Gaussian>(x, m, y, n, p0, lower, upper, opt, unc, threads,
                                              params, status, iterations, chi2, cov, se, us);
    case 1:
        return fit_model<T, models::Lorentzian>(x, m, y, n, p0, lower, upper, opt, unc, threads,
                                                params, status, iterations, chi2, cov, se, us);
    case 2:
        return fit_model<T, models::PseudoVoigt>(x, m, y, n, p0, lower, upper, opt, unc, threads,
                                                 params, status, iterations, chi2, cov, se, us);
    case 3:
        return fit_model<T, models::TwoLorentzian>(x, m, y, n, p0, lower, upper, opt, unc,
                                                   threads, params, status, iterations, chi2, cov,
                                                   se, us);
    }
    return -1;
}

template <typename T, template <typename> class Model>
int guess_model(const T *x, std::size_t m, const T *y, std::size_t n, int threads, T *params,
                T *baseline, int32_t *status) {
    constexpr std::size_t P = Model<T>::parameter_count;
    const std::span<const T> xs(x, m), ys(y, n * m);
    pool_for(threads).parallel_for(n, [&](std::size_t lo, std::size_t hi) {
        for (std::size_t i = lo; i < hi; ++i) {
            const auto g = cpu::initial_guess<T, Model>(xs, ys.subspan(i * m, m));
            for (std::size_t k = 0; k < P; ++k)
                params[i * P + k] = g.parameters[k];
            baseline[i] = g.baseline;
            status[i] = int32_t(g.status);
        }
    });
    return 0;
}

template <typename T>
int guess_any(int model, const T *x, std::size_t m, const T *y, std::size_t n, int threads,
              T *params, T *baseline, int32_t *status) {
    switch (model) {
    case 0:
        return guess_model<T, models::Gaussian>(x, m, y, n, threads, params, baseline, status);
    case 1:
        return guess_model<T, models::Lorentzian>(x, m, y, n, threads, params, baseline, status);
    case 2:
        return guess_model<T, models::PseudoVoigt>(x, m, y, n, threads, params, baseline, status);
    case 3:
        return guess_model<T, models::TwoLorentzian>(x, m, y, n, threads, params, baseline,
                                                     status);
    }
    return -1;
}

template <typename T, template <typename> class Model>
int model_one(T x, const T *p, T *value, T *jac) {
    *value = Model<T>::eval(x, p);
    Model<T>::jacobian(x, p, jac);
    return 0;
}

template <typename T> int model_any(int model, T x, const T *p, T *value, T *jac) {
    switch (model) {
    case 0:
        return model_one<T, models::Gaussian>(x, p, value, jac);
    case 1:
        return model_one<T, models::Lorentzian>(x, p, value, jac);
    case 2:
        return model_one<T, models::PseudoVoigt>(x, p, value, jac);
    case 3:
        return model_one<T, models::TwoLorentzian>(x, p, value, jac);
    }
    return -1;
}

template <typename T, std::size_t N> int cholesky_n(T *a, T *b, bool scaled) {
    cpu::detail::Matrix<T, N> matrix;
    std::array<T, N> rhs;
    for (std::size_t i = 0; i < N; ++i) {
        rhs[i] = b[i];
        for (std::size_t j = 0; j < N; ++j)
            matrix[i][j] = a[i * N + j];
    }
    const bool ok = cpu::detail::cholesky_solve<T, N>(matrix, rhs, scaled);
    for (std::size_t i = 0; i < N; ++i) {
        b[i] = rhs[i];
        for (std::size_t j = 0; j < N; ++j)
            a[i * N + j] = matrix[i][j];
    }
    return ok ? 1 : 0;
}

template <typename T> int cholesky_any(T *a, T *b, int n, int scaled) {
    switch (n) {
    case 1: return cholesky_n<T, 1>(a, b, scaled);
    case 2: return cholesky_n<T, 2>(a, b, scaled);
    case 3: return cholesky_n<T, 3>(a, b, scaled);
    case 4: return cholesky_n<T, 4>(a, b, scaled);
    case 5: return cholesky_n<T, 5>(a, b, scaled);
    case 6: return cholesky_n<T, 6>(a, b, scaled);
    case 7: return cholesky_n<T, 7>(a, b, scaled);
    case 8: return cholesky_n<T, 8>(a, b, scaled);
    }
    return -1;
}

template <typename T, template <typename> class Model>
int unc_model(const T *x, std::size_t m, const T *params, int32_t status, T chi2, const T *lower,
              const T *upper, T *cov, T *se, int32_t *us) {
    constexpr std::size_t P = Model<T>::parameter_count;
    cpu::FitResult<T, P> fit;
    for (std::size_t k = 0; k < P; ++k)
        fit.parameters[k] = params[k];
    fit.status = batchfit::FitStatus(status);
    fit.chi_squared = chi2;
    const auto o = options_of<T, P>(nullptr, lower, upper);
    const auto u = cpu::estimate_uncertainty<T, Model>(std::span<const T>(x, m), fit, o);
    for (std::size_t k = 0; k < P * P; ++k)
        cov[k] = u.covariance[k];
    for (std::size_t k = 0; k < P; ++k)
        se[k] = u.standard_errors[k];
    *us = int32_t(u.status);
    return 0;
}

template <typename T>
int unc_any(int model, const T *x, std::size_t m, const T *params, int32_t status, T chi2,
            const T *lower, const T *upper, T *cov, T *se, int32_t *us) {
    switch (model) {
    case 0:
        return unc_model<T, models::Gaussian>(x, m, params, status, chi2, lower, upper, cov, se, us);
    case 1:
        return unc_model<T, models::Lorentzian>(x, m, params, status, chi2, lower, upper, cov, se,
                                                us);
    case 2:
        return unc_model<T, models::PseudoVoigt>(x, m, params, status, chi2, lower, upper, cov, se,
                                                 us);
    case 3:
        return unc_model<T, models::TwoLorentzian>(x, m, params, status, chi2, lower, upper, cov,
                                                   se, us);
    }
    return -1;
}
}  // namespace

extern "C" {
int ref_fit_f32(int model, const float *x, size_t m, const float *y, size_t n, const float *p0,
                const float *lower, const float *upper, const bf_oracle_options *options,
                int uncertainties, int threads, float *params, int32_t *status,
                uint32_t *iterations, float *chi2, float *cov, float *stderr_, int32_t *ustatus) {
    return fit_any<float>(model, x, m, y, n, p0, lower, upper, options, uncertainties, threads,
                          params, status, iterations, chi2, cov, stderr_, ustatus);
}
int ref_fit_f64(int model, const double *x, size_t m, const double *y, size_t n,
                const double *p0, const double *lower, const double *upper,
                const bf_oracle_options *options, int uncertainties, int threads, double *params,
                int32_t *status, uint32_t *iterations, double *chi2, double *cov,
                double *stderr_, int32_t *ustatus) {
    return fit_any<double>(model, x, m, y, n, p0, lower, upper, options, uncertainties, threads,
                           params, status, iterations, chi2, cov, stderr_, ustatus);
}
int ref_guess_f32(int model, const float *x, size_t m, const float *y, size_t n, int threads,
                  float *params, float *baseline, int32_t *status) {
    return guess_any<float>(model, x, m, y, n, threads, params, baseline, status);
}
int ref_guess_f64(int model, const double *x, size_t m, const double *y, size_t n, int threads,
                  double *params, double *baseline, int32_t *status) {
    return guess_any<double>(model, x, m, y, n, threads, params, baseline, status);
}
int ref_model_f32(int model, float x, const float *p, float *value, float *jac) {
    return model_any<float>(model, x, p, value, jac);
}
int ref_model_f64(int model, double x, const double *p, double *value, double *jac) {
    return model_any<double>(model, x, p, value, jac);
}
int ref_cholesky_f32(float *a, float *b, int n, int scaled) {
    return cholesky_any<float>(a, b, n, scaled);
}
int ref_cholesky_f64(double *a, double *b, int n, int scaled) {
    return cholesky_any<double>(a, b, n, scaled);
}
int ref_uncertainty_f32(int model, const float *x, size_t m, const float *params, int32_t status,
                        float chi2, const float *lower, const float *upper, float *cov,
                        float *stderr_, int32_t *ustatus) {
    return unc_any<float>(model, x, m, params, status, chi2, lower, upper, cov, stderr_, ustatus);
}
int ref_uncertainty_f64(int model, const double *x, size_t m, const double *params,
                        int32_t status, double chi2, const double *lower, const double *upper,
                        double *cov, double *stderr_, int32_t *ustatus) {
    return unc_any<double>(model, x, m, params, status, chi2, lower, upper, cov, stderr_, ustatus);
}
}
