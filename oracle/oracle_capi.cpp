// oracle_capi.cpp — C interface over bf_oracle.hpp (TEST INFRASTRUCTURE ONLY).
//
// Batches run rows in 64-fit chunks handed out by an atomic cursor to a fixed
// set of std::thread workers, the scheduling of cpu/batch.hpp:45-63 over
// src/cpu/thread_pool.cpp:9,157-173. Each row is the unchanged single-fit
// restatement, so results do not depend on the worker count.
#include "bf_oracle.hpp"
#include "oracle_api.h"

#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

namespace {
using namespace bf_oracle;

template <typename T> Options<T> make_options(const bf_oracle_options *in, const T *lower,
                                              const T *upper, int p) {
    Options<T> o;
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
        o.compensated = in->compensated != 0;
        o.scaled = in->scaled != 0;
    }
    for (int k = 0; k < p; ++k) {
        if (lower)
            o.lower[k] = lower[k];
        if (upper)
            o.upper[k] = upper[k];
    }
    return o;
}

template <typename F> void chunked(std::size_t n, int threads, F &&body) {
    constexpr std::size_t grain = 64;
    const std::size_t chunks = (n + grain - 1) / grain;
    std::atomic<std::size_t> cursor{0};
    auto work = [&] {
        for (;;) {
            const std::size_t c = cursor.fetch_add(1, std::memory_order_relaxed);
            if (c >= chunks)
                return;
            const std::size_t lo = c * grain, hi = std::min(n, lo + grain);
            for (std::size_t i = lo; i < hi; ++i)
                body(i);
        }
    };
    if (threads <= 1 || chunks <= 1) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back(work);
    for (auto &t : pool)
        t.join();
}

template <typename T, int MODEL>
void fit_rows(const T *x, std::size_t m, const T *y, std::size_t n, const T *p0,
              const Options<T> &o, int uncertainties, int threads, T *params, int32_t *status,
              uint32_t *iterations, T *chi2, T *cov, T *se, int32_t *us) {
    constexpr int P = param_count<MODEL>();
    chunked(n, threads, [&](std::size_t i) {
        const T *row = y + i * m;
        T init[8];
        if (p0) {
            for (int k = 0; k < P; ++k)
                init[k] = p0[i * P + k];
        } else {
            const Guess<T> g = guess_one<T, MODEL>(x, row, m);
            for (int k = 0; k < P; ++k)
                init[k] = g.params[k];
        }
        const Fit<T> f = fit_one<T, MODEL>(x, row, m, init, o);
        for (int k = 0; k < P; ++k)
            params[i * P + k] = f.params[k];
        status[i] = f.status;
        iterations[i] = f.iterations;
        chi2[i] = f.chi2;
        if (uncertainties) {
            const Uncertainty<T> u = covariance<T, MODEL>(x, m, f, o);
            for (int k = 0; k < P * P; ++k)
                cov[i * P * P + k] = u.cov[k];
            for (int k = 0; k < P; ++k)
                se[i * P + k] = u.stderr_[k];
            us[i] = u.status;
        }
    });
}

template <typename T>
int fit_any(int model, const T *x, std::size_t m, const T *y, std::size_t n, const T *p0,
            const T *lower, const T *upper, const bf_oracle_options *opt, int unc, int threads,
            T *params, int32_t *status, uint32_t *iterations, T *chi2, T *cov, T *se,
            int32_t *us) {
    switch (model) {
    case GAUSSIAN:
        fit_rows<T, GAUSSIAN>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 3), unc,
                              threads, params, status, iterations, chi2, cov, se, us);
        return 0;
    case LORENTZIAN:
        fit_rows<T, LORENTZIAN>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 3), unc,
                                threads, params, status, iterations, chi2, cov, se, us);
        return 0;
    case PSEUDO_VOIGT:
        fit_rows<T, PSEUDO_VOIGT>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 4), unc,
                                  threads, params, status, iterations, chi2, cov, se, us);
        return 0;
    case TWO_LORENTZIAN:
        fit_rows<T, TWO_LORENTZIAN>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 8), unc,
                                    threads, params, status, iterations, chi2, cov, se, us);
        return 0;
    case CONSTANT:
    case EXPONENTIAL:
    case AFFINE:
    case LINE:
    case DUPLICATE:
        if (!p0)
            return -1;  // the data-only guess exists for spectral models only
        switch (model) {
        case CONSTANT:
            fit_rows<T, CONSTANT>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 1), unc,
                                  threads, params, status, iterations, chi2, cov, se, us);
            return 0;
        case EXPONENTIAL:
            fit_rows<T, EXPONENTIAL>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 1), unc,
                                     threads, params, status, iterations, chi2, cov, se, us);
            return 0;
        case AFFINE:
            fit_rows<T, AFFINE>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 2), unc,
                                threads, params, status, iterations, chi2, cov, se, us);
            return 0;
        case LINE:
            fit_rows<T, LINE>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 2), unc, threads,
                              params, status, iterations, chi2, cov, se, us);
            return 0;
        default:
            fit_rows<T, DUPLICATE>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 2), unc,
                                   threads, params, status, iterations, chi2, cov, se, us);
            return 0;
        }
    }
    return -1;
}

// Fits (from p0 or the guess) with the values of their last stop tests,
// 11 doubles per row in the order of struct Stops.
template <typename T, int MODEL>
void stops_rows(const T *x, std::size_t m, const T *y, std::size_t n, const T *p0,
                const Options<T> &o, int threads, double *out) {
    constexpr int P = param_count<MODEL>();
    const T spacing = min_spacing(x, m);
    chunked(n, threads, [&](std::size_t i) {
        const T *row = y + i * m;
        T init[8];
        if (p0) {
            for (int k = 0; k < P; ++k)
                init[k] = p0[i * P + k];
        } else {
            const Guess<T> g = guess_one<T, MODEL>(x, row, m);
            for (int k = 0; k < P; ++k)
                init[k] = g.params[k];
        }
        Stops st;
        const Fit<T> f = fit_one<T, MODEL>(x, row, m, init, o, &st);
        st.width_over_spacing = width_ratio<T, MODEL>(f.params, spacing);
        const double v[11] = {st.gradient, st.gradient_tolerance, st.step, st.step_threshold,
                              st.reduction, st.cost_threshold, st.predicted,
                              st.predicted_threshold, st.width_over_spacing, st.lambda, st.fired};
        for (int k = 0; k < 11; ++k)
            out[i * 11 + k] = v[k];
    });
}

template <typename T>
int stops_any(int model, const T *x, std::size_t m, const T *y, std::size_t n, const T *p0,
              const T *lower, const T *upper, const bf_oracle_options *opt, int threads,
              double *out) {
    switch (model) {
    case GAUSSIAN:
        stops_rows<T, GAUSSIAN>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 3), threads, out);
        return 0;
    case LORENTZIAN:
        stops_rows<T, LORENTZIAN>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 3), threads,
                                  out);
        return 0;
    case PSEUDO_VOIGT:
        stops_rows<T, PSEUDO_VOIGT>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 4), threads,
                                    out);
        return 0;
    case TWO_LORENTZIAN:
        stops_rows<T, TWO_LORENTZIAN>(x, m, y, n, p0, make_options<T>(opt, lower, upper, 8),
                                      threads, out);
        return 0;
    }
    return -1;
}

template <typename T, int MODEL>
void guess_rows(const T *x, std::size_t m, const T *y, std::size_t n, int threads, T *params,
                T *baseline, int32_t *status) {
    constexpr int P = param_count<MODEL>();
    chunked(n, threads, [&](std::size_t i) {
        const Guess<T> g = guess_one<T, MODEL>(x, y + i * m, m);
        for (int k = 0; k < P; ++k)
            params[i * P + k] = g.params[k];
        baseline[i] = g.baseline;
        status[i] = g.status;
    });
}

template <typename T>
int guess_any(int model, const T *x, std::size_t m, const T *y, std::size_t n, int threads,
              T *params, T *baseline, int32_t *status) {
    switch (model) {
    case GAUSSIAN:
        guess_rows<T, GAUSSIAN>(x, m, y, n, threads, params, baseline, status);
        return 0;
    case LORENTZIAN:
        guess_rows<T, LORENTZIAN>(x, m, y, n, threads, params, baseline, status);
        return 0;
    case PSEUDO_VOIGT:
        guess_rows<T, PSEUDO_VOIGT>(x, m, y, n, threads, params, baseline, status);
        return 0;
    case TWO_LORENTZIAN:
        guess_rows<T, TWO_LORENTZIAN>(x, m, y, n, threads, params, baseline, status);
        return 0;
    }
    return -1;
}

template <typename T> int model_any(int model, T x, const T *p, T *value, T *jac) {
    switch (model) {
    case GAUSSIAN:
        *value = model_value<T, GAUSSIAN>(x, p);
        model_jacobian<T, GAUSSIAN>(x, p, jac);
        return 0;
    case LORENTZIAN:
        *value = model_value<T, LORENTZIAN>(x, p);
        model_jacobian<T, LORENTZIAN>(x, p, jac);
        return 0;
    case PSEUDO_VOIGT:
        *value = model_value<T, PSEUDO_VOIGT>(x, p);
        model_jacobian<T, PSEUDO_VOIGT>(x, p, jac);
        return 0;
    case TWO_LORENTZIAN:
        *value = model_value<T, TWO_LORENTZIAN>(x, p);
        model_jacobian<T, TWO_LORENTZIAN>(x, p, jac);
        return 0;
    }
    return -1;
}

template <typename T, int MODEL>
void unc_one(const T *x, std::size_t m, const T *params, int32_t status, T chi2,
             const T *lower, const T *upper, T *cov, T *se, int32_t *us) {
    constexpr int P = param_count<MODEL>();
    Fit<T> f;
    for (int k = 0; k < P; ++k)
        f.params[k] = params[k];
    f.status = status;
    f.chi2 = chi2;
    const Uncertainty<T> u = covariance<T, MODEL>(x, m, f, make_options<T>(nullptr, lower, upper, P));
    for (int k = 0; k < P * P; ++k)
        cov[k] = u.cov[k];
    for (int k = 0; k < P; ++k)
        se[k] = u.stderr_[k];
    *us = u.status;
}

template <typename T>
int unc_any(int model, const T *x, std::size_t m, const T *params, int32_t status, T chi2,
            const T *lower, const T *upper, T *cov, T *se, int32_t *us) {
    switch (model) {
    case GAUSSIAN:
        unc_one<T, GAUSSIAN>(x, m, params, status, chi2, lower, upper, cov, se, us);
        return 0;
    case LORENTZIAN:
        unc_one<T, LORENTZIAN>(x, m, params, status, chi2, lower, upper, cov, se, us);
        return 0;
    case PSEUDO_VOIGT:
        unc_one<T, PSEUDO_VOIGT>(x, m, params, status, chi2, lower, upper, cov, se, us);
        return 0;
    case TWO_LORENTZIAN:
        unc_one<T, TWO_LORENTZIAN>(x, m, params, status, chi2, lower, upper, cov, se, us);
        return 0;
    case LINE:
        unc_one<T, LINE>(x, m, params, status, chi2, lower, upper, cov, se, us);
        return 0;
    case DUPLICATE:
        unc_one<T, DUPLICATE>(x, m, params, status, chi2, lower, upper, cov, se, us);
        return 0;
    }
    return -1;
}
}  // namespace

extern "C" {
int oracle_fit_f32(int model, const float *x, size_t m, const float *y, size_t n,
                   const float *p0, const float *lower, const float *upper,
                   const bf_oracle_options *options, int uncertainties, int threads,
                   float *params, int32_t *status, uint32_t *iterations, float *chi2, float *cov,
                   float *stderr_, int32_t *ustatus) {
    return fit_any<float>(model, x, m, y, n, p0, lower, upper, options, uncertainties, threads,
                          params, status, iterations, chi2, cov, stderr_, ustatus);
}
int oracle_fit_f64(int model, const double *x, size_t m, const double *y, size_t n,
                   const double *p0, const double *lower, const double *upper,
                   const bf_oracle_options *options, int uncertainties, int threads,
                   double *params, int32_t *status, uint32_t *iterations, double *chi2,
                   double *cov, double *stderr_, int32_t *ustatus) {
    return fit_any<double>(model, x, m, y, n, p0, lower, upper, options, uncertainties, threads,
                           params, status, iterations, chi2, cov, stderr_, ustatus);
}
int oracle_fit_stops_f32(int model, const float *x, size_t m, const float *y, size_t n,
                         const float *p0, const float *lower, const float *upper,
                         const bf_oracle_options *options, int threads, double *stops) {
    return stops_any<float>(model, x, m, y, n, p0, lower, upper, options, threads, stops);
}
int oracle_fit_stops_f64(int model, const double *x, size_t m, const double *y, size_t n,
                         const double *p0, const double *lower, const double *upper,
                         const bf_oracle_options *options, int threads, double *stops) {
    return stops_any<double>(model, x, m, y, n, p0, lower, upper, options, threads, stops);
}
int oracle_guess_f32(int model, const float *x, size_t m, const float *y, size_t n, int threads,
                     float *params, float *baseline, int32_t *status) {
    return guess_any<float>(model, x, m, y, n, threads, params, baseline, status);
}
int oracle_guess_f64(int model, const double *x, size_t m, const double *y, size_t n,
                     int threads, double *params, double *baseline, int32_t *status) {
    return guess_any<double>(model, x, m, y, n, threads, params, baseline, status);
}
int oracle_model_f32(int model, float x, const float *p, float *value, float *jac) {
    return model_any<float>(model, x, p, value, jac);
}
int oracle_model_f64(int model, double x, const double *p, double *value, double *jac) {
    return model_any<double>(model, x, p, value, jac);
}
int oracle_cholesky_f32(float *a, float *b, int n, int scaled) {
    return chol_solve<float>(a, b, n, scaled != 0) ? 1 : 0;
}
int oracle_cholesky_f64(double *a, double *b, int n, int scaled) {
    return chol_solve<double>(a, b, n, scaled != 0) ? 1 : 0;
}
int oracle_uncertainty_f32(int model, const float *x, size_t m, const float *params,
                           int32_t status, float chi2, const float *lower, const float *upper,
                           float *cov, float *stderr_, int32_t *ustatus) {
    return unc_any<float>(model, x, m, params, status, chi2, lower, upper, cov, stderr_, ustatus);
}
int oracle_uncertainty_f64(int model, const double *x, size_t m, const double *params,
                           int32_t status, double chi2, const double *lower, const double *upper,
                           double *cov, double *stderr_, int32_t *ustatus) {
    return unc_any<double>(model, x, m, params, status, chi2, lower, upper, cov, stderr_,
                           ustatus);
}
float oracle_kahan_sum_f32(const float *terms, size_t n, float start) {
    float sum = start, carry = 0;
    for (size_t i = 0; i < n; ++i)
        comp_add(terms[i], sum, carry);
    return sum;
}
}
