// bf_oracle.hpp — CPU restatement of batchfit's bounded Levenberg–Marquardt
// path, in float and double.
//
// TEST INFRASTRUCTURE ONLY. This header is the parity checker for the HIP
// library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it (through oracle/liboracle.so). The product path in
// batchfit_amd/ never includes, links or calls anything under oracle/.
//
// Every function restates the reference's arithmetic with the same operation
// order, so that a build with g++ -O2 -ffp-contract=off (no FMA contraction)
// reproduces the reference's float and double results bit for bit. That claim
// is pinned by tests/test_oracle_golden.py against vectors written by the
// reference's own headers (oracle/_ref) and by closed-form checks.
//
// Reference files followed (paths relative to the reference checkout):
//   models.hpp:78-185            model values and analytic Jacobians
//   numerics.hpp:12-18           compensated (Kahan) addition
//   cpu/solver.hpp:16-342        statuses, options, Cholesky, cost, linearize,
//                                fit_single
//   cpu/initial_guess.hpp:15-249 data-only initial guess
//   cpu/uncertainty.hpp:13-110   covariance at the accepted point
//   cpu/batch.hpp:23-65          row-parallel batch (64-fit chunks)
//
// Standard-library semantics the reference depends on are restated from
// libstdc++ (GCC 11.4, the toolchain in this image), which implements the
// C++20 definitions of P0811R3: std::lerp, std::midpoint; and std::clamp,
// std::min, std::max, std::max_element/min_element tie rules.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

namespace bf_oracle {

enum ModelId : int { GAUSSIAN = 0, LORENTZIAN = 1, PSEUDO_VOIGT = 2, TWO_LORENTZIAN = 3 };
// Non-spectral models of the reference's own checks, used only to pin the
// solver and covariance paths with closed forms: Constant, Exponential and
// Affine (tests/solver_checks.hpp:22-42), Line and Duplicate
// (tests/uncertainty_checks.hpp:8-18).
enum TestModelId : int { CONSTANT = 100, EXPONENTIAL = 101, AFFINE = 102, LINE = 103, DUPLICATE = 104 };
template <int MODEL> constexpr bool spectral() { return MODEL >= GAUSSIAN && MODEL <= TWO_LORENTZIAN; }

// FitStatus codes (cpu/solver.hpp:16-23); the int32 ABI of the Python outputs.
enum Status : int32_t {
    CONVERGED = 0,
    MAX_ITERATIONS = 1,
    SINGULAR_MATRIX = 2,
    INVALID_INPUT = 3,
    DEGENERATE_DATA = 4,
    UNRESOLVED_PEAK = 5
};
// UncertaintyStatus codes (uncertainty_types.hpp:3-10).
enum UStatus : int32_t {
    U_AVAILABLE = 0,
    U_NOT_CONVERGED = 1,
    U_INSUFFICIENT_DOF = 2,
    U_SINGULAR_MATRIX = 3,
    U_ACTIVE_BOUND = 4,
    U_INVALID_INPUT = 5
};
// GuessStatus codes (cpu/initial_guess.hpp:16).
enum GStatus : int32_t { G_SUCCESS = 0, G_INVALID_INPUT = 1, G_NO_PEAK = 2, G_INSUFFICIENT = 3 };

template <int MODEL> constexpr int param_count() {
    return MODEL == PSEUDO_VOIGT     ? 4
           : MODEL == TWO_LORENTZIAN ? 8
           : MODEL == CONSTANT || MODEL == EXPONENTIAL ? 1
           : MODEL == AFFINE || MODEL == LINE || MODEL == DUPLICATE ? 2
                                                     : 3;
}

// ---------------------------------------------------------------------------
// Standard-library helpers with libstdc++ tie and rounding rules.
template <typename T> inline T lo_of(T a, T b) { return (b < a) ? b : a; }    // std::min
template <typename T> inline T hi_of(T a, T b) { return (a < b) ? b : a; }    // std::max
template <typename T> inline T clamp_to(T v, T lo, T hi) {                     // std::clamp
    return (v < lo) ? lo : (hi < v) ? hi : v;
}
template <typename T> inline T mid(T a, T b) {  // std::midpoint, floating branch
    const T small = std::numeric_limits<T>::min() * 2;
    const T big = std::numeric_limits<T>::max() / 2;
    const T abs_a = a < 0 ? -a : a;
    const T abs_b = b < 0 ? -b : b;
    if (abs_a <= big && abs_b <= big)
        return (a + b) / 2;
    if (abs_a < small)
        return a + b / 2;
    if (abs_b < small)
        return a / 2 + b;
    return a / 2 + b / 2;
}
template <typename T> inline T interp(T a, T b, T t) {  // std::lerp
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0))
        return t * b + (1 - t) * a;
    if (t == 1)
        return b;
    const T v = a + t * (b - a);
    return (t > 1) == (b > a) ? (b < v ? v : b) : (b > v ? v : b);
}

template <typename T> inline T ln2() { return T(0.693147180559945309417232121458176568); }

// ---------------------------------------------------------------------------
// Peak models (models.hpp:78-185). The expression trees below are kept in the
// reference's evaluation order; every product chain is left-associated.
template <typename T, int MODEL> inline T model_value(T x, const T *p) {
    if constexpr (MODEL == GAUSSIAN) {
        const T z = (x - p[1]) / p[2];
        return p[0] * std::exp(-T(0.5f) * z * z);
    } else if constexpr (MODEL == LORENTZIAN) {
        const T z = (x - p[1]) / p[2];
        return p[0] / (T(1) + z * z);
    } else if constexpr (MODEL == PSEUDO_VOIGT) {
        const T z = (x - p[1]) / p[2];
        const T g = std::exp(-T(4) * ln2<T>() * z * z);
        const T l = T(1) / (T(1) + T(4) * z * z);
        return p[0] * ((T(1) - p[3]) * g + p[3] * l);
    } else if constexpr (MODEL == TWO_LORENTZIAN) {
        return model_value<T, LORENTZIAN>(x, p) + model_value<T, LORENTZIAN>(x, p + 3) + p[6] +
               p[7] * x;
    } else if constexpr (MODEL == CONSTANT) {
        return p[0];
    } else if constexpr (MODEL == EXPONENTIAL) {
        return std::exp(x * p[0]);
    } else {
        return p[0] + p[1] * x;
    }
}

template <typename T, int MODEL> inline void model_jacobian(T x, const T *p, T *jac) {
    if constexpr (MODEL == GAUSSIAN) {
        const T z = (x - p[1]) / p[2];
        const T e = std::exp(-T(0.5f) * z * z);
        jac[0] = e;
        jac[1] = p[0] * e * z / p[2];
        jac[2] = p[0] * e * z * z / p[2];
    } else if constexpr (MODEL == LORENTZIAN) {
        const T z = (x - p[1]) / p[2];
        const T s = T(1) / (T(1) + z * z);
        const T k = T(2) * p[0] * s * s / p[2];
        jac[0] = s;
        jac[1] = k * z;
        jac[2] = k * z * z;
    } else if constexpr (MODEL == PSEUDO_VOIGT) {
        const T z = (x - p[1]) / p[2];
        const T g = std::exp(-T(4) * ln2<T>() * z * z);
        const T l = T(1) / (T(1) + T(4) * z * z);
        const T k = p[0] * T(8) / p[2] * ((T(1) - p[3]) * ln2<T>() * g + p[3] * l * l);
        jac[0] = (T(1) - p[3]) * g + p[3] * l;
        jac[1] = k * z;
        jac[2] = k * z * z;
        jac[3] = p[0] * (l - g);
    } else if constexpr (MODEL == TWO_LORENTZIAN) {
        model_jacobian<T, LORENTZIAN>(x, p, jac);
        model_jacobian<T, LORENTZIAN>(x, p + 3, jac + 3);
        jac[6] = T(1);
        jac[7] = x;
    } else if constexpr (MODEL == CONSTANT) {
        jac[0] = T(1);
    } else if constexpr (MODEL == EXPONENTIAL) {
        jac[0] = x * model_value<T, EXPONENTIAL>(x, p);
    } else if constexpr (MODEL == DUPLICATE) {
        jac[0] = jac[1] = T(1);
    } else {
        jac[0] = T(1);
        jac[1] = x;
    }
}

// Kahan compensated addition (numerics.hpp:12-18). Must not be reassociated.
template <typename T> inline void comp_add(T term, T &sum, T &carry) {
    const T adjusted = term - carry;
    const T next = sum + adjusted;
    carry = (next - sum) - adjusted;
    sum = next;
}

// ---------------------------------------------------------------------------
// Options and results (cpu/solver.hpp:26-48).
template <typename T> struct Options {
    unsigned max_iterations = 100;
    T initial_lambda = T(1e-3);
    T gradient_tolerance = std::is_same_v<T, float> ? T(1e-5) : T(1e-10);
    T step_tolerance = std::is_same_v<T, float> ? T(1e-6) : T(1e-10);
    T cost_tolerance = std::is_same_v<T, float> ? T(1e-6) : T(1e-12);
    bool compensated = false;
    bool scaled = false;
    T lower[8];
    T upper[8];
    Options() {
        for (int k = 0; k < 8; ++k) {
            lower[k] = -std::numeric_limits<T>::infinity();
            upper[k] = std::numeric_limits<T>::infinity();
        }
    }
};

template <typename T> struct Fit {
    T params[8];
    int32_t status = INVALID_INPUT;
    uint32_t iterations = 0;
    T chi2 = std::numeric_limits<T>::quiet_NaN();
};

template <typename T> struct Uncertainty {
    T cov[64];
    T stderr_[8];
    int32_t status = U_NOT_CONVERGED;
};

template <typename T> struct Guess {
    T params[8];
    T baseline = std::numeric_limits<T>::quiet_NaN();
    int32_t status = G_INVALID_INPUT;
};

// ---------------------------------------------------------------------------
// Shared-grid helpers (cpu/solver.hpp:58-96).
template <typename T> T min_spacing(const T *x, std::size_t m) {
    T spacing = std::numeric_limits<T>::infinity();
    bool rising = true, falling = true;
    for (std::size_t j = 1; j < m; ++j) {
        const T d = x[j] - x[j - 1];
        rising = rising && d >= T(0);
        falling = falling && d <= T(0);
        if (d != T(0))
            spacing = lo_of(spacing, std::abs(d));
    }
    return (rising || falling) && std::isfinite(spacing) ? spacing : T(0);
}

template <typename T, int MODEL> bool shape_is_free(const Options<T> &o) {
    if constexpr (!spectral<MODEL>()) {
        return false;
    } else {
        bool free = o.lower[1] != o.upper[1] || o.lower[2] != o.upper[2];
        if constexpr (param_count<MODEL>() == 4)
            free = free || o.lower[3] != o.upper[3];
        if constexpr (param_count<MODEL>() == 8)
            free = free || o.lower[4] != o.upper[4] || o.lower[5] != o.upper[5];
        return free;
    }
}

template <typename T, int MODEL> T width_factor() {
    return MODEL == GAUSSIAN ? T(2.354820045030949382) : MODEL == PSEUDO_VOIGT ? T(1) : T(2);
}

template <typename T, int MODEL> int32_t resolved(const T *p, T spacing) {
    if constexpr (!spectral<MODEL>()) {
        return CONVERGED;
    } else {
        const T factor = width_factor<T, MODEL>();
        if (spacing > T(0) && std::abs(p[2]) * factor < spacing)
            return UNRESOLVED_PEAK;
        if constexpr (param_count<MODEL>() == 8)
            if (spacing > T(0) && std::abs(p[5]) * factor < spacing)
                return UNRESOLVED_PEAK;
        return CONVERGED;
    }
}

// Checked, unpivoted Cholesky solve (cpu/solver.hpp:99-143). The lower
// triangle of a[] (row-major, n×n) becomes the factor; b becomes the solution.
template <typename T> bool chol_solve(T *a, T *b, int n, bool scaled) {
    T d[8];
    if (scaled) {
        for (int i = 0; i < n; ++i) {
            if (!std::isfinite(a[i * n + i]) || a[i * n + i] <= T(0))
                return false;
            d[i] = std::sqrt(a[i * n + i]);
        }
        for (int i = 0; i < n; ++i) {
            b[i] /= d[i];
            for (int j = 0; j < n; ++j)
                a[i * n + j] = (a[i * n + j] / d[i]) / d[j];
        }
    }
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j <= i; ++j) {
            T v = a[i * n + j];
            for (int k = 0; k < j; ++k)
                v -= a[i * n + k] * a[j * n + k];
            if (!std::isfinite(v) || (i == j && v <= T(0)))
                return false;
            a[i * n + j] = i == j ? std::sqrt(v) : v / a[j * n + j];
        }
        T v = b[i];
        for (int j = 0; j < i; ++j)
            v -= a[i * n + j] * b[j];
        b[i] = v / a[i * n + i];
    }
    for (int i = n; i-- > 0;) {
        T v = b[i];
        for (int j = i + 1; j < n; ++j)
            v -= a[j * n + i] * b[j];
        b[i] = v / a[i * n + i];
        if (!std::isfinite(b[i]))
            return false;
    }
    if (scaled)
        for (int i = 0; i < n; ++i) {
            b[i] /= d[i];
            if (!std::isfinite(b[i]))
                return false;
        }
    return true;
}

// χ² = Σ (y − f)² (cpu/solver.hpp:145-157).
template <typename T, int MODEL>
T chi_squared(const T *x, const T *y, std::size_t m, const T *p, bool compensated) {
    T sum = 0, carry = 0;
    for (std::size_t j = 0; j < m; ++j) {
        const T r = y[j] - model_value<T, MODEL>(x[j], p);
        if (compensated)
            comp_add(r * r, sum, carry);
        else
            sum += r * r;
    }
    return sum;
}

// Normal equations A = JᵀJ (full, mirrored), g = Jᵀr (cpu/solver.hpp:159-192).
template <typename T, int MODEL>
bool normal_equations(const T *x, const T *y, std::size_t m, const T *p, T *a, T *g,
                      bool compensated) {
    constexpr int P = param_count<MODEL>();
    T ac[P * P], gc[P];
    for (int k = 0; k < P * P; ++k)
        a[k] = ac[k] = 0;
    for (int k = 0; k < P; ++k)
        g[k] = gc[k] = 0;
    for (std::size_t j = 0; j < m; ++j) {
        const T r = y[j] - model_value<T, MODEL>(x[j], p);
        T jac[P];
        model_jacobian<T, MODEL>(x[j], p, jac);
        for (int k = 0; k < P; ++k) {
            if (compensated)
                comp_add(jac[k] * r, g[k], gc[k]);
            else
                g[k] += jac[k] * r;
            for (int l = 0; l <= k; ++l)
                if (compensated)
                    comp_add(jac[k] * jac[l], a[k * P + l], ac[k * P + l]);
                else
                    a[k * P + l] += jac[k] * jac[l];
        }
    }
    for (int k = 0; k < P; ++k) {
        if (!std::isfinite(g[k]))
            return false;
        for (int l = 0; l <= k; ++l) {
            if (!std::isfinite(a[k * P + l]))
                return false;
            a[l * P + k] = a[k * P + l];
        }
    }
    return true;
}

template <typename T> bool all_equal_to_first(const T *v, std::size_t m) {
    for (std::size_t j = 0; j < m; ++j)
        if (!(v[j] == v[0]))
            return false;
    return true;
}

// Values of the stop tests of a fit's last iteration (diagnostics for
// logging why two fits of the same row stopped differently; they do not
// change the fit): the gradient norm and its tolerance, the step norm and its
// threshold, |χ² reduction| of the last trial with the cost threshold, the
// predicted reduction with its threshold, the smallest width·factor over the
// grid spacing at the returned parameters (below 1: unresolved peak), the
// final λ and which test ended the fit (0 none, 1 gradient, 2 step, 3 cost).
struct Stops {
    double gradient = NAN, gradient_tolerance = NAN, step = NAN, step_threshold = NAN;
    double reduction = NAN, cost_threshold = NAN, predicted = NAN, predicted_threshold = NAN;
    double width_over_spacing = NAN, lambda = NAN, fired = 0;
};

// One bounded LM fit (cpu/solver.hpp:196-342).
template <typename T, int MODEL>
Fit<T> fit_one(const T *x, const T *y, std::size_t m, const T *init, const Options<T> &o,
               Stops *stops = nullptr) {
    constexpr int P = param_count<MODEL>();
    Fit<T> out;
    for (int k = 0; k < 8; ++k)
        out.params[k] = std::numeric_limits<T>::quiet_NaN();
    for (int k = 0; k < P; ++k)
        out.params[k] = init[k];
    if (m == 0 || !std::isfinite(o.initial_lambda) || o.initial_lambda <= T(0))
        return out;
    for (const T tol : {o.gradient_tolerance, o.step_tolerance, o.cost_tolerance})
        if (!std::isfinite(tol) || tol < T(0))
            return out;
    for (int k = 0; k < P; ++k)
        if (!std::isfinite(init[k]) || std::isnan(o.lower[k]) || std::isnan(o.upper[k]) ||
            o.lower[k] > o.upper[k] || o.lower[k] == std::numeric_limits<T>::infinity() ||
            o.upper[k] == -std::numeric_limits<T>::infinity())
            return out;
    for (std::size_t j = 0; j < m; ++j)
        if (!std::isfinite(x[j]) || !std::isfinite(y[j]))
            return out;
    for (int k = 0; k < P; ++k)
        out.params[k] = clamp_to(init[k], o.lower[k], o.upper[k]);
    out.chi2 = chi_squared<T, MODEL>(x, y, m, out.params, o.compensated);
    if (!std::isfinite(out.chi2)) {
        out.chi2 = std::numeric_limits<T>::quiet_NaN();
        return out;
    }
    const T spacing = min_spacing(x, m);
    bool varied_x = false;
    for (std::size_t j = 0; j < m; ++j)
        varied_x = varied_x || x[j] != x[0];
    if (spectral<MODEL>() && varied_x && all_equal_to_first(y, m) && shape_is_free<T, MODEL>(o)) {
        out.status = DEGENERATE_DATA;
        return out;
    }
    out.status = MAX_ITERATIONS;
    T lambda = o.initial_lambda;
    const T lambda_cap = hi_of(o.initial_lambda, T(1) / std::numeric_limits<T>::epsilon());
    for (unsigned it = 0; it < o.max_iterations; ++it) {
        T a[P * P], g[P];
        if (!normal_equations<T, MODEL>(x, y, m, out.params, a, g, o.compensated)) {
            out.status = INVALID_INPUT;
            return out;
        }
        bool pinned[P];
        T gnorm = 0;
        for (int k = 0; k < P; ++k) {
            pinned[k] = o.lower[k] == o.upper[k] || (out.params[k] == o.lower[k] && g[k] < T(0)) ||
                        (out.params[k] == o.upper[k] && g[k] > T(0));
            if (!pinned[k])
                gnorm = hi_of(gnorm, std::abs(g[k]));
        }
        if (stops) {
            stops->gradient = double(gnorm);
            stops->gradient_tolerance = double(o.gradient_tolerance);
            stops->lambda = double(lambda);
        }
        if (gnorm <= o.gradient_tolerance) {
            if (stops)
                stops->fired = 1;
            out.status = resolved<T, MODEL>(out.params, spacing);
            return out;
        }
        T damped[P * P], step[P];
        for (int k = 0; k < P * P; ++k)
            damped[k] = a[k];
        for (int k = 0; k < P; ++k)
            step[k] = g[k];
        for (int k = 0; k < P; ++k) {
            if (pinned[k]) {
                for (int l = 0; l < P; ++l)
                    damped[k * P + l] = damped[l * P + k] = T(0);
                damped[k * P + k] = T(1);
                step[k] = T(0);
            } else {
                damped[k * P + k] += lambda * a[k * P + k];
            }
        }
        out.iterations = it + 1;
        if (!chol_solve(damped, step, P, o.scaled)) {
            out.status = SINGULAR_MATRIX;
            return out;
        }
        T trial[P];
        T step_norm = 0, param_norm = 0, predicted = 0;
        bool finite_trial = true;
        for (int k = 0; k < P; ++k) {
            step_norm = hi_of(step_norm, std::abs(step[k]));
            trial[k] = clamp_to(out.params[k] + step[k], o.lower[k], o.upper[k]);
            finite_trial = finite_trial && std::isfinite(trial[k]);
            step[k] = trial[k] - out.params[k];
            param_norm = hi_of(param_norm, std::abs(out.params[k]));
        }
        if (stops) {
            stops->step = double(step_norm);
            stops->step_threshold =
                double(o.step_tolerance * (param_norm + o.step_tolerance) / hi_of(T(1), lambda));
        }
        if (step_norm <=
            o.step_tolerance * (param_norm + o.step_tolerance) / hi_of(T(1), lambda)) {
            if (stops)
                stops->fired = 2;
            out.status = resolved<T, MODEL>(out.params, spacing);
            return out;
        }
        for (int k = 0; k < P; ++k) {
            predicted += T(2) * g[k] * step[k];
            for (int l = 0; l < P; ++l)
                predicted -= step[k] * a[k * P + l] * step[l];
        }
        const T trial_chi2 = finite_trial
                                 ? chi_squared<T, MODEL>(x, y, m, trial, o.compensated)
                                 : std::numeric_limits<T>::quiet_NaN();
        const T reduction = out.chi2 - trial_chi2;
        const T threshold = o.cost_tolerance * out.chi2;
        const bool small_change = std::isfinite(trial_chi2) && std::isfinite(predicted) &&
                                  predicted >= T(0) && std::abs(reduction) <= threshold &&
                                  predicted <= threshold / hi_of(T(1), lambda);
        if (stops) {
            stops->reduction = double(std::abs(reduction));
            stops->cost_threshold = double(threshold);
            stops->predicted = double(predicted);
            stops->predicted_threshold = double(threshold / hi_of(T(1), lambda));
        }
        if (std::isfinite(trial_chi2) && reduction > T(0) && std::isfinite(predicted) &&
            predicted > T(0)) {
            for (int k = 0; k < P; ++k)
                out.params[k] = trial[k];
            out.chi2 = trial_chi2;
            const T rho = reduction / predicted;
            const T q = T(2) * rho - T(1);
            const T factor = rho >= T(1) ? T(1) / T(3) : hi_of(T(1) / T(3), T(1) - q * q * q);
            lambda = clamp_to(lambda * factor, std::numeric_limits<T>::epsilon(), lambda_cap);
        } else {
            lambda = lo_of(lambda_cap, lambda * T(10));
        }
        if (small_change) {
            if (stops)
                stops->fired = 3;
            out.status = resolved<T, MODEL>(out.params, spacing);
            return out;
        }
    }
    return out;
}

// Smallest width·factor over the spacing of a spectral model's peaks (the
// quantity of resolved(): below 1 means unresolved_peak).
template <typename T, int MODEL> double width_ratio(const T *p, T spacing) {
    if constexpr (!spectral<MODEL>()) {
        return NAN;
    } else {
        const double f = double(width_factor<T, MODEL>());
        double r = std::abs(double(p[2])) * f / double(spacing);
        if constexpr (param_count<MODEL>() == 8)
            r = std::min(r, std::abs(double(p[5])) * f / double(spacing));
        return r;
    }
}

// ---------------------------------------------------------------------------
// Covariance at the accepted point (cpu/uncertainty.hpp:13-110).
template <typename T, int MODEL>
Uncertainty<T> covariance(const T *x, std::size_t m, const Fit<T> &fit, const Options<T> &o) {
    constexpr int P = param_count<MODEL>();
    const T nan = std::numeric_limits<T>::quiet_NaN();
    Uncertainty<T> u;
    for (int k = 0; k < 64; ++k)
        u.cov[k] = nan;
    for (int k = 0; k < 8; ++k)
        u.stderr_[k] = nan;
    if (fit.status != CONVERGED)
        return u;
    u.status = U_INVALID_INPUT;
    if (!std::isfinite(fit.chi2) || fit.chi2 < T(0))
        return u;
    bool fixed[P] = {};
    unsigned free = 0;
    for (int k = 0; k < P; ++k) {
        if (!std::isfinite(fit.params[k]) || std::isnan(o.lower[k]) || std::isnan(o.upper[k]) ||
            o.lower[k] > o.upper[k] || fit.params[k] < o.lower[k] || fit.params[k] > o.upper[k])
            return u;
        fixed[k] = o.lower[k] == o.upper[k];
        free += !fixed[k];
        if (!fixed[k] && (fit.params[k] == o.lower[k] || fit.params[k] == o.upper[k])) {
            u.status = U_ACTIVE_BOUND;
            return u;
        }
    }
    if (m <= free) {
        u.status = U_INSUFFICIENT_DOF;
        return u;
    }
    T info[P * P] = {}, carry[P * P] = {};
    for (std::size_t j = 0; j < m; ++j) {
        if (!std::isfinite(x[j]))
            return u;
        T jac[P];
        model_jacobian<T, MODEL>(x[j], fit.params, jac);
        for (int i = 0; i < P; ++i) {
            if (!std::isfinite(jac[i]))
                return u;
            for (int l = 0; l <= i; ++l)
                comp_add(jac[i] * jac[l], info[i * P + l], carry[i * P + l]);
        }
    }
    for (int i = 0; i < P; ++i)
        for (int l = 0; l <= i; ++l) {
            if (!std::isfinite(info[i * P + l]))
                return u;
            info[l * P + i] = info[i * P + l];
        }
    for (int i = 0; i < P; ++i)
        if (fixed[i]) {
            for (int l = 0; l < P; ++l)
                info[i * P + l] = info[l * P + i] = T(0);
            info[i * P + i] = T(1);
        }
    const T variance = fit.chi2 / T(m - free);
    if (!std::isfinite(variance))
        return u;
    T cov[P * P] = {};
    for (int col = 0; col < P; ++col) {
        if (fixed[col])
            continue;
        T factor[P * P], rhs[P] = {};
        for (int k = 0; k < P * P; ++k)
            factor[k] = info[k];
        rhs[col] = T(1);
        if (!chol_solve(factor, rhs, P, true)) {
            u.status = U_SINGULAR_MATRIX;
            return u;
        }
        for (int i = 0; i < P; ++i) {
            if (!fixed[i] && factor[i * P + i] * factor[i * P + i] <=
                                 T(64) * std::numeric_limits<T>::epsilon()) {
                u.status = U_SINGULAR_MATRIX;
                return u;
            }
            cov[i * P + col] = fixed[i] ? T(0) : rhs[i] * variance;
            if (!std::isfinite(cov[i * P + col]))
                return u;
        }
    }
    for (int i = 0; i < P; ++i) {
        if (cov[i * P + i] < T(0))
            return u;
        for (int l = 0; l < i; ++l) {
            const T s = cov[i * P + l] / T(2) + cov[l * P + i] / T(2);
            cov[i * P + l] = cov[l * P + i] = s;
        }
    }
    for (int k = 0; k < P * P; ++k)
        u.cov[k] = cov[k];
    for (int i = 0; i < P; ++i)
        u.stderr_[i] = std::sqrt(cov[i * P + i]);
    u.status = U_AVAILABLE;
    return u;
}

// ---------------------------------------------------------------------------
// Data-only initial guess (cpu/initial_guess.hpp:25-249).

// Median of a multiset: the middle order statistic, or the midpoint of the
// two middle ones for an even count (initial_guess.hpp:25-30).
template <typename T> T median_of(std::vector<T> v) {
    std::sort(v.begin(), v.end());
    const std::size_t n = v.size();
    return n % 2 ? v[n / 2] : mid(v[n / 2 - 1], v[n / 2]);
}

// [¼, ½, ¼] smoothing with unsmoothed end points (initial_guess.hpp:32-36).
template <typename T> T smoothed(const T *s, std::size_t m, std::size_t i) {
    if (i == 0 || i + 1 == m)
        return s[i];
    return T(0.25) * s[i - 1] + T(0.5) * s[i] + T(0.25) * s[i + 1];
}

// Height above the higher of the two bounding minima (initial_guess.hpp:38-54).
template <typename T> T prominence_of(const T *s, std::size_t m, std::size_t index) {
    const T height = smoothed(s, m, index);
    T left = height, right = height;
    for (std::size_t j = index; j > 0;) {
        const T v = smoothed(s, m, --j);
        if (v > height)
            break;
        left = lo_of(left, v);
    }
    for (std::size_t j = index + 1; j < m; ++j) {
        const T v = smoothed(s, m, j);
        if (v > height)
            break;
        right = lo_of(right, v);
    }
    return height - hi_of(left, right);
}

template <typename T> struct PeakShape {
    T amplitude = 0, center = 0, width = 0, left_radius = 0, right_radius = 0;
    bool valid = false;
};

// Centroid and half-height crossings around one maximum (initial_guess.hpp:61-109).
template <typename T>
PeakShape<T> measure_peak(const T *x, const T *s, std::size_t m, std::size_t index,
                          std::size_t begin, std::size_t end, bool mirror) {
    PeakShape<T> r;
    r.amplitude = s[index];
    if (!(r.amplitude > T(0)))
        return r;
    const T half = r.amplitude / T(2);
    std::size_t left = index, right = index;
    while (left > begin && s[left] > half)
        --left;
    while (right < end && s[right] > half)
        ++right;
    const bool has_left = s[left] <= half, has_right = s[right] <= half;
    if ((!has_left && (!mirror || begin == 0)) || (!has_right && (!mirror || end + 1 == m)) ||
        (!has_left && !has_right))
        return r;
    std::size_t low = index > 2 ? index - 2 : 0;
    std::size_t high = lo_of(index + 2, end);
    low = hi_of(low, has_left ? left + 1 : begin);
    high = lo_of(high, has_right ? right - 1 : end);
    T total = 0;
    r.center = x[index];
    for (std::size_t j = low; j <= high; ++j) {
        const T w = hi_of(T(0), s[j]) / r.amplitude;
        if (w > T(0)) {
            total += w;
            r.center = interp(r.center, x[j], w / total);
        }
    }
    const auto crossing = [&](std::size_t a, std::size_t b) {
        const T den = s[b] - s[a];
        if (!std::isfinite(den) || den == T(0))
            return std::numeric_limits<T>::quiet_NaN();
        return interp(x[a], x[b], (half - s[a]) / den);
    };
    if (has_left)
        r.left_radius = r.center - crossing(left, left + 1);
    if (has_right)
        r.right_radius = crossing(right - 1, right) - r.center;
    if (!has_left)
        r.left_radius = r.right_radius;
    if (!has_right)
        r.right_radius = r.left_radius;
    r.width = r.left_radius + r.right_radius;
    r.valid = std::isfinite(r.center) && std::isfinite(r.width) && r.left_radius > T(0) &&
              r.right_radius > T(0);
    return r;
}

// First maximum of the raw signal over [i-1, i+1] (initial_guess.hpp:111-115).
template <typename T> std::size_t raw_peak(const T *s, std::size_t i) {
    std::size_t best = i - 1;
    for (std::size_t j = i; j <= i + 1; ++j)
        if (s[best] < s[j])
            best = j;
    return best;
}

template <typename T, int MODEL> Guess<T> guess_one(const T *x, const T *y, std::size_t m) {
    constexpr bool paired = MODEL == TWO_LORENTZIAN;
    constexpr int P = param_count<MODEL>();
    Guess<T> g;
    for (int k = 0; k < 8; ++k)
        g.params[k] = std::numeric_limits<T>::quiet_NaN();
    if (m < 5)
        return g;
    T spacing = std::numeric_limits<T>::max();
    for (std::size_t j = 0; j < m; ++j) {
        if (!std::isfinite(x[j]) || !std::isfinite(y[j]) || (j > 0 && x[j] <= x[j - 1]))
            return g;
        if (j > 0)
            spacing = lo_of(spacing, x[j] - x[j - 1]);
    }
    if (!std::isfinite(x[m - 1] - x[0]))
        return g;
    const std::size_t count = hi_of(std::size_t(1), m / 10);
    std::vector<T> left_edge(y, y + count), right_edge(y + m - count, y + m);
    std::vector<T> both(left_edge);
    both.insert(both.end(), right_edge.begin(), right_edge.end());
    const T left_level = median_of(left_edge);
    const T right_level = median_of(right_edge);
    g.baseline = median_of(both);
    const T left_x = mid(x[(count - 1) / 2], x[count / 2]);
    const T right_x = mid(x[m - count + (count - 1) / 2], x[m - count + count / 2]);
    const T slope = paired ? (right_level - left_level) / (right_x - left_x) : T(0);
    const T level = paired ? left_level : g.baseline;
    if (!std::isfinite(slope))
        return g;
    std::vector<T> signal(m);
    for (std::size_t j = 0; j < m; ++j) {
        signal[j] = y[j] - (level + slope * (x[j] - left_x));
        if (!std::isfinite(signal[j]))
            return g;
    }
    std::vector<T> spread(2 * count);
    for (std::size_t j = 0; j < count; ++j) {
        spread[j] = std::abs(y[j] - left_level);
        spread[count + j] = std::abs(y[m - count + j] - right_level);
    }
    const T noise = T(1.4826) * median_of(spread);
    if (!std::isfinite(noise))
        return g;
    const T *s = signal.data();
    std::size_t first = m;
    T height = 0;
    for (std::size_t j = 1; j + 1 < m; ++j) {
        const T v = smoothed(s, m, j);
        if (v > height && v > smoothed(s, m, j - 1) && v >= smoothed(s, m, j + 1)) {
            first = j;
            height = v;
        }
    }
    g.status = G_NO_PEAK;
    if (first == m || height <= T(3) * noise)
        return g;
    first = raw_peak(s, first);
    PeakShape<T> primary = measure_peak(x, s, m, first, 0, m - 1, false);
    if (!primary.valid)
        return g;
    if constexpr (!paired) {
        g.params[0] = primary.amplitude;
        g.params[1] = primary.center;
        if constexpr (MODEL == GAUSSIAN)
            g.params[2] = primary.width / T(2.354820045030949382);
        else if constexpr (MODEL == LORENTZIAN)
            g.params[2] = primary.width / T(2);
        else {
            g.params[2] = primary.width;
            g.params[3] = T(0.5);
        }
    } else {
        g.status = G_INSUFFICIENT;
        const T separation =
            hi_of(T(3) * spacing, T(2) * lo_of(primary.left_radius, primary.right_radius));
        std::size_t second = m;
        T second_height = 0;
        for (std::size_t j = 1; j + 1 < m; ++j) {
            const T v = smoothed(s, m, j);
            if (v <= second_height || v <= smoothed(s, m, j - 1) || v < smoothed(s, m, j + 1))
                continue;
            const std::size_t index = raw_peak(s, j);
            const std::size_t bins = index > first ? index - first : first - index;
            if (bins < 3 || std::abs(x[index] - primary.center) < separation ||
                prominence_of(s, m, j) <= T(3) * noise)
                continue;
            second = index;
            second_height = v;
        }
        if (second == m)
            return g;
        if (first > second)
            std::swap(first, second);
        std::size_t valley = first + 1;
        for (std::size_t j = first + 1; j < second; ++j)
            if (s[j] < s[valley])
                valley = j;
        primary = measure_peak(x, s, m, first, 0, valley, true);
        const PeakShape<T> secondary = measure_peak(x, s, m, second, valley, m - 1, true);
        if (!primary.valid || !secondary.valid || !(primary.center < secondary.center))
            return g;
        const T values[8] = {primary.amplitude,   primary.center,
                             primary.width / T(2), secondary.amplitude,
                             secondary.center,    secondary.width / T(2),
                             left_level - slope * left_x, slope};
        for (int k = 0; k < 8; ++k)
            g.params[k] = values[k];
    }
    for (int k = 0; k < P; ++k)
        if (!std::isfinite(g.params[k])) {
            g.status = G_INVALID_INPUT;
            return g;
        }
    if (!(g.params[2] > T(0)) || (paired && !(g.params[5] > T(0)))) {
        g.status = G_INVALID_INPUT;
        return g;
    }
    g.status = G_SUCCESS;
    return g;
}

}  // namespace bf_oracle
