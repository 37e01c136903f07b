/* oracle_api.h — C interface of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Two libraries export this interface under two prefixes:
 *   oracle/liboracle.so     prefix "oracle_": the restatement in bf_oracle.hpp
 *   oracle/_ref/libbfref.so prefix "ref_":    the reference's own CPU headers
 *                                             (cpu/solver.hpp, initial_guess.hpp,
 *                                             uncertainty.hpp, batch.hpp) and its
 *                                             ThreadPool, compiled in place.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load them.
 *
 * Model ids: 0 gaussian, 1 lorentzian, 2 pseudo_voigt, 3 two_lorentzian.
 * Layouts follow cpu/batch.hpp:21-24: shared x[M], fit-major y[N*M], p0[N*P].
 * A null p0 runs the data-only guess per row first and fits from its
 * parameters whatever its status, as python/bindings.cpp:69-80 does.
 */
#ifndef BF_ORACLE_API_H
#define BF_ORACLE_API_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Solver options (cpu/solver.hpp:26-41). Bits of `defaults` select the dtype
 * default instead of the given value: 1 lambda (1e-3), 2 gradient, 4 step,
 * 8 cost tolerance (float 1e-5/1e-6/1e-6, double 1e-10/1e-10/1e-12). */
typedef struct {
    uint32_t max_iterations; /* default 100 */
    uint32_t defaults;
    double initial_lambda;
    double gradient_tolerance;
    double step_tolerance;
    double cost_tolerance;
    int32_t compensated; /* Kahan sums in cost/linearize */
    int32_t scaled;      /* equilibrated Cholesky */
} bf_oracle_options;

#define BF_ORACLE_DECLARE(PREFIX)                                                              \
    int PREFIX##fit_f32(int model, const float *x, size_t m, const float *y, size_t n,         \
                        const float *p0, const float *lower, const float *upper,               \
                        const bf_oracle_options *options, int uncertainties, int threads,      \
                        float *params, int32_t *status, uint32_t *iterations, float *chi2,     \
                        float *cov, float *stderr_, int32_t *ustatus);                         \
    int PREFIX##fit_f64(int model, const double *x, size_t m, const double *y, size_t n,       \
                        const double *p0, const double *lower, const double *upper,            \
                        const bf_oracle_options *options, int uncertainties, int threads,      \
                        double *params, int32_t *status, uint32_t *iterations, double *chi2,   \
                        double *cov, double *stderr_, int32_t *ustatus);                       \
    int PREFIX##guess_f32(int model, const float *x, size_t m, const float *y, size_t n,       \
                          int threads, float *params, float *baseline, int32_t *status);       \
    int PREFIX##guess_f64(int model, const double *x, size_t m, const double *y, size_t n,     \
                          int threads, double *params, double *baseline, int32_t *status);     \
    int PREFIX##model_f32(int model, float x, const float *p, float *value, float *jac);       \
    int PREFIX##model_f64(int model, double x, const double *p, double *value, double *jac);   \
    int PREFIX##cholesky_f32(float *a, float *b, int n, int scaled);                           \
    int PREFIX##cholesky_f64(double *a, double *b, int n, int scaled);                         \
    int PREFIX##uncertainty_f32(int model, const float *x, size_t m, const float *params,      \
                                int32_t status, float chi2, const float *lower,                \
                                const float *upper, float *cov, float *stderr_,                \
                                int32_t *ustatus);                                             \
    int PREFIX##uncertainty_f64(int model, const double *x, size_t m, const double *params,    \
                                int32_t status, double chi2, const double *lower,              \
                                const double *upper, double *cov, double *stderr_,             \
                                int32_t *ustatus);

BF_ORACLE_DECLARE(oracle_)
BF_ORACLE_DECLARE(ref_)

/* Restatement only: the fits of oracle_fit_* with the values of each row's
 * last stop tests, 11 doubles per row: gradient norm, gradient tolerance,
 * step norm, step threshold, |chi2 reduction|, cost threshold, predicted
 * reduction, predicted threshold, smallest width*factor/spacing at the
 * result, final lambda, test that ended the fit (0 none, 1 gradient, 2 step,
 * 3 cost). Diagnostics for logging status disagreements with their margins. */
int oracle_fit_stops_f32(int model, const float *x, size_t m, const float *y, size_t n,
                         const float *p0, const float *lower, const float *upper,
                         const bf_oracle_options *options, int threads, double *stops);
int oracle_fit_stops_f64(int model, const double *x, size_t m, const double *y, size_t n,
                         const double *p0, const double *lower, const double *upper,
                         const bf_oracle_options *options, int threads, double *stops);

/* Restatement only: Kahan accumulation of n terms (numerics.hpp:12-18). */
float oracle_kahan_sum_f32(const float *terms, size_t n, float start);

#ifdef __cplusplus
}
#endif
#endif
