// bf_ragged_runtime.hip — host side of ragged batches (bf_hip_ragged_*):
// validation, class bucketing with longest-first order, upload, and one
// captured hipGraph per pointer set: queue reset → grid table → guesses →
// persistent LM kernels, all on the plan's stream.
#include "bf_host.hpp"
#include "bf_ragged.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace bfhip {

template <int MODEL> hipError_t launch_ragged_lm(int, bool, const RaggedArgs &, uint32_t,
                                                 hipStream_t);
template <int MODEL>
hipError_t launch_ragged_guess(const bf_fit_desc *, const uint32_t *, uint32_t,
                               const bf_grid_desc *, const GridInfo *, const float *,
                               const float *, float *, uint32_t, uint32_t, hipStream_t);

namespace {

__global__ void __launch_bounds__(64) reset_kernel(uint32_t *cursors, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += 64)
        cursors[i] = 0;
}

__global__ void __launch_bounds__(256) grid_table_kernel(const bf_grid_desc *grids, const float *x,
                                                         GridInfo *info) {
    const bf_grid_desc g = grids[blockIdx.x];
    grid_block(x + g.x_offset, g.points, info + blockIdx.x);
}

hipError_t ragged_lm(uint32_t model, int cls, bool mixed, const RaggedArgs &a, uint32_t blocks,
                     hipStream_t s) {
    switch (model) {
    case 0: return launch_ragged_lm<0>(cls, mixed, a, blocks, s);
    case 1: return launch_ragged_lm<1>(cls, mixed, a, blocks, s);
    case 2: return launch_ragged_lm<2>(cls, mixed, a, blocks, s);
    case 3: return launch_ragged_lm<3>(cls, mixed, a, blocks, s);
    }
    return hipErrorInvalidValue;
}

hipError_t ragged_guess(uint32_t model, const bf_fit_desc *fits, const uint32_t *order,
                        uint32_t count, const bf_grid_desc *grids, const GridInfo *info,
                        const float *x, const float *y, float *guess, uint32_t m_pad,
                        uint32_t sort_n, hipStream_t s) {
    switch (model) {
    case 0: return launch_ragged_guess<0>(fits, order, count, grids, info, x, y, guess, m_pad, sort_n, s);
    case 1: return launch_ragged_guess<1>(fits, order, count, grids, info, x, y, guess, m_pad, sort_n, s);
    case 2: return launch_ragged_guess<2>(fits, order, count, grids, info, x, y, guess, m_pad, sort_n, s);
    case 3: return launch_ragged_guess<3>(fits, order, count, grids, info, x, y, guess, m_pad, sort_n, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

struct RaggedClassRun {
    uint32_t model;
    int cls;
    uint32_t offset, count;  // slice of the order array
    uint32_t guesses;        // fits in the slice that start from the device guess
    uint32_t max_points;
};

}  // namespace bfhip

using namespace bfhip;

struct bf_ragged_plan_s {
    hipStream_t stream = nullptr;
    uint32_t fits = 0, grids = 0, configs = 0;
    bool mixed = false, uncertainties = false, any_guess = false, any_p0 = false;
    uint32_t out_rows = 0;
    bf_fit_desc *d_fits = nullptr;
    bf_grid_desc *d_grids = nullptr;
    bf_config *d_configs = nullptr;
    GridInfo *d_info = nullptr;
    uint32_t *d_order = nullptr;
    uint32_t *d_cursors = nullptr;
    float *d_guess = nullptr;
    std::vector<RaggedClassRun> runs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *key[10] = {};
    hipEvent_t ev[3] = {};
    bool timed = false;
};

namespace {

void release(bf_ragged_plan p) {
    if (!p)
        return;
    if (p->exec)
        (void)hipGraphExecDestroy(p->exec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    for (void *ptr : {(void *)p->d_fits, (void *)p->d_grids, (void *)p->d_configs,
                      (void *)p->d_info, (void *)p->d_order, (void *)p->d_cursors,
                      (void *)p->d_guess})
        if (ptr)
            (void)hipFree(ptr);
    for (hipEvent_t e : p->ev)
        if (e)
            (void)hipEventDestroy(e);
    delete p;
}

template <typename T> hipError_t upload(T **dst, const T *src, size_t n, hipStream_t s) {
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), n ? n * sizeof(T) : 1);
    if (e != hipSuccess || !n)
        return e;
    e = hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

int enqueue(bf_ragged_plan p, const float *x, const float *y, const float *p0, float *params,
            int32_t *status, uint32_t *iterations, float *chi2, float *cov, float *se,
            int32_t *us, char *err, size_t errlen) {
    hipStream_t s = p->stream;
    BF_TRY(hipEventRecord(p->ev[0], s), "hipEventRecord");
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, p->d_cursors,
                       uint32_t(p->runs.size()));
    BF_TRY(hipGetLastError(), "queue reset launch");
    hipLaunchKernelGGL(grid_table_kernel, dim3(p->grids), dim3(256), 0, s, p->d_grids, x,
                       p->d_info);
    BF_TRY(hipGetLastError(), "grid table launch");
    for (const RaggedClassRun &r : p->runs) {
        if (!r.guesses)
            continue;
        const uint32_t count = r.max_points / 10 > 1 ? r.max_points / 10 : 1;
        const uint32_t sort_n = count <= 128 ? 0 : uint32_t(pow2_ceil(2 * count));
        BF_TRY(ragged_guess(r.model, p->d_fits, p->d_order + r.offset, r.count, p->d_grids,
                            p->d_info, x, y, p->d_guess, r.max_points, sort_n, s),
               "ragged guess launch");
    }
    BF_TRY(hipEventRecord(p->ev[1], s), "hipEventRecord");
    for (size_t i = 0; i < p->runs.size(); ++i) {
        const RaggedClassRun &r = p->runs[i];
        RaggedArgs a{};
        a.fits = p->d_fits;
        a.grids = p->d_grids;
        a.grid_info = p->d_info;
        a.configs = p->d_configs;
        a.order = p->d_order + r.offset;
        a.count = r.count;
        a.cursor = p->d_cursors + i;
        a.x = x;
        a.y = y;
        a.p0 = p0;
        a.guess = p->d_guess;
        a.params = params;
        a.status = status;
        a.iterations = iterations;
        a.chi2 = chi2;
        a.cov = p->uncertainties ? cov : nullptr;
        a.stderr_ = se;
        a.ustatus = us;
        const int G = kRaggedClasses[r.cls].G;
        const uint32_t groups_per_block = 256 / uint32_t(G);
        const uint32_t needed = (r.count + groups_per_block - 1) / groups_per_block;
        const uint32_t blocks = std::min<uint32_t>(needed, 256u * 8u);
        BF_TRY(ragged_lm(r.model, r.cls, p->mixed, a, blocks, s), "ragged LM launch");
    }
    BF_TRY(hipEventRecord(p->ev[2], s), "hipEventRecord");
    return BF_OK;
}

}  // namespace

extern "C" {

int bf_hip_ragged_create(bf_ragged_plan *plan, uint32_t fits, const bf_fit_desc *fit_table,
                         uint32_t grids, const bf_grid_desc *grid_table, uint32_t configs,
                         const bf_config *config_table, uint32_t numerics, int uncertainties,
                         void *stream, char *err, size_t errlen) {
    if (!plan)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "plan pointer is required");
    *plan = nullptr;
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP fitting is unavailable: " + rt.why);
    if (numerics != 0 && numerics != (BF_COMPENSATED | BF_MIXED_FP64))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "ragged batches support numerics 0 (FP32) or 5 (compensated mixed)");
    if (fits && (!fit_table || !grid_table || !config_table || !grids || !configs))
        return fail(err, errlen, BF_INVALID_ARGUMENT, "descriptor tables are required");
    for (uint32_t g = 0; g < grids; ++g)
        if (grid_table[g].points == 0)
            return fail(err, errlen, BF_INVALID_ARGUMENT, "every grid needs M>0");
    auto p = new bf_ragged_plan_s();
    p->stream = stream ? static_cast<hipStream_t>(stream) : rt.stream;
    p->fits = fits;
    p->grids = grids;
    p->configs = configs;
    p->mixed = numerics != 0;
    p->uncertainties = uncertainties != 0;
    // Bucket by (model, class); longest first inside each bucket.
    std::vector<std::vector<uint32_t>> bucket(16);
    for (uint32_t i = 0; i < fits; ++i) {
        const bf_fit_desc &d = fit_table[i];
        if (d.model > 3 || d.grid >= grids || d.config >= configs ||
            d.points != grid_table[d.grid].points) {
            release(p);
            return fail(err, errlen, BF_INVALID_ARGUMENT,
                        "fit " + std::to_string(i) +
                            ": model, grid or config out of range, or M differs from its grid");
        }
        if (d.p0_row == 0xffffffffu && d.points > 13000) {
            release(p);
            return fail(err, errlen, BF_INVALID_ARGUMENT,
                        "the device initial guess supports M up to 13,000 points; pass p0");
        }
        p->any_guess = p->any_guess || d.p0_row == 0xffffffffu;
        p->any_p0 = p->any_p0 || d.p0_row != 0xffffffffu;
        p->out_rows = std::max(p->out_rows, d.out_row + 1);
        bucket[d.model * 4 + ragged_class_of(d.points)].push_back(i);
    }
    std::vector<uint32_t> order;
    order.reserve(fits);
    for (int b = 0; b < 16; ++b) {
        auto &v = bucket[b];
        if (v.empty())
            continue;
        std::stable_sort(v.begin(), v.end(), [&](uint32_t a, uint32_t c) {
            return fit_table[a].points > fit_table[c].points;
        });
        RaggedClassRun r{uint32_t(b / 4), b % 4, uint32_t(order.size()), uint32_t(v.size()), 0, 0};
        for (uint32_t i : v) {
            r.guesses += fit_table[i].p0_row == 0xffffffffu;
            r.max_points = std::max(r.max_points, fit_table[i].points);
        }
        order.insert(order.end(), v.begin(), v.end());
        p->runs.push_back(r);
    }
    hipStream_t s = p->stream;
    auto bad = [&](hipError_t e, const char *what) {
        release(p);
        return fail(err, errlen, BF_RUNTIME_ERROR,
                    std::string(what) + " failed: " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = upload(&p->d_fits, fit_table, fits, s)) != hipSuccess)
        return bad(e, "upload of fit descriptors");
    if ((e = upload(&p->d_grids, grid_table, grids, s)) != hipSuccess)
        return bad(e, "upload of grid descriptors");
    if ((e = upload(&p->d_configs, config_table, configs, s)) != hipSuccess)
        return bad(e, "upload of configs");
    if ((e = upload(&p->d_order, order.data(), order.size(), s)) != hipSuccess)
        return bad(e, "upload of the work order");
    if ((e = hipMalloc(reinterpret_cast<void **>(&p->d_info), sizeof(GridInfo) * (grids ? grids : 1))) !=
        hipSuccess)
        return bad(e, "hipMalloc(grid info)");
    if ((e = hipMalloc(reinterpret_cast<void **>(&p->d_cursors), sizeof(uint32_t) * 16)) != hipSuccess)
        return bad(e, "hipMalloc(queue counters)");
    if (p->any_guess &&
        (e = hipMalloc(reinterpret_cast<void **>(&p->d_guess), size_t(p->out_rows) * 8 * sizeof(float))) !=
            hipSuccess)
        return bad(e, "hipMalloc(guess rows)");
    for (hipEvent_t &ev : p->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess)
            return bad(e, "hipEventCreate");
    *plan = p;
    return BF_OK;
}

int bf_hip_ragged_run(bf_ragged_plan p, const float *d_x, const float *d_y, const float *d_p0,
                      float *d_params, int32_t *d_status, uint32_t *d_iterations, float *d_chi2,
                      float *d_cov, float *d_stderr, int32_t *d_ustatus, char *err,
                      size_t errlen) {
    if (!p)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "plan is required");
    if (!p->fits)
        return BF_OK;
    if (!d_x || !d_y || !d_params || !d_status || !d_iterations || !d_chi2)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "x, y and every output must be non-null");
    if (p->any_p0 && !d_p0)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "descriptors name p0 rows but p0 is null");
    if (p->uncertainties && !(d_cov && d_stderr && d_ustatus))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "the plan computes uncertainties: covariance, stderr and status are required");
    const void *key[10] = {d_x, d_y, d_p0, d_params, d_status, d_iterations, d_chi2,
                           d_cov, d_stderr, d_ustatus};
    if (!p->exec || !std::equal(key, key + 10, p->key)) {
        if (p->exec) {
            (void)hipGraphExecDestroy(p->exec);
            (void)hipGraphDestroy(p->graph);
            p->exec = nullptr;
            p->graph = nullptr;
        }
        hipStream_t s = p->stream;
        BF_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const int rc = enqueue(p, d_x, d_y, d_p0, d_params, d_status, d_iterations, d_chi2, d_cov,
                               d_stderr, d_ustatus, err, errlen);
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(s, &graph);
        if (rc != BF_OK) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return rc;
        }
        BF_TRY(end, "hipStreamEndCapture");
        p->graph = graph;
        BF_TRY(hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        std::copy(key, key + 10, p->key);
    }
    BF_TRY(hipGraphLaunch(p->exec, p->stream), "hipGraphLaunch");
    p->timed = true;
    return BF_OK;
}

int bf_hip_ragged_timing(bf_ragged_plan p, double *times, int count) {
    for (int i = 0; i < count; ++i)
        times[i] = 0.0;
    if (!p || !p->timed)
        return BF_INVALID_ARGUMENT;
    if (hipEventSynchronize(p->ev[2]) != hipSuccess)
        return BF_RUNTIME_ERROR;
    for (int i = 0; i < count && i < 2; ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]) != hipSuccess)
            return BF_RUNTIME_ERROR;
        times[i] = ms;
    }
    return BF_OK;
}

int bf_hip_ragged_destroy(bf_ragged_plan p) {
    if (p)
        (void)hipStreamSynchronize(p->stream);
    release(p);
    return BF_OK;
}

}  // extern "C"
