// bf_ragged_runtime.hip — host side of ragged batches (bf_hip_ragged_*):
// validation, class bucketing with longest-first order, upload, and one
// captured hipGraph per pointer set: queue reset → grid table → guesses →
// persistent LM kernels, all on the plan's stream.
#include "bf_host.hpp"
#include "bf_ragged.hpp"

#include <algorithm>
#include <limits>
#include <cstdlib>
#include <numeric>
#include <vector>

namespace bfhip {

namespace {

// Queue counters of the runs cleared, and the batch's narrow-grid flag (the
// word after the counters) set; grid_table_kernel clears the flag for a grid
// beyond kFixedBaselineMaxX.
__global__ void __launch_bounds__(64) reset_kernel(uint32_t *cursors, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += 64)
        cursors[i] = 0;
    if (threadIdx.x == 0)
        cursors[kRaggedCounters] = 1;
}

__global__ void __launch_bounds__(256) grid_table_kernel(const bf_grid_desc *grids, const float *x,
                                                         GridInfo *info, uint32_t *narrow) {
    const bf_grid_desc g = grids[blockIdx.x];
    grid_block(x + g.x_offset, g.points, info + blockIdx.x);
    if (threadIdx.x == 0 && !(info[blockIdx.x].max_abs_x <= kFixedBaselineMaxX))
        *narrow = 0;  // every writer stores 0
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
                        uint32_t extra_n, hipStream_t s) {
    switch (model) {
    case 0: return launch_ragged_guess<0>(fits, order, count, grids, info, x, y, guess, m_pad, extra_n, s);
    case 1: return launch_ragged_guess<1>(fits, order, count, grids, info, x, y, guess, m_pad, extra_n, s);
    case 2: return launch_ragged_guess<2>(fits, order, count, grids, info, x, y, guess, m_pad, extra_n, s);
    case 3: return launch_ragged_guess<3>(fits, order, count, grids, info, x, y, guess, m_pad, extra_n, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

namespace {

// Stable sort of fit indices by M, largest first: a 16-bit counting sort per
// pass (one pass when every M < 65,536) for large classes, so a batch of
// millions of fits is laid out in linear time.
void sort_longest_first(std::vector<uint32_t> &v, const bf_fit_desc *t) {
    if (v.size() < 4096) {
        std::stable_sort(v.begin(), v.end(), [&](uint32_t a, uint32_t c) {
            return t[a].points > t[c].points;
        });
        return;
    }
    uint32_t top = 0;
    for (uint32_t i : v)
        top = std::max(top, t[i].points);
    std::vector<uint32_t> tmp(v.size()), count(65537);
    for (int shift = 0; shift < 32; shift += 16) {
        if (shift && (top >> 16) == 0)
            break;
        std::fill(count.begin(), count.end(), 0u);
        for (uint32_t i : v)
            ++count[(((top - t[i].points) >> shift) & 0xffffu) + 1];
        for (size_t d = 1; d < count.size(); ++d)
            count[d] += count[d - 1];
        for (uint32_t i : v)
            tmp[count[((top - t[i].points) >> shift) & 0xffffu]++] = i;
        v.swap(tmp);
    }
}

}  // namespace

std::string ragged_layout(uint32_t fits, const bf_fit_desc *fit_table, uint32_t grids,
                          const bf_grid_desc *grid_table, uint32_t configs,
                          const bf_config *config_table, RaggedLayout &out, bool split) {
    out = RaggedLayout{};
    if (fits && (!fit_table || !grid_table || !grids || !configs || !config_table))
        return "descriptor tables are required";
    for (uint32_t g = 0; g < grids; ++g)
        if (grid_table[g].points == 0)
            return "every grid needs M>0";
    // Buckets: (model, class), and for the two-peak model a second set of
    // classes for fits that fix both baseline parameters.
    std::vector<std::vector<uint32_t>> bucket(5 * kRaggedClassCount);
    for (uint32_t i = 0; i < fits; ++i) {
        const bf_fit_desc &d = fit_table[i];
        if (d.model > 3 || d.grid >= grids || d.config >= configs ||
            d.points != grid_table[d.grid].points)
            return "fit " + std::to_string(i) +
                   ": model, grid or config out of range, or M differs from its grid";
        if (d.out_row == 0xffffffffu)
            return "fit " + std::to_string(i) + ": output row 0xffffffff is out of range";
        out.any_guess = out.any_guess || d.p0_row == 0xffffffffu;
        out.any_p0 = out.any_p0 || d.p0_row != 0xffffffffu;
        out.out_rows = std::max(out.out_rows, uint64_t(d.out_row) + 1);
        const int cls = ragged_class_of(d.model, d.points);
        bucket[(fixed_baseline(d.model, config_table[d.config]) ? 4u : d.model) *
                   kRaggedClassCount + cls]
            .push_back(i);
    }
    // Two fits writing one output row would race on it (and on its guess
    // row): every fit needs a row of its own.
    {
        std::vector<uint32_t> rows(fits);
        for (uint32_t i = 0; i < fits; ++i)
            rows[i] = fit_table[i].out_row;
        std::sort(rows.begin(), rows.end());
        for (uint32_t i = 1; i < fits; ++i)
            if (rows[i] == rows[i - 1])
                return "output row " + std::to_string(rows[i]) + " is named by more than one fit";
    }
    out.order.reserve(fits);
    double total = 0.0;
    for (uint32_t i = 0; i < fits; ++i)
        total += double(fit_table[i].points) * kRaggedWeight[fit_table[i].model];
    // A bucket heavier than an eighth of the batch becomes several runs over
    // consecutive slices of its longest-first order, so that the branches of
    // the graph can balance (a fit's result does not depend on its run).
    // Pieces hold at least 2^26 weighted points (about 0.7 ms of work on
    // MI355X), so small batches are not cut into launches that cost more
    // than the balance gains.
    const double piece_cost =
        split ? std::max(total / 8.0, 67108864.0) : std::numeric_limits<double>::infinity();
    auto add_run = [&](uint32_t model, int cls, bool fixb, const uint32_t *v, size_t n) {
        RaggedRun r{model, cls, fixb, uint32_t(out.order.size()), uint32_t(n), 0, 0, 0.0, {}};
        double points = 0.0;
        for (size_t k = 0; k < n; ++k) {
            r.guesses += fit_table[v[k]].p0_row == 0xffffffffu;
            r.max_points = std::max(r.max_points, fit_table[v[k]].points);
            points += fit_table[v[k]].points;
        }
        // v is longest first, so each octave of M is a contiguous slice.
        for (size_t a = 0; a < n;) {
            const uint32_t top = fit_table[v[a]].points;
            const uint32_t below = guess_slice_floor(model, top);
            size_t end = a;
            uint32_t guessed = 0;
            for (; end < n && fit_table[v[end]].points > below; ++end)
                guessed += fit_table[v[end]].p0_row == 0xffffffffu;
            if (guessed)
                r.guess_slices.push_back({uint32_t(a), uint32_t(end - a), top});
            a = end;
        }
        r.cost = points * kRaggedWeight[model];
        out.order.insert(out.order.end(), v, v + n);
        out.runs.push_back(std::move(r));
    };
    for (int b = 0; b < 5 * kRaggedClassCount; ++b) {
        auto &v = bucket[b];
        if (v.empty())
            continue;
        sort_longest_first(v, fit_table);
        const bool fixb = b >= 4 * kRaggedClassCount;
        const uint32_t model = fixb ? uint32_t(BF_TWO_LORENTZIAN) : uint32_t(b / kRaggedClassCount);
        const int cls = b % kRaggedClassCount;
        size_t a = 0;
        double acc = 0.0;
        for (size_t k = 0; k < v.size(); ++k) {
            acc += double(fit_table[v[k]].points) * kRaggedWeight[model];
            if (acc >= piece_cost && k + 1 < v.size()) {
                add_run(model, cls, fixb, v.data() + a, k + 1 - a);
                a = k + 1;
                acc = 0.0;
            }
        }
        add_run(model, cls, fixb, v.data() + a, v.size() - a);
    }
    if (out.runs.size() > size_t(kRaggedCounters))
        return "internal: more runs than queue counters";
    out.cost_order.resize(out.runs.size());
    std::iota(out.cost_order.begin(), out.cost_order.end(), size_t(0));
    std::stable_sort(out.cost_order.begin(), out.cost_order.end(),
                     [&](size_t a, size_t b) { return out.runs[a].cost > out.runs[b].cost; });
    return {};
}

hipError_t ragged_prepare(const RaggedLayout &layout, bool mixed) {
    hipFuncAttributes attr;
    hipError_t e;
    if ((e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&reset_kernel))) != hipSuccess ||
        (e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&grid_table_kernel))) !=
            hipSuccess)
        return e;
    for (const RaggedRun &r : layout.runs) {
        const bool guess = r.guesses != 0;
        switch (r.model) {
        case 0: e = prepare_ragged<0>(r.cls, mixed, guess); break;
        case 1: e = prepare_ragged<1>(r.cls, mixed, guess); break;
        case 2: e = prepare_ragged<2>(r.cls, mixed, guess); break;
        case 3: e = prepare_ragged<3>(r.cls, mixed, guess); break;
        default: e = hipErrorInvalidValue;
        }
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

namespace {

// Guess (when any fit of the run starts from it) and LM of run i on `s`.
hipError_t enqueue_run(const RaggedLayout &layout, size_t i, const RaggedDevice &dev, bool mixed,
                       const RaggedIO &io, hipStream_t s) {
    const RaggedRun &r = layout.runs[i];
    hipError_t e;
    for (const GuessSlice &g : r.guess_slices) {
        const uint32_t extra_n = guess_extra_floats(int(r.model), g.max_points);
        if ((e = ragged_guess(r.model, dev.fits, dev.order + r.offset + g.offset, g.count,
                              dev.grids, dev.info, io.x, io.y, dev.guess, g.max_points, extra_n,
                              s)) != hipSuccess)
            return e;
    }
    RaggedArgs a{};
    a.fits = dev.fits;
    a.grids = dev.grids;
    a.grid_info = dev.info;
    a.configs = dev.configs;
    a.order = dev.order + r.offset;
    a.count = r.count;
    a.cursor = dev.cursors + i;
    a.flat_check = r.guesses < r.count;
    a.x = io.x;
    a.y = io.y;
    a.p0 = io.p0;
    a.guess = dev.guess;
    a.params = io.params;
    a.status = io.status;
    a.iterations = io.iterations;
    a.chi2 = io.chi2;
    a.cov = io.cov;
    a.stderr_ = io.stderr_;
    a.ustatus = io.ustatus;
    a.narrow = dev.cursors + kRaggedCounters;
    a.fixed_baseline = r.fixed_baseline;
    a.wide_only = 0;
    const uint32_t per_wave = 64 / uint32_t(kRaggedClasses[r.cls].G);
    const uint32_t waves = (r.count + per_wave - 1) / per_wave;
    const uint32_t blocks = std::min<uint32_t>((waves + 3) / 4, 256u * 8u);
    return ragged_lm(r.model, r.cls, mixed, a, blocks, s);
}

}  // namespace

hipError_t ragged_enqueue(const RaggedLayout &layout, const RaggedDevice &dev, bool mixed,
                          const RaggedIO &io, hipStream_t s, const RaggedFork *fork) {
    hipError_t e;
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, dev.cursors,
                       uint32_t(layout.runs.size()));
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(grid_table_kernel, dim3(dev.grid_count), dim3(256), 0, s, dev.grids, io.x,
                       dev.info, dev.cursors + kRaggedCounters);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    const int branches = fork ? std::min<int>(fork->streams, int(layout.runs.size())) : 0;
    if (branches < 2) {
        for (size_t k = 0; k < layout.runs.size(); ++k)
            if ((e = enqueue_run(layout, layout.cost_order[k], dev, mixed, io, s)) != hipSuccess)
                return e;
        return hipSuccess;
    }
    // Fork: every branch waits for the prologue, runs its share of the runs
    // (heaviest first, each to the least loaded branch) and joins back.
    if ((e = hipEventRecord(fork->fork, s)) != hipSuccess)
        return e;
    for (int b = 0; b < branches; ++b)
        if ((e = hipStreamWaitEvent(fork->side[b], fork->fork, 0)) != hipSuccess)
            return e;
    std::vector<double> load(branches, 0.0);
    for (size_t k = 0; k < layout.runs.size(); ++k) {
        const size_t i = layout.cost_order[k];
        const int b = int(std::min_element(load.begin(), load.end()) - load.begin());
        load[b] += layout.runs[i].cost;
        if ((e = enqueue_run(layout, i, dev, mixed, io, fork->side[b])) != hipSuccess)
            return e;
    }
    for (int b = 0; b < branches; ++b)
        if ((e = hipEventRecord(fork->join[b], fork->side[b])) != hipSuccess ||
            (e = hipStreamWaitEvent(s, fork->join[b], 0)) != hipSuccess)
            return e;
    return hipSuccess;
}

}  // namespace bfhip

using namespace bfhip;

struct bf_ragged_plan_s {
    hipStream_t stream = nullptr;
    uint32_t fits = 0, grids = 0, configs = 0;
    bool mixed = false, uncertainties = false, any_guess = false, any_p0 = false;
    RaggedLayout layout;
    bf_fit_desc *d_fits = nullptr;
    bf_grid_desc *d_grids = nullptr;
    bf_config *d_configs = nullptr;
    GridInfo *d_info = nullptr;
    uint32_t *d_order = nullptr;
    uint32_t *d_cursors = nullptr;
    float *d_guess = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *key[10] = {};
    hipEvent_t ev[2] = {};
    bool timed = false;
    // Branches of the captured graph: every (model, class) run is a chain
    // guess → LM on one of these streams, so the runs' tails overlap.
    static constexpr int kBranches = 4;
    hipStream_t side[kBranches] = {};
    hipEvent_t fork = nullptr, join[kBranches] = {};
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
    for (hipEvent_t e : p->join)
        if (e)
            (void)hipEventDestroy(e);
    if (p->fork)
        (void)hipEventDestroy(p->fork);
    for (hipStream_t st : p->side)
        if (st)
            (void)hipStreamDestroy(st);
    delete p;
}

template <typename T> hipError_t upload(T **dst, const T *src, size_t n, hipStream_t s) {
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), n ? n * sizeof(T) : 1);
    if (e != hipSuccess || !n)
        return e;
    e = hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// Concurrent branches of a plan's graph: BF_HIP_RAGGED_BRANCHES (1 to 4,
// default 4). One branch runs the runs one after another, so that a profile
// shows each kernel's own time.
int ragged_branches() {
    static const int n = [] {
        const char *v = std::getenv("BF_HIP_RAGGED_BRANCHES");
        const int k = v ? std::atoi(v) : bf_ragged_plan_s::kBranches;
        return k < 1 ? 1 : k > bf_ragged_plan_s::kBranches ? bf_ragged_plan_s::kBranches : k;
    }();
    return n;
}

int enqueue(bf_ragged_plan p, const float *x, const float *y, const float *p0, float *params,
            int32_t *status, uint32_t *iterations, float *chi2, float *cov, float *se,
            int32_t *us, char *err, size_t errlen) {
    const RaggedDevice dev{p->d_fits, p->d_grids, p->d_configs, p->d_info,
                           p->d_order, p->d_cursors, p->d_guess, p->grids};
    // Events inside the capture only mark the fork and join points of the
    // branches (an event recorded on a capturing stream is not replayed).
    const RaggedIO io{x, y, p0, params, status, iterations, chi2,
                      p->uncertainties ? cov : nullptr, se, us};
    const RaggedFork fork{ragged_branches(), p->side, p->fork, p->join};
    BF_TRY(ragged_enqueue(p->layout, dev, p->mixed, io, p->stream, &fork),
           "ragged batch enqueue");
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
    // The tables are checked on the host before any device work.
    if (numerics != 0 && numerics != (BF_COMPENSATED | BF_MIXED_FP64))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "ragged batches support numerics 0 (FP32) or 5 (compensated mixed)");
    if (fits && !config_table)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "descriptor tables are required");
    auto p = new bf_ragged_plan_s();
    const std::string why =
        ragged_layout(fits, fit_table, grids, grid_table, configs, config_table, p->layout, true);
    if (!why.empty()) {
        delete p;
        return fail(err, errlen, BF_INVALID_ARGUMENT, why);
    }
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK) {
        delete p;
        return fail(err, errlen, BF_UNAVAILABLE, "HIP fitting is unavailable: " + rt.why);
    }
    p->stream = stream ? static_cast<hipStream_t>(stream) : rt.stream;
    p->fits = fits;
    p->grids = grids;
    p->configs = configs;
    p->mixed = numerics != 0;
    p->uncertainties = uncertainties != 0;
    p->any_guess = p->layout.any_guess;
    p->any_p0 = p->layout.any_p0;
    const std::vector<uint32_t> &order = p->layout.order;
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
    if ((e = hipMalloc(reinterpret_cast<void **>(&p->d_cursors),
                       sizeof(uint32_t) * kRaggedCounterWords)) != hipSuccess)
        return bad(e, "hipMalloc(queue counters)");
    if (p->any_guess &&
        (e = hipMalloc(reinterpret_cast<void **>(&p->d_guess),
                       size_t(p->layout.out_rows) * 8 * sizeof(float))) != hipSuccess)
        return bad(e, "hipMalloc(guess rows)");
    for (hipEvent_t &ev : p->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess)
            return bad(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming)) != hipSuccess)
        return bad(e, "hipEventCreate");
    for (int b = 0; b < bf_ragged_plan_s::kBranches; ++b)
        if ((e = hipStreamCreateWithFlags(&p->side[b], hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&p->join[b], hipEventDisableTiming)) != hipSuccess)
            return bad(e, "branch stream");
    if ((e = ragged_prepare(p->layout, p->mixed)) != hipSuccess)
        return bad(e, "loading the ragged kernels");
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
    BF_TRY(hipEventRecord(p->ev[0], p->stream), "hipEventRecord");
    BF_TRY(hipGraphLaunch(p->exec, p->stream), "hipGraphLaunch");
    BF_TRY(hipEventRecord(p->ev[1], p->stream), "hipEventRecord");
    p->timed = true;
    return BF_OK;
}

int bf_hip_ragged_timing(bf_ragged_plan p, double *times, int count) {
    for (int i = 0; i < count; ++i)
        times[i] = 0.0;
    if (!p || !p->timed)
        return BF_INVALID_ARGUMENT;
    if (hipEventSynchronize(p->ev[1]) != hipSuccess)
        return BF_RUNTIME_ERROR;
    float ms = 0.0f;
    if (count < 1 || hipEventElapsedTime(&ms, p->ev[0], p->ev[1]) != hipSuccess)
        return BF_RUNTIME_ERROR;
    times[0] = ms;
    return BF_OK;
}

int bf_hip_ragged_destroy(bf_ragged_plan p) {
    if (p)
        (void)hipStreamSynchronize(p->stream);
    release(p);
    return BF_OK;
}

}  // extern "C"
