// bf_serve.hip — continuous-batching fit server (bf_serve_*).
//
// Requests (one model, one grid, N spectra, their own bounds/options) are
// queued by any number of host threads. A scheduler thread admits pending
// requests in arrival order into the next free slot while the bytes the
// admitted batches hold in HBM stay within the budget, packs them into one
// ragged batch (each request one grid and one config row), and enqueues
// staging copies, guesses and LM on that slot's stream. Up to `slots` batches
// are on the device at once, so requests that arrive while others run are
// admitted as slots and memory free up, and their work overlaps on the
// device. A completion thread waits for slots in order, scatters outputs to
// the requests and releases their bytes. Each fit's arithmetic depends only
// on its own data, model and M, so a request's result is the same alone and
// in any mix.
//
// The reference serialises native fitting behind one global mutex
// (python/bindings.cpp:13,66) and streams NPY chunks through a ≤3-slot ring
// (src/metal/pipeline.cpp:99-169); this server generalises that ring to
// arbitrary request streams with admission control.
#include "bf_host.hpp"
#include "bf_ragged.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

namespace bfhip {
namespace {

using clock_type = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    uint32_t model = 0, n = 0, m = 0;
    bf_config cfg{};
    bool device = false;
    const float *x = nullptr, *y = nullptr, *p0 = nullptr;
    float *params = nullptr, *chi2 = nullptr, *cov = nullptr, *se = nullptr;
    int32_t *status = nullptr, *us = nullptr;
    uint32_t *iterations = nullptr;
    uint64_t bytes = 0;
    uint32_t batch = 0;
    clock_type::time_point submitted, admitted, finished;
    int rc = BF_OK;
    std::string message;
    bool done = false;
    // placement inside the batch
    uint64_t x_off = 0, y_off = 0, p0_row = 0, out_row = 0;
};

// Grow-only device buffer.
struct DeviceArena {
    char *ptr = nullptr;
    size_t size = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= size)
            return hipSuccess;
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        size = 0;
        const hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), bytes);
        if (e == hipSuccess)
            size = bytes;
        return e;
    }
    ~DeviceArena() {
        if (ptr)
            (void)hipFree(ptr);
    }
};

// Grow-only pinned host buffer (allocated pinned, never registered).
struct PinnedArena {
    char *ptr = nullptr;
    size_t size = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= size)
            return hipSuccess;
        if (ptr)
            (void)hipHostFree(ptr);
        ptr = nullptr;
        size = 0;
        const hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&ptr), bytes);
        if (e == hipSuccess)
            size = bytes;
        return e;
    }
    ~PinnedArena() {
        if (ptr)
            (void)hipHostFree(ptr);
    }
};

constexpr size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

struct Slot {
    static constexpr int kBranches = 4;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipStream_t side[kBranches] = {};  // branches of a batch's runs
    hipEvent_t fork = nullptr, join[kBranches] = {};
    DeviceArena dev;
    PinnedArena host;
    std::vector<Request *> batch;
    uint64_t bytes = 0;
    uint32_t rows = 0;
    bool uncertainties = false;
    bool busy = false;
    int rc = BF_OK;
    std::string message;
    // pinned output regions
    size_t out_params = 0, out_status = 0, out_iter = 0, out_chi2 = 0, out_cov = 0, out_se = 0,
           out_us = 0;
};

}  // namespace

struct Server {
    bf_serve_options opt{};
    bool mixed = true;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_complete;
    std::deque<Request *> pending;
    std::deque<Slot *> in_flight;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> table;
    uint64_t next_ticket = 1, bytes_in_flight = 0, batches = 0, fits_done = 0, requests_done = 0,
             max_bytes = 0;
    uint32_t max_batches_in_flight = 0;
    bool stopping = false;
    std::thread scheduler, completer;

    void schedule_loop();
    void complete_loop();
    void launch(Slot &slot);
    void finish_batch(Slot &slot);
};

namespace {

uint64_t request_bytes(const Request &r) {
    const uint64_t n = r.n, m = r.m;
    // Host requests are staged into the arena; device requests are read in
    // place, so only their tables and outputs count.
    uint64_t b = (r.device ? 0 : 4 * (m + n * m)) + 4 * n * (8 + 8 + 3) + 32 * n + 160;
    if (r.p0)
        b += 32 * n;
    if (r.cov)
        b += 4 * n * (64 + 8 + 1);
    return b;
}

}  // namespace

// One request of a batch as the device sees it: where its fits' rows,
// observations, work order and p0 rows lie, and (device requests) where its
// outputs go.
struct ServeReq {
    uint64_t y_off;      // first observation, in floats from the batch's y base
    uint32_t m, model;
    uint32_t first_row;  // batch row of its first fit (rows of a batch increase with the request)
    uint32_t n;
    uint32_t order_pos;  // position of its first fit in the work order
    uint32_t p0_row;     // batch p0 row of its first fit, or 0xffffffff (device guess)
    uint32_t P, pad;
    const float *p0_dev;  // a device request's own p0 [n, P], gathered into rows of 8
    float *params, *chi2;  // a device request's outputs (null for host requests)
    int32_t *status;
    uint32_t *iterations;
};

// The request that holds batch row i: the last one whose first row is at or
// below i (every request of a batch has at least one fit).
__device__ __forceinline__ uint32_t request_of(const ServeReq *req, uint32_t nreq, uint32_t i) {
    uint32_t lo = 0, hi = nreq - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (req[mid].first_row <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Fit descriptors, the work order and device requests' p0 rows of a batch,
// one thread per row, from the request table.
__global__ void __launch_bounds__(256) serve_expand_kernel(const ServeReq *req, uint32_t nreq,
                                                           uint32_t rows, bf_fit_desc *fits,
                                                           uint32_t *order, float *p0_rows) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows)
        return;
    const uint32_t q = request_of(req, nreq, i);
    const ServeReq &r = req[q];
    const uint32_t k = i - r.first_row;
    bf_fit_desc d;
    d.y_offset = r.y_off + uint64_t(k) * r.m;
    d.points = r.m;
    d.grid = q;
    d.model = r.model;
    d.config = q;
    d.p0_row = r.p0_row == 0xffffffffu ? 0xffffffffu : r.p0_row + k;
    d.out_row = i;
    fits[i] = d;
    order[r.order_pos + k] = i;
    if (r.p0_dev)
        for (uint32_t p = 0; p < r.P; ++p)
            p0_rows[size_t(d.p0_row) * 8 + p] = r.p0_dev[size_t(k) * r.P + p];
}

// Outputs of the device requests of a batch, one thread per row, from the
// batch's rows (params in rows of 8).
__global__ void __launch_bounds__(256) serve_scatter_kernel(const ServeReq *req, uint32_t nreq,
                                                            uint32_t rows, const float *params,
                                                            const int32_t *status,
                                                            const uint32_t *iters,
                                                            const float *chi2) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows)
        return;
    const ServeReq &r = req[request_of(req, nreq, i)];
    if (!r.params)
        return;
    const uint32_t k = i - r.first_row;
    for (uint32_t p = 0; p < r.P; ++p)
        r.params[size_t(k) * r.P + p] = params[size_t(i) * 8 + p];
    r.status[k] = status[i];
    r.iterations[k] = iters[i];
    r.chi2[k] = chi2[i];
}

// Pack and enqueue one admitted batch on its slot (called without the lock).
// The host works per request: it stages host requests' arrays and the small
// tables (one upload), buckets the requests by (model, mapping class) and
// orders each bucket longest first, which gives the runs and their guess
// slices (a request's fits share M, so it never straddles a slice); the
// device expands the requests into fit descriptors and the work order, and
// scatters device requests' outputs, in one kernel each.
void Server::launch(Slot &slot) {
    auto &batch = slot.batch;
    uint64_t xs = 0, ys = 0, p0s = 0, rows = 0;
    bool unc = false, any_host = false, any_device = false;
    for (Request *r : batch) {
        r->out_row = rows;
        if (!r->device) {
            r->x_off = xs;
            r->y_off = ys;
            xs += r->m;
            ys += uint64_t(r->n) * r->m;
            any_host = true;
        } else {
            any_device = true;
        }
        if (r->p0) {
            r->p0_row = p0s;
            p0s += r->n;
        }
        rows += r->n;
        unc = unc || r->cov;
    }
    slot.rows = uint32_t(rows);
    slot.uncertainties = unc;
    const uint32_t nreq = uint32_t(batch.size());
    // Both arenas begin with the uploaded prefix x | y | p0 | grids | configs
    // | requests (host requests' arrays staged, device requests' p0 rows
    // gathered on the device afterwards); the device goes on with its work
    // tables and the outputs, the pinned arena with the same outputs for
    // host requests (one download).
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes);
        return at;
    };
    const size_t o_x = carve(xs * 4), o_y = carve(ys * 4), o_p0 = carve(p0s * 32);
    const size_t o_grids = carve(nreq * sizeof(bf_grid_desc));
    const size_t o_cfg = carve(nreq * sizeof(bf_config)), o_req = carve(nreq * sizeof(ServeReq));
    const size_t upload = off;
    const size_t d_fits = carve(rows * sizeof(bf_fit_desc)), d_order = carve(rows * 4);
    const size_t d_info = carve(nreq * sizeof(GridInfo)), d_cur = carve(kRaggedCounterWords * 4);
    const size_t d_guess = carve(rows * 32);
    const size_t d_out = off;
    const size_t r_params = carve(rows * 32) - d_out, r_status = carve(rows * 4) - d_out,
                 r_iter = carve(rows * 4) - d_out, r_chi2 = carve(rows * 4) - d_out;
    const size_t r_cov = unc ? carve(rows * 256) - d_out : 0,
                 r_se = unc ? carve(rows * 32) - d_out : 0, r_us = unc ? carve(rows * 4) - d_out : 0;
    const size_t out_bytes = off - d_out;
    const size_t device_bytes = off;
    const size_t h_out = upload;  // the pinned arena: the prefix, then the outputs
    const size_t host_bytes = upload + (any_host ? out_bytes : 0);
    slot.out_params = h_out + r_params;
    slot.out_status = h_out + r_status;
    slot.out_iter = h_out + r_iter;
    slot.out_chi2 = h_out + r_chi2;
    slot.out_cov = unc ? h_out + r_cov : 0;
    slot.out_se = unc ? h_out + r_se : 0;
    slot.out_us = unc ? h_out + r_us : 0;
    auto failed = [&](hipError_t e, const char *what) {
        slot.rc = BF_RUNTIME_ERROR;
        slot.message = std::string(what) + " failed: " + hipGetErrorString(e);
        // Copies already enqueued may still read the staging buffers: drain
        // the slot's stream before the next batch reuses them.
        if (slot.stream)
            (void)hipStreamSynchronize(slot.stream);
    };
    hipError_t e;
    if ((e = slot.dev.reserve(device_bytes)) != hipSuccess)
        return failed(e, "hipMalloc(batch arena)");
    if ((e = slot.host.reserve(host_bytes)) != hipSuccess)
        return failed(e, "hipHostMalloc(batch staging)");
    char *H = slot.host.ptr, *D = slot.dev.ptr;
    hipStream_t s = slot.stream;
    // Device requests are read in place: x and y offsets of every request
    // are taken from the lowest address among the staged region and the
    // device arrays (all float arrays, so every difference is whole floats).
    uintptr_t xbase = any_host ? uintptr_t(D + o_x) : UINTPTR_MAX;
    uintptr_t ybase = any_host ? uintptr_t(D + o_y) : UINTPTR_MAX;
    for (Request *r : batch)
        if (r->device) {
            xbase = std::min(xbase, uintptr_t(r->x));
            ybase = std::min(ybase, uintptr_t(r->y));
        }
    // Requests by (model, mapping class), each bucket longest first (stable
    // in arrival order): the runs of the batch, one per bucket; two-peak
    // requests that fix both baseline parameters have buckets of their own.
    std::vector<std::vector<uint32_t>> bucket(5 * kRaggedClassCount);
    for (uint32_t q = 0; q < nreq; ++q) {
        const Request &r = *batch[q];
        bucket[(fixed_baseline(r.model, r.cfg) ? 4u : r.model) * kRaggedClassCount +
               ragged_class_of(r.model, r.m)]
            .push_back(q);
    }
    auto *grids = reinterpret_cast<bf_grid_desc *>(H + o_grids);
    auto *cfgs = reinterpret_cast<bf_config *>(H + o_cfg);
    auto *reqs = reinterpret_cast<ServeReq *>(H + o_req);
    RaggedLayout layout;
    uint32_t pos = 0;
    for (int b = 0; b < 5 * kRaggedClassCount; ++b) {
        auto &v = bucket[b];
        if (v.empty())
            continue;
        std::stable_sort(v.begin(), v.end(),
                         [&](uint32_t a, uint32_t c) { return batch[a]->m > batch[c]->m; });
        const bool fixb = b >= 4 * kRaggedClassCount;
        const uint32_t model = fixb ? uint32_t(BF_TWO_LORENTZIAN) : uint32_t(b / kRaggedClassCount);
        RaggedRun run{model, b % kRaggedClassCount, fixb, pos, 0, 0, 0, 0.0, {}};
        for (size_t a = 0; a < v.size();) {
            const uint32_t top = batch[v[a]]->m;
            const uint32_t below = guess_slice_floor(model, top);
            const uint32_t slice_at = pos - run.offset;
            uint32_t count = 0, guessed = 0;
            for (; a < v.size() && batch[v[a]]->m > below; ++a) {
                const Request &r = *batch[v[a]];
                reqs[v[a]].order_pos = pos;
                pos += r.n;
                count += r.n;
                guessed += r.p0 ? 0 : r.n;
                run.cost += double(r.n) * r.m * kRaggedWeight[model];
            }
            if (guessed)
                run.guess_slices.push_back({slice_at, count, top});
            run.guesses += guessed;
            run.count += count;
            run.max_points = std::max(run.max_points, top);
        }
        layout.runs.push_back(std::move(run));
    }
    layout.cost_order.resize(layout.runs.size());
    std::iota(layout.cost_order.begin(), layout.cost_order.end(), size_t(0));
    std::stable_sort(layout.cost_order.begin(), layout.cost_order.end(), [&](size_t a, size_t c) {
        return layout.runs[a].cost > layout.runs[c].cost;
    });
    layout.out_rows = rows;
    for (uint32_t q = 0; q < nreq; ++q) {
        const Request &r = *batch[q];
        const uintptr_t xa = r.device ? uintptr_t(r.x) : uintptr_t(D + o_x + r.x_off * 4);
        const uintptr_t ya = r.device ? uintptr_t(r.y) : uintptr_t(D + o_y + r.y_off * 4);
        grids[q] = bf_grid_desc{(xa - xbase) / 4, r.m, 0};
        cfgs[q] = r.cfg;
        ServeReq &t = reqs[q];
        t.y_off = (ya - ybase) / 4;
        t.m = r.m;
        t.model = r.model;
        t.first_row = uint32_t(r.out_row);
        t.n = r.n;
        t.p0_row = r.p0 ? uint32_t(r.p0_row) : 0xffffffffu;
        t.P = uint32_t(kParams[r.model]);
        t.pad = 0;
        t.p0_dev = r.device ? r.p0 : nullptr;
        t.params = r.device ? r.params : nullptr;
        t.chi2 = r.device ? r.chi2 : nullptr;
        t.status = r.device ? r.status : nullptr;
        t.iterations = r.device ? r.iterations : nullptr;
        if (r.device)
            continue;
        std::memcpy(H + o_x + r.x_off * 4, r.x, size_t(r.m) * 4);
        std::memcpy(H + o_y + r.y_off * 4, r.y, size_t(r.n) * r.m * 4);
        if (r.p0) {
            const int P = kParams[r.model];
            float *dst = reinterpret_cast<float *>(H + o_p0) + r.p0_row * 8;
            for (uint32_t k = 0; k < r.n; ++k)
                std::memcpy(dst + size_t(k) * 8, r.p0 + size_t(k) * P, size_t(P) * 4);
        }
    }
    if ((e = hipMemcpyAsync(D, H, upload, hipMemcpyHostToDevice, s)) != hipSuccess)
        return failed(e, "staging copy");
    const auto *dreq = reinterpret_cast<const ServeReq *>(D + o_req);
    const uint32_t blocks = uint32_t((rows + 255) / 256);
    hipLaunchKernelGGL(serve_expand_kernel, dim3(blocks), dim3(256), 0, s, dreq, nreq,
                       uint32_t(rows), reinterpret_cast<bf_fit_desc *>(D + d_fits),
                       reinterpret_cast<uint32_t *>(D + d_order),
                       reinterpret_cast<float *>(D + o_p0));
    if ((e = hipGetLastError()) != hipSuccess)
        return failed(e, "batch expansion");
    const RaggedDevice dev{reinterpret_cast<bf_fit_desc *>(D + d_fits),
                           reinterpret_cast<bf_grid_desc *>(D + o_grids),
                           reinterpret_cast<bf_config *>(D + o_cfg),
                           reinterpret_cast<GridInfo *>(D + d_info),
                           reinterpret_cast<uint32_t *>(D + d_order),
                           reinterpret_cast<uint32_t *>(D + d_cur),
                           reinterpret_cast<float *>(D + d_guess), nreq};
    char *O = D + d_out;
    auto F = [&](size_t at) { return reinterpret_cast<float *>(O + at); };
    const RaggedIO io{reinterpret_cast<const float *>(xbase),
                      reinterpret_cast<const float *>(ybase),
                      reinterpret_cast<const float *>(D + o_p0), F(r_params),
                      reinterpret_cast<int32_t *>(O + r_status),
                      reinterpret_cast<uint32_t *>(O + r_iter), F(r_chi2),
                      unc ? F(r_cov) : nullptr, unc ? F(r_se) : nullptr,
                      unc ? reinterpret_cast<int32_t *>(O + r_us) : nullptr};
    const RaggedFork fork{Slot::kBranches, slot.side, slot.fork, slot.join};
    if ((e = ragged_enqueue(layout, dev, mixed, io, s, &fork)) != hipSuccess)
        return failed(e, "batch enqueue");
    if (any_device) {
        hipLaunchKernelGGL(serve_scatter_kernel, dim3(blocks), dim3(256), 0, s, dreq, nreq,
                           uint32_t(rows), F(r_params),
                           reinterpret_cast<const int32_t *>(O + r_status),
                           reinterpret_cast<const uint32_t *>(O + r_iter), F(r_chi2));
        if ((e = hipGetLastError()) != hipSuccess)
            return failed(e, "device output scatter");
    }
    if (any_host &&
        (e = hipMemcpyAsync(H + h_out, O, out_bytes, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return failed(e, "output copy");
    if ((e = hipEventRecord(slot.done, s)) != hipSuccess)
        return failed(e, "hipEventRecord");
}

// Scatter the batch's outputs to its requests (called without the lock).
void Server::finish_batch(Slot &slot) {
    const hipError_t e = hipEventSynchronize(slot.done);
    if (slot.rc == BF_OK && e != hipSuccess) {
        slot.rc = BF_RUNTIME_ERROR;
        slot.message = std::string("batch completion failed: ") + hipGetErrorString(e);
    }
    const char *H = slot.host.ptr;
    for (Request *r : slot.batch) {
        r->rc = slot.rc;
        r->message = slot.message;
        if (slot.rc != BF_OK)
            continue;
        const int P = kParams[r->model];
        const float *params = reinterpret_cast<const float *>(H + slot.out_params);
        const int32_t *status = reinterpret_cast<const int32_t *>(H + slot.out_status);
        const uint32_t *iters = reinterpret_cast<const uint32_t *>(H + slot.out_iter);
        const float *chi2 = reinterpret_cast<const float *>(H + slot.out_chi2);
        const uint64_t r0 = r->out_row;
        if (r->device) {
            // Already copied device to device in stream order.
        } else {
            for (uint32_t k = 0; k < r->n; ++k)
                std::memcpy(r->params + size_t(k) * P, params + (r0 + k) * 8, size_t(P) * 4);
            std::memcpy(r->status, status + r0, size_t(r->n) * 4);
            std::memcpy(r->iterations, iters + r0, size_t(r->n) * 4);
            std::memcpy(r->chi2, chi2 + r0, size_t(r->n) * 4);
            if (r->cov) {
                const float *cov = reinterpret_cast<const float *>(H + slot.out_cov);
                const float *se = reinterpret_cast<const float *>(H + slot.out_se);
                const int32_t *us = reinterpret_cast<const int32_t *>(H + slot.out_us);
                for (uint32_t k = 0; k < r->n; ++k) {
                    std::memcpy(r->cov + size_t(k) * P * P, cov + (r0 + k) * 64, size_t(P * P) * 4);
                    std::memcpy(r->se + size_t(k) * P, se + (r0 + k) * 8, size_t(P) * 4);
                }
                std::memcpy(r->us, us + r0, size_t(r->n) * 4);
            }
        }
    }
}

void Server::schedule_loop() {
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
        cv_sched.wait(lock, [&] {
            if (stopping && pending.empty())
                return true;
            if (pending.empty())
                return false;
            for (auto &s : slots)
                if (!s->busy)
                    return bytes_in_flight + pending.front()->bytes <= opt.hbm_budget ||
                           bytes_in_flight == 0;
            return false;
        });
        if (stopping && pending.empty())
            return;
        Slot *slot = nullptr;
        for (auto &s : slots)
            if (!s->busy) {
                slot = s.get();
                break;
            }
        // Admit in arrival order within the budget and the batch cap.
        slot->batch.clear();
        slot->bytes = 0;
        slot->rc = BF_OK;
        slot->message.clear();
        uint64_t fits = 0;
        const auto now = clock_type::now();
        while (!pending.empty()) {
            Request *r = pending.front();
            const bool first = slot->batch.empty();
            if (!first && (bytes_in_flight + slot->bytes + r->bytes > opt.hbm_budget ||
                           fits + r->n > opt.max_batch_fits))
                break;
            if (first && bytes_in_flight + r->bytes > opt.hbm_budget && bytes_in_flight != 0)
                break;
            pending.pop_front();
            r->admitted = now;
            r->batch = uint32_t(batches);
            slot->batch.push_back(r);
            slot->bytes += r->bytes;
            fits += r->n;
        }
        if (slot->batch.empty())
            continue;
        slot->busy = true;
        bytes_in_flight += slot->bytes;
        max_bytes = std::max(max_bytes, bytes_in_flight);
        ++batches;
        lock.unlock();
        launch(*slot);
        lock.lock();
        in_flight.push_back(slot);
        uint32_t busy = 0;
        for (auto &s : slots)
            busy += s->busy;
        max_batches_in_flight = std::max(max_batches_in_flight, busy);
        cv_complete.notify_one();
    }
}

void Server::complete_loop() {
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
        cv_complete.wait(lock, [&] { return !in_flight.empty() || (stopping && pending.empty()); });
        if (in_flight.empty()) {
            bool any_busy = false;
            for (auto &s : slots)
                any_busy = any_busy || s->busy;
            if (stopping && pending.empty() && !any_busy)
                return;
            continue;
        }
        Slot *slot = in_flight.front();
        in_flight.pop_front();
        lock.unlock();
        finish_batch(*slot);
        lock.lock();
        const auto now = clock_type::now();
        for (Request *r : slot->batch) {
            r->finished = now;
            r->done = true;
            fits_done += r->n;
            ++requests_done;
        }
        bytes_in_flight -= slot->bytes;
        slot->busy = false;
        slot->batch.clear();
        cv_done.notify_all();
        cv_sched.notify_one();
    }
}

}  // namespace bfhip

using namespace bfhip;

namespace {
void destroy_slots(Server &s) {
    for (auto &slot : s.slots) {
        if (slot->stream)
            (void)hipStreamSynchronize(slot->stream);
        for (hipStream_t st : slot->side)
            if (st) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(st);
            }
        if (slot->stream)
            (void)hipStreamDestroy(slot->stream);
        for (hipEvent_t e : {slot->done, slot->fork, slot->join[0], slot->join[1], slot->join[2],
                             slot->join[3]})
            if (e)
                (void)hipEventDestroy(e);
    }
    s.slots.clear();
}
}  // namespace

struct bf_server_s {
    Server server;
};

extern "C" {

int bf_serve_create(bf_server *out, const bf_serve_options *options, char *err, size_t errlen) {
    if (!out)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "server pointer is required");
    *out = nullptr;
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP fitting is unavailable: " + rt.why);
    bf_serve_options o{};
    o.hbm_budget = uint64_t(64) << 30;
    o.slots = 3;
    o.max_batch_fits = 1u << 22;
    o.numerics = BF_COMPENSATED | BF_MIXED_FP64;
    if (options) {
        if (options->hbm_budget)
            o.hbm_budget = options->hbm_budget;
        if (options->slots)
            o.slots = options->slots;
        if (options->max_batch_fits)
            o.max_batch_fits = options->max_batch_fits;
        o.numerics = options->numerics;
    }
    if (o.numerics != 0 && o.numerics != (BF_COMPENSATED | BF_MIXED_FP64))
        return fail(err, errlen, BF_INVALID_ARGUMENT, "numerics must be 0 or 5");
    if (o.slots > 16)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "at most 16 slots");
    auto holder = std::make_unique<bf_server_s>();
    Server &s = holder->server;
    s.opt = o;
    s.mixed = o.numerics != 0;
    for (uint32_t i = 0; i < o.slots; ++i) {
        auto slot = std::make_unique<Slot>();
        bool ok = hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&slot->done, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&slot->fork, hipEventDisableTiming) == hipSuccess;
        for (int b = 0; ok && b < Slot::kBranches; ++b)
            ok = hipStreamCreateWithFlags(&slot->side[b], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&slot->join[b], hipEventDisableTiming) == hipSuccess;
        s.slots.push_back(std::move(slot));
        if (!ok) {
            destroy_slots(s);
            return fail(err, errlen, BF_RUNTIME_ERROR, "could not create slot streams/events");
        }
    }
    s.scheduler = std::thread([&s] { s.schedule_loop(); });
    s.completer = std::thread([&s] { s.complete_loop(); });
    *out = holder.release();
    return BF_OK;
}

int bf_serve_submit(bf_server server, uint32_t model, const bf_config *cfg, int on_device,
                    const float *x, const float *y, const float *p0, float *params,
                    int32_t *status, uint32_t *iterations, float *chi2, float *cov,
                    float *stderr_, int32_t *ustatus, uint64_t *ticket, char *err,
                    size_t errlen) {
    if (!server || !ticket)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "server and ticket are required");
    if (model > 3)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "unknown fit model");
    if (!cfg || cfg->points == 0)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "expected x[M] and y[N,M] with M>0");
    const bool unc = cov || stderr_ || ustatus;
    if (unc && !(cov && stderr_ && ustatus))
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "covariance, stderr and uncertainty status are requested together");
    if (on_device && unc)
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "uncertainties are served for host-array requests only");
    if (on_device && ((uintptr_t(x) | uintptr_t(y) | uintptr_t(p0)) & 3u))
        return fail(err, errlen, BF_INVALID_ARGUMENT, "device arrays must be 4-byte aligned");
    if (cfg->fits && (!x || !y || !params || !status || !iterations || !chi2))
        return fail(err, errlen, BF_INVALID_ARGUMENT, "x, y and every output must be non-null");
    Server &s = server->server;
    auto r = std::make_unique<Request>();
    r->model = model;
    r->n = cfg->fits;
    r->m = cfg->points;
    r->cfg = *cfg;
    r->cfg.numerics = s.opt.numerics;
    r->device = on_device != 0;
    r->x = x;
    r->y = y;
    r->p0 = p0;
    r->params = params;
    r->status = status;
    r->iterations = iterations;
    r->chi2 = chi2;
    r->cov = cov;
    r->se = stderr_;
    r->us = ustatus;
    r->bytes = request_bytes(*r);
    r->submitted = clock_type::now();
    if (r->bytes > s.opt.hbm_budget)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "request exceeds the HBM admission budget");
    std::lock_guard<std::mutex> lock(s.mu);
    if (s.stopping)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "server is shutting down");
    r->ticket = s.next_ticket++;
    *ticket = r->ticket;
    Request *raw = r.get();
    s.table.emplace(raw->ticket, std::move(r));
    if (raw->n == 0) {
        raw->admitted = raw->finished = raw->submitted;
        raw->done = true;
        s.cv_done.notify_all();
        return BF_OK;
    }
    s.pending.push_back(raw);
    s.cv_sched.notify_one();
    return BF_OK;
}

int bf_serve_wait(bf_server server, uint64_t ticket, double timeout_s, bf_request_info *info) {
    if (!server)
        return BF_INVALID_ARGUMENT;
    Server &s = server->server;
    std::unique_lock<std::mutex> lock(s.mu);
    auto it = s.table.find(ticket);
    if (it == s.table.end())
        return BF_INVALID_ARGUMENT;
    Request *r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_s < 0)
        s.cv_done.wait(lock, ready);
    else if (!s.cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), ready))
        return BF_TIMEOUT;
    if (info) {
        using sec = std::chrono::duration<double>;
        info->rc = r->rc;
        info->batch = r->batch;
        info->queued_s = sec(r->admitted - r->submitted).count();
        info->service_s = sec(r->finished - r->admitted).count();
        info->latency_s = sec(r->finished - r->submitted).count();
        std::snprintf(info->message, sizeof info->message, "%s", r->message.c_str());
    }
    const int rc = r->rc;
    s.table.erase(it);
    return rc;
}

int bf_serve_stats(bf_server server, bf_serve_counters *out) {
    if (!server || !out)
        return BF_INVALID_ARGUMENT;
    Server &s = server->server;
    std::lock_guard<std::mutex> lock(s.mu);
    out->requests_done = s.requests_done;
    out->fits_done = s.fits_done;
    out->batches = s.batches;
    out->pending = s.pending.size();
    out->bytes_in_flight = s.bytes_in_flight;
    out->max_bytes_in_flight = s.max_bytes;
    out->max_batches_in_flight = s.max_batches_in_flight;
    return BF_OK;
}

int bf_serve_destroy(bf_server server) {
    if (!server)
        return BF_OK;
    Server &s = server->server;
    {
        std::lock_guard<std::mutex> lock(s.mu);
        s.stopping = true;
    }
    s.cv_sched.notify_all();
    s.cv_complete.notify_all();
    if (s.scheduler.joinable())
        s.scheduler.join();
    {
        std::lock_guard<std::mutex> lock(s.mu);
    }
    s.cv_complete.notify_all();
    if (s.completer.joinable())
        s.completer.join();
    destroy_slots(s);
    delete server;
    return BF_OK;
}

}  // extern "C"
