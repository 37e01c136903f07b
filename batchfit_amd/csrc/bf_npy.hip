// bf_npy.hip — NPY inputs and the streaming fit pipeline (bf_npy_inspect,
// bf_hip_fit_npy).
//
// The reference streams NPY files through its GPU in fixed chunks
// (metal::fit_npy / stream_npy, src/metal/pipeline.cpp:28-265; NPY reading in
// src/cpu/npy.cpp:51-84 and src/cpu/mmap.cpp:91-145): memory-map a window of
// y, convert float64 to float32 and guess on the host, transpose to
// point-major on the host, submit, and hand each finished chunk to an ordered
// consumer, with at most three chunks in flight.
//
// Here the host only moves bytes: each chunk's window of the file is read
// by several threads (pread, no page mapping) straight into a pinned slot,
// sent to HBM as it is
// (float64 stays float64 over PCIe and is converted by a kernel), and fitted
// by the uniform device path on the slot's own stream — the data-only guess
// runs on the device and y stays fit-major, so there is no host transpose.
// Outputs come back to pinned memory; a reader thread waits for chunks in
// order and calls the consumer. The next window is announced to the kernel
// (POSIX_FADV_WILLNEED) while the current one is fitted.
#include "bf_host.hpp"
#include "bf_npy_header.hpp"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <fcntl.h>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

namespace bfhip {
namespace {

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// A read-only file read one window at a time: y windows by parallel pread
// into pinned memory, small windows (grid, p0) mapped as in
// src/cpu/mmap.cpp:91-145 (page-aligned, sequential advice); read-ahead of
// the next y window.
struct MappedFile {
    std::string path;
    NpyHeader h;
    int fd = -1;
    explicit MappedFile(const std::string &p) : path(p), h(read_header(p)) {
        if (h.dtype == 2)
            invalid("mapped NPY inputs require float32/float64");
        fd = ::open(p.c_str(), O_RDONLY | O_CLOEXEC);
        if (fd < 0)
            runtime("NPY mmap open failed: " + p);
    }
    ~MappedFile() {
        if (fd >= 0)
            ::close(fd);
    }
    struct Window {
        void *base = nullptr;
        size_t bytes = 0, delta = 0;
        ~Window() {
            if (base)
                ::munmap(base, bytes);
        }
        const char *data() const { return static_cast<const char *>(base) + delta; }
    };
    // Elements [first, first+count) mapped; returns the mapping.
    std::unique_ptr<Window> window(uint64_t first, uint64_t count) const {
        auto w = std::make_unique<Window>();
        if (count == 0)
            return w;
        const long page = ::sysconf(_SC_PAGESIZE);
        const uint64_t off = h.offset + first * h.item;
        const uint64_t aligned = off / uint64_t(page) * uint64_t(page);
        w->delta = size_t(off - aligned);
        w->bytes = size_t(count * h.item) + w->delta;
        void *b = ::mmap(nullptr, w->bytes, PROT_READ, MAP_PRIVATE, fd, off_t(aligned));
        if (b == MAP_FAILED)
            runtime("NPY mmap failed: " + path);
        w->base = b;
        (void)::madvise(b, w->bytes, MADV_SEQUENTIAL);
        return w;
    }
    // Elements [first, first+count) read into dst by up to `threads` threads,
    // each with pread over its own piece (page cache to pinned memory without
    // the page faults of a mapping).
    void read(uint64_t first, uint64_t count, char *dst, unsigned threads) const {
        const uint64_t off = h.offset + first * h.item;
        const size_t bytes = size_t(count * h.item);
        const size_t min_piece = size_t(8) << 20;
        const unsigned t =
            unsigned(std::min<size_t>(threads, std::max<size_t>(1, bytes / min_piece)));
        std::atomic<int> failed{0};
        auto piece = [&](size_t a, size_t b) {
            while (a < b) {
                const ssize_t got = ::pread(fd, dst + a, b - a, off_t(off + a));
                if (got < 0 && errno == EINTR)
                    continue;
                if (got <= 0) {
                    failed.store(1);
                    return;
                }
                a += size_t(got);
            }
        };
        if (t <= 1) {
            piece(0, bytes);
        } else {
            std::vector<std::thread> pool;
            const size_t step = (bytes + t - 1) / t;
            for (unsigned k = 0; k < t; ++k) {
                const size_t a = k * step, b = std::min(bytes, a + step);
                if (a < b)
                    pool.emplace_back(piece, a, b);
            }
            for (auto &th : pool)
                th.join();
        }
        if (failed.load())
            runtime("NPY read failed: " + path);
    }
    void prefetch(uint64_t first, uint64_t count) const {
        if (count)
            (void)::posix_fadvise(fd, off_t(h.offset + first * h.item), off_t(count * h.item),
                                  POSIX_FADV_WILLNEED);
    }
};

// float64 → float32, round to nearest (the conversion of
// src/cpu/mmap.cpp:100-117 and of NumPy's astype).
__global__ void __launch_bounds__(256) narrow_kernel(const double *in, float *out, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * 256)
        out[i] = __double2float_rn(in[i]);
}

struct PipeSlot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char *pinned = nullptr;  // staged input window(s) | outputs
    size_t pinned_bytes = 0;
    char *device = nullptr;  // raw input | f32 y | p0 | outputs
    size_t device_bytes = 0;
    uint64_t first = 0, count = 0;
    bool busy = false;
    double submitted = 0;
    size_t out_params = 0, out_status = 0, out_iter = 0, out_chi2 = 0, out_cov = 0, out_se = 0,
           out_us = 0;
    ~PipeSlot() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (pinned)
            (void)hipHostFree(pinned);
        if (device)
            (void)hipFree(device);
        if (done)
            (void)hipEventDestroy(done);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

constexpr size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Ring slots of finished calls, kept for the next call: pinning host memory
// and touching it the first time cost more than streaming the chunks through
// it. Held until the process ends (never freed, so nothing runs at exit);
// a call takes slots out and gives them back, so concurrent calls never share
// one.
struct SlotPool {
    std::mutex mu;
    std::vector<std::unique_ptr<PipeSlot>> idle;
};
SlotPool &slot_pool() {
    static SlotPool *pool = new SlotPool;
    return *pool;
}

void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        runtime(std::string(what) + " failed: " + hipGetErrorString(e));
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

extern "C" {

int bf_npy_inspect(const char *path, uint64_t *shape, uint32_t *rank, uint32_t *dtype,
                   uint64_t *offset, char *err, size_t errlen) {
    if (!path || !shape || !rank || !dtype)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "path and outputs are required");
    try {
        const NpyHeader h = read_header(path);
        shape[0] = h.shape[0];
        shape[1] = h.shape[1];
        *rank = h.rank;
        *dtype = h.dtype;
        if (offset)
            *offset = h.offset;
        return BF_OK;
    } catch (const NpyError &e) {
        return fail(err, errlen, e.code, e.what());
    }
}

uint64_t bf_npy_chunk_fits(uint64_t fits, uint64_t points, uint32_t parameters,
                           uint32_t input_bytes, uint32_t p0_bytes, const bf_npy_options *opt) {
    return npy_chunk_fits(fits, points, parameters, input_bytes, p0_bytes, opt);
}

int bf_hip_fit_npy(uint32_t model, const char *x_path, const char *y_path, const char *p0_path,
                   const bf_config *cfg_in, const bf_npy_options *opt_in,
                   bf_npy_consumer consume, void *user, bf_npy_stats *stats_out, char *err,
                   size_t errlen) {
    const double start = now_seconds();
    if (model > 3)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "unknown fit model");
    if (!x_path || !y_path || !cfg_in || !consume)
        return fail(err, errlen, BF_INVALID_ARGUMENT,
                    "x and y paths, a config and a result consumer are required");
    bf_npy_options opt{};
    opt.in_flight = 3;
    if (opt_in)
        opt = *opt_in;
    if (opt.in_flight == 0 || opt.in_flight > 8)
        return fail(err, errlen, BF_INVALID_ARGUMENT, "in_flight must be 1..8");
    Runtime &rt = Runtime::get();
    if (rt.status != BF_OK)
        return fail(err, errlen, BF_UNAVAILABLE, "HIP fitting is unavailable: " + rt.why);
    bf_npy_stats stats{};
    try {
        MappedFile xf(x_path), yf(y_path);
        if (yf.h.rank != 2 || yf.h.shape[1] == 0 || xf.h.rank != 1 ||
            xf.h.shape[0] != yf.h.shape[1])
            invalid("pipeline requires x[M], y[N,M] NPY arrays");
        const uint64_t n = yf.h.shape[0], m = yf.h.shape[1];
        const uint32_t P = uint32_t(kParams[model]);
        std::unique_ptr<MappedFile> pf;
        if (p0_path && *p0_path) {
            pf = std::make_unique<MappedFile>(p0_path);
            if (pf->h.rank != 2 || pf->h.shape[0] != n || pf->h.shape[1] != P)
                invalid("pipeline initial NPY must be [N,P]");
        }
        if (m > UINT32_MAX)
            invalid("M exceeds 2^32 - 1 points");
        stats.fits = n;
        const uint64_t chunk = bf_npy_chunk_fits(n, m, P, uint32_t(yf.h.item),
                                                 pf ? uint32_t(pf->h.item) : 0, &opt);
        stats.chunk_fits = chunk;
        if (n == 0) {
            stats.wall_seconds = now_seconds() - start;
            if (stats_out)
                *stats_out = stats;
            return BF_OK;
        }
        // The grid, once, in float32 on the device.
        std::vector<float> grid(m);
        {
            auto w = xf.window(0, m);
            for (uint64_t j = 0; j < m; ++j) {
                if (xf.h.dtype == 0)
                    std::memcpy(&grid[j], w->data() + 4 * j, 4);
                else {
                    double v;
                    std::memcpy(&v, w->data() + 8 * j, 8);
                    grid[j] = float(v);
                }
            }
        }
        float *d_x = nullptr;
        check(hipMalloc(reinterpret_cast<void **>(&d_x), m * 4), "hipMalloc(grid)");
        struct XFree {
            float *p;
            ~XFree() { (void)hipFree(p); }
        } x_free{d_x};
        check(hipMemcpy(d_x, grid.data(), m * 4, hipMemcpyHostToDevice), "grid upload");

        const bool unc = opt.uncertainties != 0;
        const uint64_t in_bytes = chunk * m * yf.h.item;
        const uint64_t p0_in = pf ? chunk * P * pf->h.item : 0;
        std::vector<std::unique_ptr<PipeSlot>> slots;
        // However this call ends, its slots go back to the pool once their
        // streams are idle.
        struct GiveBack {
            std::vector<std::unique_ptr<PipeSlot>> &slots;
            ~GiveBack() {
                SlotPool &pool = slot_pool();
                for (auto &s : slots)
                    if (s->stream)
                        (void)hipStreamSynchronize(s->stream);
                std::lock_guard<std::mutex> lock(pool.mu);
                for (auto &s : slots) {
                    s->busy = false;
                    pool.idle.push_back(std::move(s));
                }
            }
        } give_back{slots};
        size_t ring_bytes = 0;
        for (uint32_t k = 0; k < opt.in_flight; ++k) {
            std::unique_ptr<PipeSlot> s;
            {
                SlotPool &pool = slot_pool();
                std::lock_guard<std::mutex> lock(pool.mu);
                if (!pool.idle.empty()) {
                    s = std::move(pool.idle.back());
                    pool.idle.pop_back();
                }
            }
            if (!s)
                s = std::make_unique<PipeSlot>();
            slots.push_back(std::move(s));  // given back even if the setup below fails
            PipeSlot *sp = slots.back().get();
            if (!sp->stream)
                check(hipStreamCreateWithFlags(&sp->stream, hipStreamNonBlocking),
                      "hipStreamCreate");
            if (!sp->done)
                check(hipEventCreateWithFlags(&sp->done, hipEventDisableTiming), "hipEventCreate");
            sp->busy = false;
            sp->first = sp->count = 0;
            size_t off = 0;
            auto carve = [&](size_t b) {
                const size_t at = off;
                off = align256(off + b);
                return at;
            };
            carve(in_bytes);  // staged y window (as stored)
            carve(p0_in);     // staged p0 window (as stored)
            sp->out_params = carve(chunk * P * 4);
            sp->out_status = carve(chunk * 4);
            sp->out_iter = carve(chunk * 4);
            sp->out_chi2 = carve(chunk * 4);
            sp->out_cov = unc ? carve(chunk * P * P * 4) : 0;
            sp->out_se = unc ? carve(chunk * P * 4) : 0;
            sp->out_us = unc ? carve(chunk * 4) : 0;
            ring_bytes += off;
            if (sp->pinned_bytes < off) {
                if (sp->pinned)
                    check(hipHostFree(sp->pinned), "hipHostFree(ring)");
                sp->pinned = nullptr;
                sp->pinned_bytes = 0;
                check(hipHostMalloc(reinterpret_cast<void **>(&sp->pinned), off),
                      "hipHostMalloc(ring)");
                sp->pinned_bytes = off;
            }
            off = 0;
            carve(in_bytes);                                   // raw y (float64 windows)
            carve(yf.h.item == 8 ? chunk * m * 4 : 0);         // float32 y after narrowing
            carve(p0_in);                                      // raw p0
            carve(pf && pf->h.item == 8 ? chunk * P * 4 : 0);  // float32 p0
            carve(chunk * P * 4 + chunk * 12);                 // params, status, iter, chi2
            carve(unc ? chunk * (P * P + P + 1) * 4 : 0);
            if (sp->device_bytes < off) {
                if (sp->device)
                    check(hipFree(sp->device), "hipFree(ring)");
                sp->device = nullptr;
                sp->device_bytes = 0;
                check(hipMalloc(reinterpret_cast<void **>(&sp->device), off), "hipMalloc(ring)");
                sp->device_bytes = off;
            }
        }
        stats.shared_payload_bytes = ring_bytes;

        std::mutex mu;
        std::condition_variable cv;
        std::deque<PipeSlot *> queued;
        bool producer_done = false;
        std::string failure;
        int failure_code = BF_OK;
        auto set_failure = [&](int code, const std::string &m) {
            std::lock_guard<std::mutex> lock(mu);
            if (failure_code == BF_OK) {
                failure_code = code;
                failure = m;
            }
            cv.notify_all();
        };
        std::thread reader([&] {
            for (;;) {
                PipeSlot *s;
                {
                    std::unique_lock<std::mutex> lock(mu);
                    cv.wait(lock, [&] { return !queued.empty() || producer_done; });
                    if (queued.empty())
                        return;
                    s = queued.front();
                    queued.pop_front();
                }
                const hipError_t e = hipEventSynchronize(s->done);
                const double t_ready = now_seconds();
                bool cancelled;
                {
                    std::lock_guard<std::mutex> lock(mu);
                    cancelled = failure_code != BF_OK;
                }
                if (e != hipSuccess) {
                    set_failure(BF_RUNTIME_ERROR,
                                std::string("chunk completion failed: ") + hipGetErrorString(e));
                } else if (!cancelled) {
                    const char *H = s->pinned;
                    const int rc = consume(
                        user, s->first, s->count, reinterpret_cast<const float *>(H + s->out_params),
                        reinterpret_cast<const int32_t *>(H + s->out_status),
                        reinterpret_cast<const uint32_t *>(H + s->out_iter),
                        reinterpret_cast<const float *>(H + s->out_chi2),
                        unc ? reinterpret_cast<const float *>(H + s->out_cov) : nullptr,
                        unc ? reinterpret_cast<const float *>(H + s->out_se) : nullptr,
                        unc ? reinterpret_cast<const int32_t *>(H + s->out_us) : nullptr);
                    if (rc != 0)
                        set_failure(BF_RUNTIME_ERROR, "the result consumer stopped the pipeline");
                }
                const double t_done = now_seconds();
                {
                    std::lock_guard<std::mutex> lock(mu);
                    stats.readback_seconds += t_done - t_ready;
                    stats.device_seconds += t_ready - s->submitted;
                    s->busy = false;
                    cv.notify_all();
                }
            }
        });
        const unsigned copy_threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        bf_config cfg = *cfg_in;
        try {
            yf.prefetch(0, std::min(chunk, n) * m);
            uint64_t chunk_index = 0;
            for (uint64_t first = 0; first < n; ++chunk_index) {
                PipeSlot &s = *slots[chunk_index % slots.size()];
                {
                    std::unique_lock<std::mutex> lock(mu);
                    cv.wait(lock, [&] { return !s.busy || failure_code != BF_OK; });
                    if (failure_code != BF_OK)
                        break;
                    s.busy = true;
                }
                const uint64_t count = std::min(chunk, n - first);
                s.first = first;
                s.count = count;
                const double t0 = now_seconds();
                // Window → pinned, as stored.
                const size_t ybytes = size_t(count * m * yf.h.item);
                yf.read(first * m, count * m, s.pinned, copy_threads);
                stats.max_mapped_bytes = std::max<uint64_t>(stats.max_mapped_bytes, ybytes);
                const size_t pbytes = pf ? size_t(count * P * pf->h.item) : 0;
                const size_t p_pinned = align256(size_t(in_bytes));
                if (pf) {
                    auto w = pf->window(first * P, count * P);
                    std::memcpy(s.pinned + p_pinned, w->data(), pbytes);
                }
                if (first + count < n)
                    yf.prefetch((first + count) * m, std::min(chunk, n - first - count) * m);
                stats.prepare_seconds += now_seconds() - t0;
                // Pinned → HBM, narrowing on the device, fit, outputs back.
                char *D = s.device;
                size_t off = 0;
                auto next = [&](size_t b) {
                    const size_t at = off;
                    off = align256(off + b);
                    return at;
                };
                const size_t d_raw = next(in_bytes);
                const size_t d_y32 = next(yf.h.item == 8 ? chunk * m * 4 : 0);
                const size_t d_praw = next(p0_in);
                const size_t d_p32 = next(pf && pf->h.item == 8 ? chunk * P * 4 : 0);
                const size_t d_out = next(chunk * P * 4 + chunk * 12);
                const size_t d_unc = next(unc ? chunk * (P * P + P + 1) * 4 : 0);
                hipStream_t st = s.stream;
                check(hipMemcpyAsync(D + d_raw, s.pinned, ybytes, hipMemcpyHostToDevice, st),
                      "chunk upload");
                const float *dy = reinterpret_cast<const float *>(D + d_raw);
                if (yf.h.item == 8) {
                    const uint64_t cnt = count * m;
                    const unsigned blocks = unsigned(std::min<uint64_t>((cnt + 255) / 256, 8192));
                    hipLaunchKernelGGL(narrow_kernel, dim3(blocks), dim3(256), 0, st,
                                       reinterpret_cast<const double *>(D + d_raw),
                                       reinterpret_cast<float *>(D + d_y32), cnt);
                    check(hipGetLastError(), "narrowing launch");
                    dy = reinterpret_cast<const float *>(D + d_y32);
                }
                const float *dp0 = nullptr;
                if (pf) {
                    check(hipMemcpyAsync(D + d_praw, s.pinned + p_pinned, pbytes,
                                         hipMemcpyHostToDevice, st),
                          "p0 upload");
                    dp0 = reinterpret_cast<const float *>(D + d_praw);
                    if (pf->h.item == 8) {
                        hipLaunchKernelGGL(narrow_kernel, dim3(unsigned((count * P + 255) / 256)),
                                           dim3(256), 0, st,
                                           reinterpret_cast<const double *>(D + d_praw),
                                           reinterpret_cast<float *>(D + d_p32), count * P);
                        check(hipGetLastError(), "narrowing launch");
                        dp0 = reinterpret_cast<const float *>(D + d_p32);
                    }
                }
                float *o_params = reinterpret_cast<float *>(D + d_out);
                int32_t *o_status = reinterpret_cast<int32_t *>(o_params + count * P);
                uint32_t *o_iter = reinterpret_cast<uint32_t *>(o_status + count);
                float *o_chi2 = reinterpret_cast<float *>(o_iter + count);
                float *o_cov = unc ? reinterpret_cast<float *>(D + d_unc) : nullptr;
                float *o_se = unc ? o_cov + count * P * P : nullptr;
                int32_t *o_us = unc ? reinterpret_cast<int32_t *>(o_se + count * P) : nullptr;
                cfg.fits = uint32_t(count);
                cfg.points = uint32_t(m);
                bf_launch launch{};
                launch.stream = st;
                char msg[512];
                const int rc = bf_hip_fit_device(model, &cfg, &launch, d_x, dy, dp0, o_params,
                                                 o_status, o_iter, o_chi2, o_cov, o_se, o_us, msg,
                                                 sizeof msg);
                if (rc != BF_OK)
                    throw NpyError(rc, msg);
                char *H = s.pinned;
                check(hipMemcpyAsync(H + s.out_params, o_params, count * P * 4,
                                     hipMemcpyDeviceToHost, st),
                      "output download");
                check(hipMemcpyAsync(H + s.out_status, o_status, count * 4, hipMemcpyDeviceToHost,
                                     st),
                      "output download");
                check(hipMemcpyAsync(H + s.out_iter, o_iter, count * 4, hipMemcpyDeviceToHost, st),
                      "output download");
                check(hipMemcpyAsync(H + s.out_chi2, o_chi2, count * 4, hipMemcpyDeviceToHost, st),
                      "output download");
                if (unc) {
                    check(hipMemcpyAsync(H + s.out_cov, o_cov, count * P * P * 4,
                                         hipMemcpyDeviceToHost, st),
                          "output download");
                    check(hipMemcpyAsync(H + s.out_se, o_se, count * P * 4, hipMemcpyDeviceToHost,
                                         st),
                          "output download");
                    check(hipMemcpyAsync(H + s.out_us, o_us, count * 4, hipMemcpyDeviceToHost, st),
                          "output download");
                }
                check(hipEventRecord(s.done, st), "hipEventRecord");
                s.submitted = now_seconds();
                {
                    std::lock_guard<std::mutex> lock(mu);
                    queued.push_back(&s);
                    uint64_t active = 0;
                    for (auto &c : slots)
                        active += c->busy;
                    stats.max_in_flight = std::max(stats.max_in_flight, active);
                    cv.notify_all();
                }
                ++stats.chunks;
                first += count;
            }
        } catch (const NpyError &e) {
            set_failure(e.code, e.what());
        } catch (const std::exception &e) {
            set_failure(BF_RUNTIME_ERROR, e.what());
        }
        {
            std::lock_guard<std::mutex> lock(mu);
            producer_done = true;
            cv.notify_all();
        }
        reader.join();  // every queued chunk is waited for before the slots go away
        if (failure_code != BF_OK)
            return fail(err, errlen, failure_code, failure);
    } catch (const NpyError &e) {
        return fail(err, errlen, e.code, e.what());
    } catch (const std::exception &e) {
        return fail(err, errlen, BF_RUNTIME_ERROR, e.what());
    }
    stats.wall_seconds = now_seconds() - start;
    if (stats_out)
        *stats_out = stats;
    return BF_OK;
}

}  // extern "C"
