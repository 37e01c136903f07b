// serve_stress.cpp — the continuous-batching server (bf_serve_*) driven from
// several host threads at once, for an AddressSanitizer + UndefinedBehavior-
// Sanitizer build of its host code (tests/cpp/Makefile, target asan:
// bf_serve.hip compiled with the sanitizers on its host side and linked with
// this driver against libbatchfit_hip.so). Four threads submit host-array requests (with and
// without p0, some with uncertainties) and device-array requests of the four
// models at mixed M and N while a fifth polls the counters; then every
// request is waited for and its bytes are compared with bf_hip_fit on the
// same inputs. Exit status 0 when every result matches.
#include "../../include/batchfit_hip.h"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

const int kParams[4] = {3, 3, 4, 8};

// Small deterministic generator (splitmix64) for shapes, parameters and noise.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform(double a, double b) { return a + (b - a) * double(next() >> 11) * 0x1.0p-53; }
};

struct Req {
    uint32_t model = 0, n = 0, m = 0;
    bool device = false, with_p0 = false, unc = false;
    std::vector<float> x, y, p0;
    bf_config cfg{};
    // outputs (host requests) or host copies of device outputs
    std::vector<float> params, chi2, cov, se;
    std::vector<int32_t> status, us;
    std::vector<uint32_t> iters;
    void *dx = nullptr, *dy = nullptr, *dp0 = nullptr, *dparams = nullptr, *dstatus = nullptr,
         *diters = nullptr, *dchi2 = nullptr;
    uint64_t ticket = 0;
};

float model_value(uint32_t model, const float *p, float x) {
    auto lor = [](float a, float c, float w, float t) {
        const float z = (t - c) / w;
        return a / (1.0f + z * z);
    };
    switch (model) {
    case 0: {
        const float z = (x - p[1]) / p[2];
        return p[0] * std::exp(-0.5f * z * z);
    }
    case 1: return lor(p[0], p[1], p[2], x);
    case 2: {
        const float z = (x - p[1]) / p[2];
        return p[0] * ((1 - p[3]) * std::exp(-4 * std::log(2.0f) * z * z) + p[3] / (1 + 4 * z * z));
    }
    default: return lor(p[0], p[1], p[2], x) + lor(p[3], p[4], p[5], x) + p[6] + p[7] * x;
    }
}

Req make_request(Rng &rng, int index) {
    Req r;
    r.model = uint32_t(index % 4);
    r.m = uint32_t(std::exp(rng.uniform(std::log(64.0), std::log(3000.0))));
    r.n = 1 + uint32_t(std::exp(rng.uniform(0.0, std::log(300.0))));
    r.device = index % 3 == 1;
    r.with_p0 = index % 5 == 2;
    r.unc = !r.device && index % 4 == 3;
    const int P = kParams[r.model];
    r.x.resize(r.m);
    for (uint32_t j = 0; j < r.m; ++j)
        r.x[j] = -8.0f + 16.0f * float(j) / float(r.m - 1);
    r.y.resize(size_t(r.n) * r.m);
    r.p0.resize(size_t(r.n) * P);
    for (uint32_t i = 0; i < r.n; ++i) {
        float p[8] = {float(rng.uniform(1, 4)), float(rng.uniform(-2, 2)), float(rng.uniform(0.4, 1.2)),
                      float(rng.uniform(0.2, 0.8)), 0, 0, 0, 0};
        if (r.model == 3) {
            p[3] = float(rng.uniform(1, 4));
            p[4] = p[1] + float(rng.uniform(2, 4));
            p[5] = float(rng.uniform(0.4, 1.0));
        }
        for (uint32_t j = 0; j < r.m; ++j) {
            const double u1 = rng.uniform(1e-12, 1), u2 = rng.uniform(0, 1);
            const double g = std::sqrt(-2 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            r.y[size_t(i) * r.m + j] = model_value(r.model, p, r.x[j]) + float(0.02 * g);
        }
        for (int k = 0; k < P; ++k)
            r.p0[size_t(i) * P + k] = p[k] * (k == 1 ? 1.0f : 1.1f);
    }
    bf_hip_default_config(&r.cfg);
    r.cfg.fits = r.n;
    r.cfg.points = r.m;
    for (int k = 0; k < P; ++k) {
        r.cfg.lower[k] = -INFINITY;
        r.cfg.upper[k] = INFINITY;
    }
    r.cfg.lower[2] = 1e-3f;
    if (r.model == 3)
        r.cfg.lower[5] = 1e-3f;
    r.params.resize(size_t(r.n) * P);
    r.status.resize(r.n);
    r.iters.resize(r.n);
    r.chi2.resize(r.n);
    if (r.unc) {
        r.cov.resize(size_t(r.n) * P * P);
        r.se.resize(size_t(r.n) * P);
        r.us.resize(r.n);
    }
    return r;
}

bool same(const void *a, const void *b, size_t bytes) { return std::memcmp(a, b, bytes) == 0; }

}  // namespace

int main(int argc, char **argv) {
    const int requests = argc > 1 ? std::atoi(argv[1]) : 96;
    char err[512] = {};
    if (!bf_hip_available()) {
        bf_server probe = nullptr;
        bf_serve_create(&probe, nullptr, err, sizeof err);
        std::printf("no device: %s\n", err);
        return 2;
    }
    Rng rng{20261204};
    std::vector<Req> reqs;
    for (int i = 0; i < requests; ++i)
        reqs.push_back(make_request(rng, i));
    for (Req &r : reqs) {
        if (!r.device)
            continue;
        const int P = kParams[r.model];
        auto dev = [&](void **p, const void *src, size_t bytes) {
            if (bf_hip_malloc(p, bytes, err, sizeof err) != BF_OK)
                return false;
            return !src || bf_hip_memcpy(*p, src, bytes, 1, err, sizeof err) == BF_OK;
        };
        if (!dev(&r.dx, r.x.data(), r.x.size() * 4) || !dev(&r.dy, r.y.data(), r.y.size() * 4) ||
            !dev(&r.dp0, r.p0.data(), r.p0.size() * 4) ||
            !dev(&r.dparams, nullptr, size_t(r.n) * P * 4) || !dev(&r.dstatus, nullptr, r.n * 4u) ||
            !dev(&r.diters, nullptr, r.n * 4u) || !dev(&r.dchi2, nullptr, r.n * 4u)) {
            std::printf("device setup failed: %s\n", err);
            return 1;
        }
    }
    bf_server srv = nullptr;
    bf_serve_options opt{};
    opt.slots = 3;
    opt.max_batch_fits = 512;  // many batches, so admission and completion interleave
    if (bf_serve_create(&srv, &opt, err, sizeof err) != BF_OK) {
        std::printf("create failed: %s\n", err);
        return 1;
    }
    std::atomic<int> submit_errors{0};
    std::atomic<bool> done{false};
    std::thread poller([&] {
        bf_serve_counters c;
        while (!done.load())
            bf_serve_stats(srv, &c);
    });
    std::vector<std::thread> submitters;
    for (int t = 0; t < 4; ++t)
        submitters.emplace_back([&, t] {
            char e[256];
            for (int i = t; i < requests; i += 4) {
                Req &r = reqs[size_t(i)];
                const int rc =
                    r.device
                        ? bf_serve_submit(srv, r.model, &r.cfg, 1, (const float *)r.dx,
                                          (const float *)r.dy,
                                          r.with_p0 ? (const float *)r.dp0 : nullptr,
                                          (float *)r.dparams, (int32_t *)r.dstatus,
                                          (uint32_t *)r.diters, (float *)r.dchi2, nullptr,
                                          nullptr, nullptr, &r.ticket, e, sizeof e)
                        : bf_serve_submit(srv, r.model, &r.cfg, 0, r.x.data(), r.y.data(),
                                          r.with_p0 ? r.p0.data() : nullptr, r.params.data(),
                                          r.status.data(), r.iters.data(), r.chi2.data(),
                                          r.unc ? r.cov.data() : nullptr,
                                          r.unc ? r.se.data() : nullptr,
                                          r.unc ? r.us.data() : nullptr, &r.ticket, e, sizeof e);
                if (rc != BF_OK)
                    submit_errors.fetch_add(1);
            }
        });
    for (auto &t : submitters)
        t.join();
    int wait_errors = 0;
    for (Req &r : reqs) {
        bf_request_info info;
        if (bf_serve_wait(srv, r.ticket, 120.0, &info) != BF_OK)
            ++wait_errors;
    }
    done.store(true);
    poller.join();
    bf_serve_counters c{};
    bf_serve_stats(srv, &c);
    bf_serve_destroy(srv);
    int mismatches = 0;
    for (Req &r : reqs) {
        const int P = kParams[r.model];
        if (r.device) {
            bf_hip_memcpy(r.params.data(), r.dparams, r.params.size() * 4, 2, err, sizeof err);
            bf_hip_memcpy(r.status.data(), r.dstatus, r.n * 4u, 2, err, sizeof err);
            bf_hip_memcpy(r.iters.data(), r.diters, r.n * 4u, 2, err, sizeof err);
            bf_hip_memcpy(r.chi2.data(), r.dchi2, r.n * 4u, 2, err, sizeof err);
        }
        std::vector<float> params(r.params.size()), chi2(r.n), cov(r.cov.size()), se(r.se.size());
        std::vector<int32_t> status(r.n), us(r.us.size());
        std::vector<uint32_t> iters(r.n);
        bf_launch launch{};
        if (bf_hip_fit(r.model, &r.cfg, &launch, r.x.data(), r.y.data(),
                       r.with_p0 ? r.p0.data() : nullptr, params.data(), status.data(),
                       iters.data(), chi2.data(), r.unc ? cov.data() : nullptr,
                       r.unc ? se.data() : nullptr, r.unc ? us.data() : nullptr, err,
                       sizeof err) != BF_OK) {
            std::printf("reference fit failed: %s\n", err);
            return 1;
        }
        bool ok = same(params.data(), r.params.data(), params.size() * 4) &&
                  same(status.data(), r.status.data(), r.n * 4u) &&
                  same(iters.data(), r.iters.data(), r.n * 4u) &&
                  same(chi2.data(), r.chi2.data(), r.n * 4u);
        if (r.unc)
            ok = ok && same(cov.data(), r.cov.data(), cov.size() * 4) &&
                 same(se.data(), r.se.data(), se.size() * 4) && same(us.data(), r.us.data(), r.n * 4u);
        if (!ok && mismatches++ < 5)
            std::printf("mismatch: model %u M=%u N=%u device %d p0 %d\n", r.model, r.m, r.n,
                        int(r.device), int(r.with_p0));
        (void)P;
        for (void *p : {r.dx, r.dy, r.dp0, r.dparams, r.dstatus, r.diters, r.dchi2})
            if (p)
                bf_hip_free(p);
    }
    std::printf("served %d requests in %llu batches (max %u in flight): submit errors %d, wait "
                "errors %d, mismatches %d\n",
                requests, (unsigned long long)c.batches, c.max_batches_in_flight,
                submit_errors.load(), wait_errors, mismatches);
    // Leave without the runtime's exit handlers: the HIP runtime frees its
    // allocations from them after the sanitizer's device-allocator hooks are
    // gone, which the sanitizer reports as a failed internal check.
    std::fflush(stdout);
    std::_Exit(submit_errors.load() || wait_errors || mismatches ? 1 : 0);
}
