// bf_npy_header.hpp — host-only NPY v1.0 header reader and chunk arithmetic
// of the NPY pipeline (bf_npy.hip), kept free of HIP so that it can be built
// and fuzzed under the address and undefined-behaviour sanitizers on its own.
#pragma once

#include "../../include/batchfit_hip.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fcntl.h>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <unistd.h>

namespace bfhip {
namespace {

struct NpyError : std::runtime_error {
    int code;
    NpyError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void invalid(const std::string &m) { throw NpyError(BF_INVALID_ARGUMENT, m); }
[[noreturn]] void runtime(const std::string &m) { throw NpyError(BF_RUNTIME_ERROR, m); }

// NPY v1.0 header (src/cpu/npy.cpp:51-84 accepts the same set): magic,
// version 1.0, header length ≤ 4096 with 16-byte alignment, a dict with the
// keys descr ('<f4', '<f8' or '<u4'), fortran_order False and a shape of rank
// 1 or 2 in this order, and a file size equal to header + payload.
struct NpyHeader {
    uint64_t shape[2] = {0, 0};
    uint32_t rank = 0;
    uint32_t dtype = 0;  // 0 float32, 1 float64, 2 uint32
    uint64_t offset = 0, elements = 0, item = 0;
};

struct Cursor {
    const std::string &s;
    size_t i = 0;
    void space() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t'))
            ++i;
    }
    bool take(char c) {
        space();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    bool quoted(const char *word) {
        space();
        if (i >= s.size() || (s[i] != '\'' && s[i] != '"'))
            return false;
        const char q = s[i];
        const size_t n = std::strlen(word);
        if (s.compare(i + 1, n, word) != 0 || i + 1 + n >= s.size() || s[i + 1 + n] != q)
            return false;
        i += n + 2;
        return true;
    }
    bool number(uint64_t &v) {
        space();
        const size_t start = i;
        v = 0;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
            const uint64_t d = uint64_t(s[i] - '0');
            if (v > (UINT64_MAX - d) / 10)
                return false;
            v = v * 10 + d;
            ++i;
        }
        return i > start;
    }
};

NpyHeader read_header(const std::string &path) {
    const int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0)
        invalid("NPY " + path + ": open for reading failed");
    struct Closer {
        int fd;
        ~Closer() { ::close(fd); }
    } closer{fd};
    unsigned char prefix[10];
    if (::pread(fd, prefix, 10, 0) != 10 || std::memcmp(prefix, "\x93NUMPY\x01\x00", 8) != 0)
        invalid("NPY " + path + ": expected NPY v1.0 magic/version");
    const size_t length = size_t(prefix[8]) + 256 * size_t(prefix[9]);
    if (length == 0 || length > 4096 || (10 + length) % 16 != 0)
        invalid("NPY " + path + ": invalid or oversized header length/alignment");
    std::string text(length, '\0');
    if (::pread(fd, text.data(), length, 10) != ssize_t(length))
        invalid("NPY " + path + ": truncated header");
    NpyHeader h;
    Cursor c{text};
    auto unsupported = [&]() {
        invalid("NPY " + path +
                ": unsupported header (requires <f4/<f8/<u4, C order, rank 1/2, "
                "alphabetical keys)");
    };
    if (!c.take('{') || !c.quoted("descr") || !c.take(':'))
        unsupported();
    if (c.quoted("<f4"))
        h.dtype = 0;
    else if (c.quoted("<f8"))
        h.dtype = 1;
    else if (c.quoted("<u4"))
        h.dtype = 2;
    else
        unsupported();
    if (!c.take(',') || !c.quoted("fortran_order") || !c.take(':'))
        unsupported();
    c.space();
    if (text.compare(c.i, 5, "False") != 0)
        unsupported();
    c.i += 5;
    if (!c.take(',') || !c.quoted("shape") || !c.take(':') || !c.take('('))
        unsupported();
    uint64_t d0 = 0, d1 = 0;
    if (!c.number(d0))
        unsupported();
    h.rank = 1;
    h.shape[0] = d0;
    if (c.take(',')) {
        if (c.number(d1)) {
            h.rank = 2;
            h.shape[1] = d1;
            c.take(',');
        }
    }
    if (!c.take(')'))
        unsupported();
    c.take(',');
    if (!c.take('}'))
        unsupported();
    c.space();
    if (c.i + 1 != text.size() || text[c.i] != '\n')
        unsupported();
    h.item = h.dtype == 1 ? 8 : 4;
    h.elements = h.shape[0];
    if (h.rank == 2) {
        if (h.shape[1] && h.elements > UINT64_MAX / h.shape[1])
            invalid("NPY " + path + ": shape overflows size_t");
        h.elements *= h.shape[1];
    }
    if (h.elements > (UINT64_MAX - 4096) / h.item)
        invalid("NPY " + path + ": array byte count overflows");
    h.offset = 10 + length;
    struct stat st;
    if (::fstat(fd, &st) != 0 || uint64_t(st.st_size) != h.offset + h.elements * h.item)
        invalid("NPY " + path + ": payload length does not match shape");
    return h;
}

// Fits per chunk (choose_chunk_fits, src/metal/pipeline.cpp:28-68): chunk_fits
// when set, else the byte budget over the ring's bytes per fit, rounded down
// to a multiple of 256; 0 for invalid dimensions. Bytes per fit that do not
// fit in 64 bits make the chunk one fit.
uint64_t npy_chunk_fits(uint64_t fits, uint64_t points, uint32_t parameters, uint32_t input_bytes,
                        uint32_t p0_bytes, const bf_npy_options *opt) {
    if (!opt || points == 0 || parameters == 0 || parameters > 8 || opt->in_flight == 0 ||
        opt->in_flight > 8 || (input_bytes != 4 && input_bytes != 8) ||
        (p0_bytes != 0 && p0_bytes != 4 && p0_bytes != 8) || fits == 0)
        return 0;
    if (opt->chunk_fits)
        return std::min<uint64_t>(opt->chunk_fits, fits);
    // Per fit and slot: the window as read and as float32, p0, outputs
    // (src/metal/pipeline.cpp:55-64 sizes its ring the same way).
    const uint64_t p = parameters;
    const uint64_t extra = opt->uncertainties ? (p * p + p + 1) * 4 : 0;
    uint64_t per_fit = 0;
    if (__builtin_mul_overflow(points, uint64_t(4 + input_bytes), &per_fit) ||
        __builtin_add_overflow(per_fit, p * (p0_bytes + 4) + 32 + 12 + extra, &per_fit) ||
        __builtin_mul_overflow(per_fit, uint64_t(opt->in_flight), &per_fit))
        per_fit = UINT64_MAX;
    const uint64_t budget = opt->budget_bytes ? opt->budget_bytes : (uint64_t(1) << 30);
    uint64_t chunk = std::max<uint64_t>(1, budget / per_fit);
    chunk = std::min<uint64_t>({chunk, fits, uint64_t(UINT32_MAX)});
    if (chunk >= 256)
        chunk = chunk / 256 * 256;
    return chunk;
}

}  // namespace
}  // namespace bfhip
