// npy_fuzz.cpp — the NPY header reader and chunk arithmetic of the streaming
// pipeline (batchfit_amd/csrc/bf_npy_header.hpp) under AddressSanitizer and
// UndefinedBehaviorSanitizer. Headers that NumPy writes, then many mutated
// ones (bytes replaced, inserted and deleted; digit runs grown past 64 bits;
// header lengths, alignment and payload sizes off by a little or a lot), are
// written to a file and read: the reader must accept (and then describe the
// file exactly) or reject with BF_INVALID_ARGUMENT, and never read or write
// outside its buffers. The chunk arithmetic is held to its contract over
// dimensions at the limits of their integer types.
//
//   g++ -std=c++17 -O1 -g -fsanitize=address,undefined -fno-sanitize-recover=all \
//       tests/cpp/npy_fuzz.cpp -o npy_fuzz && ./npy_fuzz <dir> <iterations>
#include "../../batchfit_amd/csrc/bf_npy_header.hpp"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

namespace {

int failures = 0;
void expect(bool ok, const std::string &what) {
    if (!ok && failures++ < 20)
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
}

// A v1.0 file: magic, little-endian length, header text padded with spaces
// and a newline to a 16-byte boundary (unless `pad` is false), payload.
std::string npy_file(const std::string &text, bool pad, uint64_t payload, const char *magic) {
    std::string body = text;
    if (pad) {
        const size_t n = (16 - (10 + body.size() + 1) % 16) % 16;
        body += std::string(n, ' ') + "\n";
    }
    std::string out(magic, 8);
    out += char(body.size() & 0xff);
    out += char((body.size() >> 8) & 0xff);
    out += body;
    out += std::string(size_t(payload), '\0');
    return out;
}

bool write_file(const std::string &path, const std::string &bytes) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        return false;
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: npy_fuzz <dir> <iterations>\n");
        return 2;
    }
    const std::string path = std::string(argv[1]) + "/fuzz.npy";
    const long iterations = std::atol(argv[2]);
    const char *magic = "\x93NUMPY\x01\x00";
    struct Seed {
        std::string text;
        uint64_t payload;
    };
    const std::vector<Seed> seeds = {
        {"{'descr': '<f4', 'fortran_order': False, 'shape': (3, 5), }", 60},
        {"{'descr': '<f8', 'fortran_order': False, 'shape': (7,), }", 56},
        {"{'descr': '<u4', 'fortran_order': False, 'shape': (0, 9), }", 0},
        {"{'descr': '<f4', 'fortran_order': False, 'shape': (2, 0), }", 0},
        {"{\"descr\": \"<f8\", \"fortran_order\": False, \"shape\": (1, 1)}", 8},
        {"{'descr': '<f4', 'fortran_order': False, 'shape': (18446744073709551615,), }", 0},
        {"{'descr': '<f8', 'fortran_order': False, 'shape': (4294967296, 4294967296), }", 0},
        {"{'descr': '<f4', 'fortran_order': False, 'shape': (4611686018427387904,), }", 0},
    };
    std::mt19937_64 rng(20261204);
    const std::string alphabet = "'\"(),:{} \t\n0123456789fFalseTru<>duescrhapio_\x00\xff";
    long accepted = 0, rejected = 0;
    for (long it = 0; it < iterations; ++it) {
        const Seed &seed = seeds[rng() % seeds.size()];
        std::string text = seed.text;
        const int edits = int(rng() % 4);
        for (int e = 0; e < edits && !text.empty(); ++e) {
            const size_t at = rng() % text.size();
            switch (rng() % 4) {
            case 0: text[at] = alphabet[rng() % alphabet.size()]; break;
            case 1: text.insert(at, 1, alphabet[rng() % alphabet.size()]); break;
            case 2: text.erase(at, 1 + rng() % 3); break;
            default: {  // grow a digit run
                const size_t d = text.find_first_of("0123456789", at);
                if (d != std::string::npos)
                    text.insert(d, std::string(1 + rng() % 25, char('0' + rng() % 10)));
            }
            }
        }
        uint64_t payload = seed.payload;
        switch (rng() % 5) {
        case 0: payload += 1 + rng() % 16; break;
        case 1: payload = payload ? payload - 1 : 3; break;
        default: break;
        }
        std::string file = npy_file(text, rng() % 8 != 0, payload,
                                    rng() % 16 ? magic : "\x93NUMPY\x02\x00");
        if (rng() % 16 == 0 && file.size() > 10)  // truncated
            file.resize(rng() % file.size());
        if (rng() % 16 == 0 && file.size() >= 10) {  // header length off
            file[8] = char(rng());
            file[9] = char(rng() % 32);
        }
        if (!write_file(path, file)) {
            std::fprintf(stderr, "cannot write %s\n", path.c_str());
            return 2;
        }
        try {
            const bfhip::NpyHeader h = bfhip::read_header(path);
            ++accepted;
            expect(h.rank == 1 || h.rank == 2, "rank");
            expect(h.dtype <= 2, "dtype");
            expect(h.item == (h.dtype == 1 ? 8u : 4u), "item size");
            expect(h.offset % 16 == 0 && h.offset >= 16, "aligned offset");
            expect(h.offset + h.elements * h.item == file.size(), "payload length");
            const uint64_t elements = h.rank == 2 ? h.shape[0] * h.shape[1] : h.shape[0];
            expect(elements == h.elements, "elements");
        } catch (const bfhip::NpyError &e) {
            ++rejected;
            expect(e.code == BF_INVALID_ARGUMENT, std::string("error code: ") + e.what());
        }
    }
    // Chunk arithmetic: 0 exactly for invalid dimensions; otherwise at least
    // one fit, at most the batch and 2^32 − 1, a multiple of 256 from 256 on
    // (unless set explicitly).
    const uint64_t wide[] = {0, 1, 255, 256, 257, 1024, 65535, uint64_t(1) << 32,
                             uint64_t(1) << 62, UINT64_MAX};
    long chunks = 0;
    for (long it = 0; it < iterations; ++it) {
        bf_npy_options o{};
        o.in_flight = uint32_t(rng() % 10);
        o.chunk_fits = rng() % 4 == 0 ? wide[rng() % 10] : 0;
        o.budget_bytes = rng() % 2 ? wide[rng() % 10] : 0;
        o.uncertainties = uint32_t(rng() % 2);
        const uint64_t fits = wide[rng() % 10], points = wide[rng() % 10];
        const uint32_t params = uint32_t(rng() % 10), in_b = uint32_t(rng() % 10),
                       p0_b = uint32_t(rng() % 10);
        const uint64_t c = bfhip::npy_chunk_fits(fits, points, params, in_b, p0_b, &o);
        const bool valid = points && params && params <= 8 && o.in_flight && o.in_flight <= 8 &&
                           (in_b == 4 || in_b == 8) && (p0_b == 0 || p0_b == 4 || p0_b == 8) &&
                           fits;
        expect(valid == (c != 0), "chunk validity");
        if (c) {
            expect(c <= fits, "chunk within the batch");
            if (!o.chunk_fits) {
                expect(c <= UINT32_MAX, "chunk below 2^32");
                expect(c < 256 || c % 256 == 0, "chunk a multiple of 256");
            }
        }
        ++chunks;
    }
    std::printf("headers accepted %ld, rejected %ld; chunk cases %ld; failures %d\n", accepted,
                rejected, chunks, failures);
    return failures ? 1 : 0;
}
