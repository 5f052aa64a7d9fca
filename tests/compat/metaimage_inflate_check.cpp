// CPU check of metaimage::inflate_all's chunked loop: payloads are fed and
// drained in pieces of `chunk` bytes (the reader uses UINT_MAX, zlib's 32-bit
// limit per call); small chunks exercise the same loop on small data.
// Prints "ok" and exits 0, or names the failing case and exits 1.
#include "metaimage.hpp"

#include <cstdio>
#include <random>

int main() {
    std::mt19937 rng(7);
    for (size_t n : {size_t(1), size_t(1000), size_t(123457)}) {
        std::vector<unsigned char> raw(n);
        for (size_t i = 0; i < n; ++i) raw[i] = (unsigned char)(rng() % 7);   // compressible
        uLongf cap = compressBound(uLong(n));
        std::vector<unsigned char> z(cap);
        if (compress2(z.data(), &cap, raw.data(), uLong(n), Z_DEFAULT_COMPRESSION) != Z_OK) return 1;
        z.resize(cap);
        for (size_t chunk : {size_t(1), size_t(7), size_t(4096), size_t(UINT_MAX)}) {
            if (metaimage::inflate_all(z, n, chunk) != raw) {
                std::printf("mismatch n=%zu chunk=%zu\n", n, chunk);
                return 1;
            }
            bool threw = false;   // truncated payload and wrong declared size must fail
            try {
                std::vector<unsigned char> t(z.begin(), z.end() - 1);
                metaimage::inflate_all(t, n, chunk);
            } catch (const std::exception&) { threw = true; }
            bool threw2 = false;
            try { metaimage::inflate_all(z, n + 1, chunk); } catch (const std::exception&) { threw2 = true; }
            if (!threw || !threw2) {
                std::printf("no error n=%zu chunk=%zu\n", n, chunk);
                return 1;
            }
        }
    }
    std::printf("ok\n");
    return 0;
}
