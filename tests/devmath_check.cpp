// Test helper (built and run by tests/test_devmath.py on the host): the device arithmetic of beamlab_amd/csrc/
// bl_devmath.h, compiled for the host, against the host C library on every float input (one-argument functions) or
// on drawn pairs (atan2f). Prints one line per function: "name checked mismatches [first mismatching input]".
//
//   g++ -std=c++17 -O2 -ffp-contract=off -fopenmp -I beamlab_amd/csrc tests/devmath_check.cpp -o devmath_check -lm
#include "bl_devmath.h"

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace blg;

static bool same(float a, float b) { return f2u(a) == f2u(b) || (std::isnan(a) && std::isnan(b)); }

template <class F, class G>
static void scan_all(const char* name, F mine, G libm) {
    std::atomic<uint64_t> bad{0};
    std::atomic<uint32_t> first{0xffffffffu};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t hi = 0; hi < 65536; hi++) {
        uint64_t nb = 0;
        for (uint32_t lo = 0; lo < 65536; lo++) {
            const uint32_t u = (uint32_t)(hi << 16) | lo;
            const float x = u2f(u);
            if (!same(mine(x), libm(x))) {
                nb++;
                uint32_t f = first.load();
                while (u < f && !first.compare_exchange_weak(f, u)) {
                }
            }
        }
        bad += nb;
    }
    const uint32_t f = first.load();
    std::printf("%s %llu %llu %08x\n", name, (unsigned long long)(1ull << 32), (unsigned long long)bad.load(), f);
}

static inline uint64_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// atan2f on pairs: (a) both arguments' bits uniform; (b) the quaternion log's arguments, (|v|, w) of unit
// quaternions with w >= 0, over every magnitude of |v|; (c) y and x within a factor of 4 of each other.
static void scan_atan2(uint64_t n) {
    std::atomic<uint64_t> bad{0};
    std::atomic<uint64_t> first{~0ull};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t blk = 0; blk < 4096; blk++) {
        uint64_t nb = 0;
        for (uint64_t i = 0; i < n / 4096; i++) {
            const uint64_t r = mix((uint64_t)blk * (n / 4096) + i);
            float y, x;
            switch (i % 3) {
            case 0: y = u2f((uint32_t)r), x = u2f((uint32_t)(r >> 32)); break;
            case 1: {
                const float s = u2f((uint32_t)(r & 0x3f7fffffu)); // |v| in [0, 1)
                x = std::sqrt(std::fmax(0.0f, 1.0f - s * s));
                y = s;
                if ((r >> 40) & 1) x = std::nextafter(x, 2.0f);
                break;
            }
            default: {
                const float a = u2f((uint32_t)(0x20000000u + (r & 0x3fffffffu)));
                const float b = a * (0.25f + 3.75f * (float)((r >> 34) & 0xffffff) / 16777216.0f);
                y = (r >> 62) & 1 ? -a : a;
                x = (r >> 63) & 1 ? -b : b;
            }
            }
            if (!same(bl_atan2f(y, x), atan2f(y, x))) {
                nb++;
                const uint64_t key = ((uint64_t)f2u(y) << 32) | f2u(x);
                uint64_t f = first.load();
                while (key < f && !first.compare_exchange_weak(f, key)) {
                }
            }
        }
        bad += nb;
    }
    std::printf("atan2f %llu %llu %016llx\n", (unsigned long long)(n / 4096 * 4096), (unsigned long long)bad.load(),
                (unsigned long long)first.load());
}

int main(int argc, char** argv) {
    const char* which = argc > 1 ? argv[1] : "all";
    const bool all = !std::strcmp(which, "all");
    if (all || !std::strcmp(which, "table")) {
        // the exp2 table: 2^(i/32) rounded to double, less i << 47
        int bad = 0;
        for (int i = 0; i < 32; i++) {
            const double e = (double)exp2l((long double)i / 32.0L);
            if (d2u(e) - ((uint64_t)i << 47) != BL_EXP2_TABLE[i]) bad++;
        }
        std::printf("exp2_table 32 %d 0\n", bad);
    }
    if (all || !std::strcmp(which, "expf")) scan_all("expf", [](float x) { return bl_expf(x); }, [](float x) { return expf(x); });
    if (all || !std::strcmp(which, "pow2"))
        scan_all("pow_alpha", [](float x) { return pow_alpha(x, 2.0f); }, [](float x) { return powf(x, 2.0f); });
    if (all || !std::strcmp(which, "sinf")) scan_all("sinf", [](float x) { return bl_sinf(x); }, [](float x) { return sinf(x); });
    if (all || !std::strcmp(which, "cosf")) scan_all("cosf", [](float x) { return bl_cosf(x); }, [](float x) { return cosf(x); });
    if (all || !std::strcmp(which, "atanf")) scan_all("atanf", [](float x) { return bl_atanf(x); }, [](float x) { return atanf(x); });
    if (all || !std::strcmp(which, "acosf")) scan_all("acosf", [](float x) { return bl_acosf(x); }, [](float x) { return acosf(x); });
    if (all || !std::strcmp(which, "atan2f")) scan_atan2(argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 300000000ull);
    return 0;
}
