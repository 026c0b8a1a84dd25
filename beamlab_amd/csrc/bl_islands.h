// Host-side broadphase of the island building: every pair of body boxes that overlap, each pair once. A uniform
// grid over x and z (the cell about the median box extent) holds each box in every cell its x-z span covers; a pair
// is reported by the first cell (lowest x, then z) that both boxes cover, so it comes out once. Boxes that are not
// finite, or cover more than kMaxCellsPerBox cells, are tested against every other box instead. The set of pairs is
// exactly the set the plain all-pairs test finds; only their order differs, and the islands (connected components,
// numbered by their lowest body) do not depend on it.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace blg {

// boxes: 6 floats each (min x, y, z, max x, y, z), closed intervals
inline bool boxes_overlap(const float* p, const float* q) {
    return p[0] <= q[3] && p[3] >= q[0] && p[1] <= q[4] && p[4] >= q[1] && p[2] <= q[5] && p[5] >= q[2];
}

template <class Emit>
void overlapping_pairs(const float* bx, size_t n, Emit emit) {
    constexpr size_t kMaxCellsPerBox = 4096;
    std::vector<uint8_t> odd(n, 0);
    std::vector<float> ext;
    ext.reserve(n);
    float x0 = INFINITY, z0 = INFINITY, x1 = -INFINITY, z1 = -INFINITY;
    for (size_t i = 0; i < n; i++) {
        const float* b = bx + 6 * i;
        bool fin = true;
        for (int k = 0; k < 6; k++) fin = fin && std::isfinite(b[k]);
        if (!fin || b[0] > b[3] || b[2] > b[5]) {
            odd[i] = 1;
            continue;
        }
        ext.push_back(std::max(b[3] - b[0], b[5] - b[2]));
        x0 = std::min(x0, b[0]), z0 = std::min(z0, b[2]), x1 = std::max(x1, b[3]), z1 = std::max(z1, b[5]);
    }
    if (!ext.empty()) {
        std::nth_element(ext.begin(), ext.begin() + ext.size() / 2, ext.end());
        double cell = std::max<double>(ext[ext.size() / 2], 1e-3);
        // at most about 4 cells a box on average over the whole extent
        const double span = std::max<double>(x1 - x0, 0) * std::max<double>(z1 - z0, 0);
        if (span / (cell * cell) > 4.0 * (double)n + 16) cell = std::sqrt(span / (4.0 * (double)n + 16));
        const double inv = 1.0 / cell;
        const int64_t nx = (int64_t)((x1 - x0) * inv) + 1, nz = (int64_t)((z1 - z0) * inv) + 1;
        auto cx = [&](float v) { return std::min<int64_t>(nx - 1, std::max<int64_t>(0, (int64_t)((v - x0) * inv))); };
        auto cz = [&](float v) { return std::min<int64_t>(nz - 1, std::max<int64_t>(0, (int64_t)((v - z0) * inv))); };
        // counting sort of (cell, box) entries
        std::vector<int64_t> lo(4 * n), cnt((size_t)(nx * nz) + 1, 0);
        for (size_t i = 0; i < n; i++) {
            if (odd[i]) continue;
            const float* b = bx + 6 * i;
            const int64_t a0 = cx(b[0]), a1 = cx(b[3]), c0 = cz(b[2]), c1 = cz(b[5]);
            if ((size_t)((a1 - a0 + 1) * (c1 - c0 + 1)) > kMaxCellsPerBox) {
                odd[i] = 1;
                continue;
            }
            lo[4 * i] = a0, lo[4 * i + 1] = a1, lo[4 * i + 2] = c0, lo[4 * i + 3] = c1;
            for (int64_t u = a0; u <= a1; u++)
                for (int64_t w = c0; w <= c1; w++) cnt[(size_t)(u * nz + w) + 1]++;
        }
        for (size_t k = 1; k < cnt.size(); k++) cnt[k] += cnt[k - 1];
        std::vector<uint32_t> items((size_t)cnt.back());
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (size_t i = 0; i < n; i++) {
            if (odd[i]) continue;
            for (int64_t u = lo[4 * i]; u <= lo[4 * i + 1]; u++)
                for (int64_t w = lo[4 * i + 2]; w <= lo[4 * i + 3]; w++) items[(size_t)fill[(size_t)(u * nz + w)]++] = (uint32_t)i;
        }
        for (int64_t u = 0; u < nx; u++)
            for (int64_t w = 0; w < nz; w++) {
                const size_t c = (size_t)(u * nz + w);
                for (int64_t p = cnt[c]; p < cnt[c + 1]; p++)
                    for (int64_t q = p + 1; q < cnt[c + 1]; q++) {
                        const uint32_t i = items[(size_t)p], j = items[(size_t)q];
                        // the pair's first common cell only
                        if (std::max(lo[4 * i], lo[4 * j]) != u || std::max(lo[4 * i + 2], lo[4 * j + 2]) != w) continue;
                        if (boxes_overlap(bx + 6 * (size_t)i, bx + 6 * (size_t)j)) emit(std::min(i, j), std::max(i, j));
                    }
            }
    }
    // the odd boxes against all (a pair of two odd boxes once)
    for (size_t i = 0; i < n; i++) {
        if (!odd[i]) continue;
        for (size_t j = 0; j < n; j++) {
            if (j == i || (odd[j] && j < i)) continue;
            if (boxes_overlap(bx + 6 * i, bx + 6 * j)) emit((uint32_t)std::min(i, j), (uint32_t)std::max(i, j));
        }
    }
}

} // namespace blg
