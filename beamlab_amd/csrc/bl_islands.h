// Host-side broadphase of the island building: every pair of body boxes that overlap, each pair once. A uniform
// 3-D grid (the cell about the median box extent, larger where the scene is sparse: at most about 2 cells a box)
// holds each box in every cell it covers; a pair is reported by the first cell (lowest x, then y, then z) that both
// boxes cover, so it comes out once. Boxes that are not finite, or cover more than kMaxCellsPerBox cells, are tested
// against every other box instead. The set of pairs is
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

// The grid's working arrays, kept by the caller from one call to the next (fresh arrays of this size cost page
// faults on every call).
struct PairScratch {
    std::vector<uint8_t> odd;
    std::vector<float> ext;
    std::vector<int64_t> lo, cnt, fill;
    std::vector<uint32_t> items;
};

template <class Emit>
void overlapping_pairs(const float* bx, size_t n, PairScratch& ws, Emit emit) {
    constexpr size_t kMaxCellsPerBox = 4096;
    std::vector<uint8_t>& odd = ws.odd;
    odd.assign(n, 0);
    std::vector<float>& ext = ws.ext;
    ext.clear();
    float lo3[3] = {INFINITY, INFINITY, INFINITY}, hi3[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; i++) {
        const float* b = bx + 6 * i;
        bool fin = true;
        for (int k = 0; k < 6; k++) fin = fin && std::isfinite(b[k]);
        if (!fin || b[0] > b[3] || b[1] > b[4] || b[2] > b[5]) {
            odd[i] = 1;
            continue;
        }
        ext.push_back(std::max(b[3] - b[0], std::max(b[4] - b[1], b[5] - b[2])));
        for (int k = 0; k < 3; k++) lo3[k] = std::min(lo3[k], b[k]), hi3[k] = std::max(hi3[k], b[3 + k]);
    }
    if (!ext.empty()) {
        std::nth_element(ext.begin(), ext.begin() + ext.size() / 2, ext.end());
        double cell = std::max<double>(ext[ext.size() / 2], 1e-3);
        // at most about 2 cells a box over the whole extent (a sparse scene - bodies spread on the ground around a
        // tall one - would otherwise spend its time on empty cells)
        double vol = 1;
        for (int k = 0; k < 3; k++) vol *= std::max<double>(hi3[k] - lo3[k], cell);
        const double cap = 2.0 * (double)n + 64;
        if (vol / (cell * cell * cell) > cap) cell = std::cbrt(vol / cap);
        const double inv = 1.0 / cell;
        int64_t nc[3];
        for (int k = 0; k < 3; k++) nc[k] = (int64_t)((hi3[k] - lo3[k]) * inv) + 1;
        auto ci = [&](int k, float v) {
            return std::min<int64_t>(nc[k] - 1, std::max<int64_t>(0, (int64_t)((v - lo3[k]) * inv)));
        };
        const int64_t ncell = nc[0] * nc[1] * nc[2];
        // counting sort of (cell, box) entries
        std::vector<int64_t>&lo = ws.lo, &cnt = ws.cnt;
        lo.resize(6 * n);
        cnt.assign((size_t)ncell + 1, 0);
        for (size_t i = 0; i < n; i++) {
            if (odd[i]) continue;
            const float* b = bx + 6 * i;
            int64_t r[6];
            for (int k = 0; k < 3; k++) r[2 * k] = ci(k, b[k]), r[2 * k + 1] = ci(k, b[3 + k]);
            if ((size_t)((r[1] - r[0] + 1) * (r[3] - r[2] + 1) * (r[5] - r[4] + 1)) > kMaxCellsPerBox) {
                odd[i] = 1;
                continue;
            }
            for (int k = 0; k < 6; k++) lo[6 * i + k] = r[k];
            for (int64_t u = r[0]; u <= r[1]; u++)
                for (int64_t v = r[2]; v <= r[3]; v++)
                    for (int64_t w = r[4]; w <= r[5]; w++) cnt[(size_t)((u * nc[1] + v) * nc[2] + w) + 1]++;
        }
        for (size_t k = 1; k < cnt.size(); k++) cnt[k] += cnt[k - 1];
        std::vector<uint32_t>& items = ws.items;
        items.resize((size_t)cnt.back());
        std::vector<int64_t>& fill = ws.fill;
        fill.assign(cnt.begin(), cnt.end() - 1);
        for (size_t i = 0; i < n; i++) {
            if (odd[i]) continue;
            const int64_t* r = &lo[6 * i];
            for (int64_t u = r[0]; u <= r[1]; u++)
                for (int64_t v = r[2]; v <= r[3]; v++)
                    for (int64_t w = r[4]; w <= r[5]; w++) items[(size_t)fill[(size_t)((u * nc[1] + v) * nc[2] + w)]++] = (uint32_t)i;
        }
        for (int64_t u = 0; u < nc[0]; u++)
            for (int64_t v = 0; v < nc[1]; v++)
                for (int64_t w = 0; w < nc[2]; w++) {
                    const size_t c = (size_t)((u * nc[1] + v) * nc[2] + w);
                    for (int64_t p = cnt[c]; p < cnt[c + 1]; p++)
                        for (int64_t q = p + 1; q < cnt[c + 1]; q++) {
                            const uint32_t i = items[(size_t)p], j = items[(size_t)q];
                            // the pair's first common cell only
                            if (std::max(lo[6 * i], lo[6 * j]) != u || std::max(lo[6 * i + 2], lo[6 * j + 2]) != v ||
                                std::max(lo[6 * i + 4], lo[6 * j + 4]) != w)
                                continue;
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
