// Inter-body and self contacts of slow bodies on gfx950: the candidate-pair rebuild (spatial hash of contacter
// nodes; node-triangle, node-capsule and capsule-triangle candidates), the near-pair filter, the narrow phase and
// the order-dependent contact response.
//
// Reference semantics: World::rebuild_pairs (src/phys/world.cpp:1851-2308, without the fast bodies' own schedule,
// hull triangles and FEM mid points), World::near_pairs (:2525-2596), World::pair_detect (:2807-2962),
// World::collide_pairs / contact_force (:2964-3079, :59-91), geometry of src/phys/softbody.h:158-190 and
// src/phys/world.cpp:23-47,1135-1142.
//
// Order is part of the result: the response of a contact cancels the forces accumulated before it. Every list is
// built in the reference's order - candidates by triangle (body, then index), then hash cell (z, y, x), then the
// cell's nodes in insertion order (body, then node): each bucket is filled by atomics, and a query puts each cell's
// matches back into node order; all compactions keep their order (a scan, or a one-pass look-back). A rebuild runs
// on the device end to end (island parameters, partner boxes, hash, queries, filter) and the host waits once, for
// the list sizes. The response runs in one workgroup: contacts that share no node are independent, so each contact
// is placed at the dependency level one past the latest earlier contact sharing a node with it (a sort of
// (node, contact) keys in LDS) and each level runs in parallel - the same arithmetic, in the same per-node order, as
// the serial loop.
#include "bl_contacts.h"
#include "bl_islands.h"
#include "bl_narrow.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#pragma clang fp contract(off)

namespace blg {

__constant__ ground_model k_gm_c[BL_SURF_N] = BL_GROUND_MODELS;

// ---------------------------------------------------------------------------------------------- geometry
__device__ __forceinline__ float max_edge2(f3 a, f3 b, f3 c) {
    const f3 ba = sub(b, a), cb = sub(c, b), ac = sub(a, c);
    return smax(dot(ba, ba), smax(dot(cb, cb), dot(ac, ac)));
}
// cell_hash (src/phys/world.cpp:23-25): the int products wrap as unsigned ones do
__device__ __forceinline__ uint32_t cell_hash(int x, int y, int z) {
    return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}
__device__ __forceinline__ int cell_of(float v, float inv_cell) { return (int)floorf(v * inv_cell); }



// near_triangle (src/phys/world.cpp:1135-1142): a two-sided triangle within r of p; a hull (one-sided) triangle also
// takes p up to `depth` behind it, over its face
__device__ __forceinline__ bool near_triangle(f3 p, f3 a, f3 b, f3 c, f3 tu, float r, bool two_sided, float depth) {
    const float s = dot(sub(p, a), tu);
    f3 bary;
    if (two_sided) {
        if (!(ffabs(s) <= r)) return false;
        const f3 d = sub(closest_on_triangle(p, a, b, c, bary), p);
        return dot(d, d) <= r * r;
    }
    if (s > r || s < -r - depth) return false;
    const f3 q = sub(p, mul(tu, s));
    const f3 d = sub(closest_on_triangle(q, a, b, c, bary), q);
    return dot(d, d) <= r * r;
}

__device__ __forceinline__ bool same_group(const KArgs& a, uint32_t ga, uint32_t gb) {
    const int x = a.bodies[ga].group, y = a.bodies[gb].group;
    return x != 0 && x == y;
}

// One table for every pair island: a cell's bucket is its cell_hash salted with the island and mixed. The candidate lists
// do not depend on the table: a triangle's (or capsule's) candidates are, cell by cell in z, y, x order, the points
// of that exact cell in insertion order, whatever else shares the bucket; the capsule-triangle candidates are the
// set of triangles whose boxes meet the capsule's, sorted.
__device__ __forceinline__ uint32_t bucket_of(int x, int y, int z, int isl, uint32_t mask) {
    // (cell_hash's low bits collide in runs on lattice-shaped cell sets: a finalizing mix spreads them)
    uint32_t h = cell_hash(x, y, z) ^ ((uint32_t)isl * 0x9E3779B1u);
    h ^= h >> 16, h *= 0x85EBCA6Bu, h ^= h >> 13, h *= 0xC2B2AE35u, h ^= h >> 16;
    return h & mask;
}

__device__ __forceinline__ bool box_meets(const float* q, f3 mn, f3 mx) {
    return mn.x <= q[3] && mx.x >= q[0] && mn.y <= q[4] && mx.y >= q[1] && mn.z <= q[5] && mx.z >= q[2];
}

// Partners (src/phys/world.cpp:1936-1950): two bodies of one island, not of one collision group, whose bounds
// grown by the island's reach meet.
__device__ __forceinline__ bool partners(const KArgs& a, const CArgs& c, uint32_t g, uint32_t y) {
    if (y == g || same_group(a, g, y)) return false;
    const float* q = c.bbox + 6 * y;
    const float* p = c.bbox + 6 * g;
    return box_meets(q, mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]));
}

// near_partner: a self-colliding body, or one of its partners' boxes meets [mn, mx].
__device__ bool near_partner(const KArgs& a, const CArgs& c, uint32_t g, f3 mn, f3 mx) {
    if (a.bodies[g].self_coll) return true;
    const int m = c.bmem[g];
    const uint32_t k0 = c.poff[m], k1 = k0 + c.pcnt[m] < c.cap_part ? k0 + c.pcnt[m] : c.cap_part;
    // (the next partner's index read while this one's box is tested)
    uint32_t y = k0 < k1 ? c.plist[k0] : 0u;
    for (uint32_t k = k0; k < k1; k++) {
        const uint32_t yk = y;
        if (k + 1 < k1) y = c.plist[k + 1];
        if (box_meets(c.bbox + 6 * yk, mn, mx)) return true;
    }
    return false;
}

// near_partner by a group of G lanes with the same query (G divides 64; every lane of the group calls it): the
// partners split over the lanes, the group stopping at the first pass in which a lane finds one
template <int G>
__device__ __forceinline__ bool near_partner_group(const KArgs& a, const CArgs& c, uint32_t g, f3 mn, f3 mx) {
    if (a.bodies[g].self_coll) return true;
    const int m = c.bmem[g];
    const uint32_t k0 = c.poff[m], k1 = k0 + c.pcnt[m] < c.cap_part ? k0 + c.pcnt[m] : c.cap_part;
    const uint32_t gl = threadIdx.x & (G - 1), shift = (threadIdx.x & 63) & ~(uint32_t)(G - 1);
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    for (uint32_t kb = k0; kb < k1; kb += G) {
        const uint32_t k = kb + gl;
        const bool hit = k < k1 && box_meets(c.bbox + 6 * c.plist[k], mn, mx);
        if ((__ballot(hit) >> shift) & gmask) return true;
    }
    return false;
}

// wave-aggregated add of the counts of the last pair island (the reference's pair counter keeps the last island's);
// every lane of the wave calls it
__device__ __forceinline__ void add_last(const CArgs& c, uint32_t n) {
    uint32_t s = n;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&c.st->last_pairs, s);
}

// exclusive prefix over the wave's lanes (every lane calls it); *total: the wave's sum
__device__ __forceinline__ uint32_t wave_prefix(uint32_t v, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

// LDS traffic of one wave is processed in order: a code-motion barrier with wave-scope fences is enough between
// one lane's write and another lane's read
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------------------------- one-pass compaction
// A count, its exclusive scan and the writes at the scanned offsets in one launch. Each workgroup takes a ticket (its
// place in the output: workgroups are numbered in the order they start), counts its items, publishes its count,
// and finds the count of all earlier workgroups by looking back over their published words; a word holding an
// inclusive prefix ends the look-back. A word is (kind << 62) | (epoch << 32) | value: kind 1 a workgroup's own
// count, 2 its inclusive prefix; the epoch tells this launch's words from an earlier launch's, so the words are never
// cleared. A workgroup publishes its count before it waits, and waits only for workgroups with earlier tickets, which
// have started: every wait ends. It is bounded all the same: a time-out sets bit 2 of the rebuild's unsupported word
// (the host then fails the rebuild).
__device__ __forceinline__ uint32_t wg_ticket(const Lookback& k, uint32_t* sh) {
    // (without a ticket counter: the workgroup's index. The hardware starts the workgroups of an XCD in index
    // order and deals them to the XCDs in turn, so the lowest unfinished workgroup has started and every wait
    // ends; should a wait time out all the same, the host runs the rebuild again with tickets)
    if (!k.ticket) return blockIdx.x;
    if (threadIdx.x == 0) *sh = atomicAdd(k.ticket, 1u) - k.base;
    __syncthreads();
    return *sh;
}

// exclusive prefix over a workgroup of 256 threads (every thread calls it); agg: the workgroup's sum. sh: 4 words.
__device__ __forceinline__ uint32_t wg_prefix(uint32_t v, uint32_t& agg, uint32_t* sh) {
    uint32_t wt;
    const uint32_t wp = wave_prefix(v, wt), wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wv] = wt;
    __syncthreads();
    uint32_t bp = 0;
    agg = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++) bp += w < wv ? sh[w] : 0u, agg += sh[w];
    return bp + wp;
}

// publishes ticket t's count agg and returns the sum of the counts of the tickets before it (every thread calls
// it; the first wave looks back, 64 predecessors at a time). sh: 1 word.
__device__ uint32_t wg_lookback(const Lookback& k, uint32_t t, uint32_t agg, RStat* st, uint32_t* sh) {
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        constexpr uint64_t kEpoch = 0x3fffffff00000000ull;
        const uint64_t tag = (uint64_t)(k.epoch & 0x3fffffffu) << 32;
        uint32_t pre = 0;
        if (t == 0) {
            if (lane == 0) __hip_atomic_store(&k.words[0], (2ull << 62) | tag | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&k.words[t], (1ull << 62) | tag | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int j = (int)t - 1; // the nearest predecessor not yet summed
            uint32_t spins = 0;
            bool bad = false;
            for (;;) {
                const int idx = j - (int)lane; // lane 0 the nearest
                uint64_t w = idx >= 0 ? 0ull : ((2ull << 62) | tag); // before the first workgroup: a zero prefix
                for (;;) {
                    if (idx >= 0 && (w & kEpoch) != tag)
                        w = __hip_atomic_load(&k.words[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (__all((w & kEpoch) == tag)) break;
                    if (++spins > (1u << 22)) {
                        bad = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (bad) break;
                const uint64_t pm = __ballot((w >> 62) == 2);
                const uint32_t stop = pm ? (uint32_t)__builtin_ctzll(pm) : 63u; // lanes 0..stop are summed
                uint32_t x = lane <= stop ? (uint32_t)w : 0u;
                for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
                pre += x;
                if (pm) break;
                j -= 64;
            }
            if (bad && lane == 0) atomicOr(&st->unsupported, 4u);
            if (lane == 0)
                __hip_atomic_store(&k.words[t], (2ull << 62) | tag | (uint32_t)(pre + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) *sh = pre;
    }
    __syncthreads();
    return *sh;
}

// ---------------------------------------------------------------------------------------------- rebuild stages
__global__ void k_rb_clear(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Triangles stretched past 2.5 times their longest rest edge are torn (src/phys/world.cpp:1874-1877); and every
// 7th triangle's size sample for the cell size (:1916-1929).
__global__ void k_rb_torn(KArgs a, CArgs c) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntris) return;
    const uint4 tr = a.tri_abc[t];
    if (c.bisl[tr.w] < 0) return;
    const f3 pa = ld3(a.pm[c.parity][tr.x]), pb = ld3(a.pm[c.parity][tr.y]), pc = ld3(a.pm[c.parity][tr.z]);
    const float e2 = max_edge2(pa, pb, pc);
    if (!a.tri_torn[t] && a.tri_rest2[t] > 0 && e2 > 2.5f * 2.5f * a.tri_rest2[t]) a.tri_torn[t] = 1;
    const uint32_t k = t - c.body_tri0[tr.w];
    if (k % 7 == 0) c.samples[c.sample_base[tr.w] + k / 7] = fsqrt(e2);
}

template <class T, class Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T* sh) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    T r = sh[0];
    for (uint32_t w = 1; w < blockDim.x / 64; w++) r = op(r, sh[w]);
    return r;
}

// Each pair island's search parameters (src/phys/world.cpp:1879-1935), one workgroup per island: the largest
// collision radius, the margin from the top speed of its slow bodies (a body faster than 8 m/s is flagged: its own
// pair schedule is not stepped here), the cell size from the mean size sample. The reference sums the samples in
// double, one after another. When every partial sum is exact in double - all samples are multiples of 2^emin and
// their total stays below 2^(emin + 53) - that sequential sum is the exact sum, which an integer sum in any order
// gives; otherwise one thread adds them in the reference's order.
__global__ __launch_bounds__(256) void k_rb_island(KArgs a, CArgs c) {
    __shared__ float shf[4];
    __shared__ int shi[4];
    __shared__ unsigned long long shu[4];
    const uint32_t i = blockIdx.x, tid = threadIdx.x;
    const uint32_t k0 = c.imem_off[i], k1 = c.imem_off[i + 1];
    float rmax = 0, slow = 0;
    for (uint32_t k = k0 + tid; k < k1; k += 256) {
        const uint32_t g = c.imem[k];
        rmax = fmaxf(rmax, a.bodies[g].collision_radius);
        const float ms = fsqrt(u2f(a.body_frame[g * BF_N + BF_V2]));
        if (ms <= 8.0f) slow = fmaxf(slow, ms);
    }
    auto fmx = [](float x, float y) { return fmaxf(x, y); };
    auto imx = [](int x, int y) { return x > y ? x : y; };
    auto imn = [](int x, int y) { return x < y ? x : y; };
    auto iad = [](int x, int y) { return x + y; };
    auto uad = [](unsigned long long x, unsigned long long y) { return x + y; };
    rmax = block_reduce(rmax, fmx, shf);
    slow = block_reduce(slow, fmx, shf);
    // The island's samples: its members' ranges of the samples (stored by body, then triangle). A member with few
    // samples is swept by one thread; the members with many (up to 64 of them) by all threads, one after another.
    __shared__ uint32_t bigm[64];
    __shared__ uint32_t nbig;
    if (tid == 0) nbig = 0;
    __syncthreads();
    for (uint32_t k = k0 + tid; k < k1; k += 256) {
        const uint32_t g = c.imem[k];
        if (c.sample_base[g + 1] - c.sample_base[g] > 64) {
            const uint32_t q = atomicAdd(&nbig, 1u);
            if (q < 64) bigm[q] = g;
        }
    }
    __syncthreads();
    const uint32_t nbm = nbig < 64u ? nbig : 64u;
    auto sweep = [&](auto&& visit) {
        for (uint32_t k = k0 + tid; k < k1; k += 256) {
            const uint32_t g = c.imem[k], s0 = c.sample_base[g], s1 = c.sample_base[g + 1];
            bool listed = false;
            if (s1 - s0 > 64)
                for (uint32_t q = 0; q < nbm; q++) listed = listed || bigm[q] == g;
            if (!listed)
                for (uint32_t j = s0; j < s1; j++) visit(c.samples[j]);
        }
        for (uint32_t q = 0; q < nbm; q++) {
            const uint32_t g = bigm[q], s0 = c.sample_base[g], s1 = c.sample_base[g + 1];
            for (uint32_t j = s0 + tid; j < s1; j += 256) visit(c.samples[j]);
        }
    };
    // count, smallest unit in the last place, largest value, anything not finite
    int n = 0, emin = 1 << 20, bad = 0;
    float xmax = 0;
    sweep([&](float x) {
        n++;
        const uint32_t b = f2u(x), ex = (b >> 23) & 0xff;
        if (ex == 0xff || (b >> 31)) bad = 1;
        else if (b) emin = imn(emin, ex ? (int)ex - 150 : -149), xmax = fmaxf(xmax, x);
    });
    n = block_reduce(n, iad, shi);
    emin = block_reduce(emin, imn, shi);
    bad = block_reduce(bad, imx, shi);
    xmax = block_reduce(xmax, fmx, shf);
    double sum = 0;
    bool exact = false;
    if (!bad && n > 0 && xmax > 0) {
        int ehi;
        (void)frexp((double)n * (double)xmax * (1.0 + 1e-9), &ehi); // every partial sum is below 2^ehi
        exact = ehi - emin <= 53;
    } else if (!bad) {
        exact = true; // (no samples, or all zero)
    }
    if (exact && xmax > 0) {
        unsigned long long acc = 0;
        sweep([&](float x) {
            const uint32_t b = f2u(x), ex = (b >> 23) & 0xff;
            if (!b) return;
            const unsigned long long mant = (b & 0x7fffffu) | (ex ? 0x800000u : 0u);
            const int e = ex ? (int)ex - 150 : -149;
            acc += mant << (e - emin);
        });
        acc = block_reduce(acc, uad, shu);
        sum = ldexp((double)acc, emin);
    }
    // a fast body's extra margin (src/phys/world.cpp:1893-1915): its bulk speed - the mass-weighted RMS of its nodes'
    // speeds, +25 %, within [8, 400] m/s - times the interval; the sums in double in node order, one thread a body
    for (uint32_t k = k0 + tid; k < k1; k += 256) {
        const uint32_t g = c.imem[k];
        const float ms = fsqrt(u2f(a.body_frame[g * BF_N + BF_V2]));
        float extra = 0;
        if (ms > 8.0f) {
            const uint32_t n0 = a.ch_n0[a.body_c0[g]], n1 = a.ch_n0[a.body_c0[g + 1]];
            const float4* Vv = a.vm[c.parity];
            double mv2 = 0, m = 0;
            for (uint32_t i = n0; i < n1; i++) {
                const float4 v = Vv[i];
                mv2 += (double)v.w * (double)(v.x * v.x + v.y * v.y + v.z * v.z);
                m += (double)v.w;
            }
            const float bulk = clampf(m > 0 ? 1.25f * (float)sqrt_cr(mv2 / m) : ms, 8.0f, 400.0f);
            extra = bulk * (float)c.interval * c.dt;
        }
        c.extra[g] = extra;
    }
    if (tid != 0) return;
    if (!exact) {
        // the reference's order: the members in body order, each one's samples in order
        n = 0;
        for (uint32_t k = k0; k < k1; k++) {
            const uint32_t g = c.imem[k];
            for (uint32_t j = c.sample_base[g]; j < c.sample_base[g + 1]; j++) sum += (double)c.samples[j], n++;
        }
    }
    const float speed = slow + 0.5f;
    const float margin = 2.0f * speed * (float)c.interval * c.dt + 0.02f;
    const float mean = n ? (float)(sum / (double)n) : 1.0f;
    IslandPar ip;
    ip.margin = margin, ip.rmax = rmax, ip.reach = rmax + margin, ip.speed = speed;
    ip.inv_cell = 1.0f / clampf(mean + 2.0f * (rmax + margin), 0.25f, 1.0f);
    c.ipar[i] = ip;
}

// Each island body's partner box: its bounds grown by the island's reach.
__global__ void k_rb_boxes(KArgs a, CArgs c) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nbodies || c.bisl[g] < 0) return;
    const float r = c.ipar[c.bisl[g]].reach + c.extra[g] + a.bodies[g].hull_reach;
    const uint32_t* w = c.body_box + g * BF_N;
    for (int k = 0; k < 3; k++) c.bbox[6 * g + k] = fkey_inv(w[BF_MN + k]) - r, c.bbox[6 * g + 3 + k] = fkey_inv(w[BF_MX + k]) + r;
}

// Each island body's partners (one wave per body, lanes over its island; count, then write in body order) and
// whether it searches (a partner, or self collision).
__global__ __launch_bounds__(256) void k_rb_partners(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t m = tk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool live = m < c.nmem;
    uint32_t g = 0, k0 = 0, k1 = 0;
    if (live) {
        g = c.imem[m];
        const int isl = c.bisl[g];
        k0 = c.imem_off[isl], k1 = c.imem_off[isl + 1];
    }
    uint32_t total = 0;
    for (uint32_t y0 = k0; y0 < k1; y0 += 64) {
        const uint32_t k = y0 + lane;
        total += (uint32_t)__popcll(__ballot(k < k1 && partners(a, c, g, c.imem[k])));
    }
    uint32_t agg;
    const uint32_t wb = __shfl(wg_prefix(lane == 0 ? total : 0u, agg, sh + 1), 0);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5), base = pre + wb;
    uint32_t done = 0;
    for (uint32_t y0 = k0; y0 < k1; y0 += 64) {
        const uint32_t k = y0 + lane;
        const bool ok = k < k1 && partners(a, c, g, c.imem[k]);
        uint32_t chunk;
        const uint32_t before = wave_prefix(ok ? 1u : 0u, chunk);
        if (ok && base + done + before < c.cap_part) c.plist[base + done + before] = c.imem[k];
        done += chunk;
    }
    if (live && lane == 0) {
        c.pcnt[m] = total, c.poff[m] = base;
        c.bquery[g] = (total > 0 || a.bodies[g].self_coll) ? 1u : 0u;
    }
    if (tk == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t tot = pre + agg;
        if (tot > c.cap_part) atomicOr(&c.st->overflow, 16u), c.st->npart = tot;
    }
}

// 1) the contacter nodes of the searching bodies near a partner: their records, in node order
__global__ __launch_bounds__(256) void k_rb_pts(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t i = tk * 256 + threadIdx.x;
    bool ok = false;
    f3 p = mk(0, 0, 0);
    uint32_t g = 0;
    int isl = -1;
    if (i < a.nnodes) {
        g = a.ninfo[i] >> 16;
        isl = c.bisl[g];
        // (a fast body's nodes are not hashed: they are paired by the fast bodies' own search)
        ok = isl >= 0 && c.bquery[g] && !(c.extra[g] > 0) && (a.ninfo[i] & BL_NF_CONTACTER);
        if (ok) {
            p = ld3(a.pm[c.parity][i]);
            ok = near_partner(a, c, g, p, p);
        }
    }
    uint32_t agg;
    const uint32_t bp = wg_prefix(ok ? 1u : 0u, agg, sh + 1);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5);
    if (tk == gridDim.x - 1 && threadIdx.x == 0) c.st->npts = pre + agg;
    if (!ok) return;
    const uint32_t k = pre + bp;
    const float ic = c.ipar[isl].inv_cell;
    const int x = cell_of(p.x, ic), y = cell_of(p.y, ic), z = cell_of(p.z, ic);
    c.pcell[k] = make_int4(x, y, z, (int)i);
    c.ppos[k] = make_float4(p.x, p.y, p.z, u2f(g));
    const uint32_t b = bucket_of(x, y, z, isl, c.tmask);
    c.pbucket[k] = b;
    atomicAdd(&c.tcount[b], 1u);
}

// each point's record into a slot of its bucket, counting the bucket down to zero (so the counts are clear for the
// next rebuild; the slots are taken in no particular order: a query puts each cell's matches into node order itself)
__global__ void k_rb_fill(KArgs a, CArgs c) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= c.st->npts) return;
    const uint32_t b = c.pbucket[k];
    const uint32_t at = c.tstart[b] + atomicSub(&c.tcount[b], 1u) - 1u;
    c.scell[at] = c.pcell[k], c.spos[at] = c.ppos[k];
}

// One cell's matches, written in the order its bucket holds them, put into node order (the order the points were
// hashed in, which the lists keep): an insertion sort by one lane over the few written.
__device__ __forceinline__ void sort_by_node(uint4* out, uint32_t at, uint32_t cnt, uint32_t cap) {
    const uint32_t m = at >= cap ? 0u : (cnt < cap - at ? cnt : cap - at);
    for (uint32_t i = 1; i < m; i++) {
        const uint4 x = out[at + i];
        uint32_t j = i;
        while (j > 0 && out[at + j - 1].z > x.z) {
            out[at + j] = out[at + j - 1];
            j--;
        }
        out[at + j] = x;
    }
}

// A query's cell range: index ci in z, y, x order.
struct CellRange {
    int x0, y0, z0, nx, ny;
    uint32_t n;
    __device__ void set(f3 mn, f3 mx, float ic) {
        x0 = cell_of(mn.x, ic), y0 = cell_of(mn.y, ic), z0 = cell_of(mn.z, ic);
        nx = cell_of(mx.x, ic) - x0 + 1, ny = cell_of(mx.y, ic) - y0 + 1;
        const int nz = cell_of(mx.z, ic) - z0 + 1;
        // (over 512 cells: no query, as the reference; an empty range, from coordinates that are not finite: none)
        n = nx > 0 && ny > 0 && nz > 0 && nx * ny * nz <= 512 ? (uint32_t)(nx * ny * nz) : 0u;
    }
    __device__ void cell(uint32_t ci, int& x, int& y, int& z) const {
        x = x0 + (int)(ci % (uint32_t)nx);
        y = y0 + (int)((ci / (uint32_t)nx) % (uint32_t)ny);
        z = z0 + (int)(ci / (uint32_t)(nx * ny));
    }
};

// exclusive prefix over a group of G lanes (G divides 64; every lane of the group calls it); total: the group's sum
template <int G>
__device__ __forceinline__ uint32_t group_prefix(uint32_t v, uint32_t& total) {
    const uint32_t gl = threadIdx.x & (G - 1);
    uint32_t x = v;
    for (int off = 1; off < G; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, G);
        if (gl >= (uint32_t)off) x += y;
    }
    total = __shfl(x, G - 1, G);
    return x - v;
}

// 2) each non-torn triangle of a searching body scans its cells for nodes in reach: a group of kGnt lanes per
// triangle, a lane per cell; the lanes' matches concatenate in cell order (count, then write).
struct NtQuery {
    uint4 tr;
    uint32_t t;
    int isl;
    f3 pa, pb, pc, tmn, tmx, tu;
    float rq, margin, extra, depth; // extra: the triangle body's fast margin; depth: a hull triangle's, behind it
    bool self_coll, two_sided;
};

__device__ __forceinline__ uint32_t nt_cell(const KArgs& a, const CArgs& c, const NtQuery& q, int x, int y, int z,
                                            uint4* out, uint32_t at, uint32_t cap) {
    uint32_t cnt = 0;
    const uint32_t bk = bucket_of(x, y, z, q.isl, c.tmask);
    // (the next entry's records read while this one is tested)
    const uint32_t k0 = c.tstart[bk], k1 = c.tstart[bk + 1];
    int4 ncell = make_int4(0, 0, 0, 0);
    float4 npp = make_float4(0, 0, 0, 0);
    if (k0 < k1) ncell = c.scell[k0], npp = c.spos[k0];
    for (uint32_t k = k0; k < k1; k++) {
        const int4 cell = ncell;
        const float4 pp = npp;
        if (k + 1 < k1) ncell = c.scell[k + 1], npp = c.spos[k + 1];
        if (cell.x != x || cell.y != y || cell.z != z) continue;
        const uint32_t gbn = f2u(pp.w);
        if (gbn == q.tr.w && !q.self_coll) continue;
        if (c.bisl[gbn] != q.isl) continue;
        const float rq = q.rq;
        if (pp.x < q.tmn.x - rq || pp.y < q.tmn.y - rq || pp.z < q.tmn.z - rq || pp.x > q.tmx.x + rq ||
            pp.y > q.tmx.y + rq || pp.z > q.tmx.z + rq)
            continue;
        const uint32_t node = (uint32_t)cell.w;
        const BodyParam& B = a.bodies[gbn];
        const bool self = gbn == q.tr.w;
        if (self && (!B.self_coll || q.tr.x == node || q.tr.y == node || q.tr.z == node)) continue;
        if (!self && same_group(a, gbn, q.tr.w)) continue;
        const f3 p = mk(pp.x, pp.y, pp.z);
        const float r = B.collision_radius + q.margin + q.extra;
        const float rb = r + q.depth;
        if (p.x < q.tmn.x - rb || p.y < q.tmn.y - rb || p.z < q.tmn.z - rb || p.x > q.tmx.x + rb ||
            p.y > q.tmx.y + rb || p.z > q.tmx.z + rb)
            continue;
        if (!near_triangle(p, q.pa, q.pb, q.pc, q.tu, r, q.two_sided, q.depth)) continue;
        if (out && at + cnt < cap) out[at + cnt] = make_uint4(gbn, q.tr.w, node, q.t);
        cnt++;
    }
    if (out) sort_by_node(out, at, cnt, cap);
    return cnt;
}

constexpr int kGnt = 16;
__global__ __launch_bounds__(256, 5) void k_rb_nt(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t ti = (tk * 256 + threadIdx.x) / kGnt, gl = threadIdx.x & (kGnt - 1);
    const uint32_t t = ti < c.n_itri ? c.itri[ti] : 0u;
    NtQuery q;
    CellRange cr;
    cr.n = 0;
    if (ti < c.n_itri) {
        q.tr = a.tri_abc[t];
        q.t = t;
        q.isl = c.bisl[q.tr.w];
        if (q.isl >= 0 && c.bquery[q.tr.w] && !a.tri_torn[t] && c.st->npts > 0) {
            const IslandPar ip = c.ipar[q.isl];
            q.self_coll = a.bodies[q.tr.w].self_coll != 0;
            q.margin = ip.margin, q.extra = c.extra[q.tr.w];
            q.two_sided = !a.tri_hull[t];
            q.depth = q.two_sided ? 0.0f : a.bodies[q.tr.w].hull_depth;
            const float rq0 = ip.rmax + ip.margin + q.extra;
            q.rq = rq0 + q.depth;
            q.pa = ld3(a.pm[c.parity][q.tr.x]), q.pb = ld3(a.pm[c.parity][q.tr.y]), q.pc = ld3(a.pm[c.parity][q.tr.z]);
            q.tmn = vmin(q.pa, vmin(q.pb, q.pc)), q.tmx = vmax(q.pa, vmax(q.pb, q.pc));
            const f3 r3 = mk(q.rq, q.rq, q.rq);
            if (near_partner_group<kGnt>(a, c, q.tr.w, sub(q.tmn, r3), add(q.tmx, r3))) {
                const f3 tn = cross(sub(q.pb, q.pa), sub(q.pc, q.pa));
                const float tn2 = dot(tn, tn);
                q.tu = tn2 > 1e-20f ? mul(tn, 1.0f / fsqrt(tn2)) : mk(0, 0, 0);
                cr.set(sub(q.tmn, r3), add(q.tmx, r3), ip.inv_cell);
            }
        }
    }
    // count, then the group's place in the list, then write (a lane keeps the counts of its first two cells; those
    // of further cells are counted again)
    uint32_t total = 0, keep0 = 0, keep1 = 0;
    for (uint32_t c0 = 0; c0 < cr.n; c0 += kGnt) {
        const uint32_t ci = c0 + gl;
        uint32_t cnt = 0;
        if (ci < cr.n) {
            int x, y, z;
            cr.cell(ci, x, y, z);
            cnt = nt_cell(a, c, q, x, y, z, nullptr, 0, 0);
        }
        if (c0 == 0) keep0 = cnt;
        else if (c0 == kGnt) keep1 = cnt;
        uint32_t chunk;
        (void)group_prefix<kGnt>(cnt, chunk);
        total += chunk;
    }
    uint32_t agg;
    const uint32_t gb = __shfl(wg_prefix(gl == 0 ? total : 0u, agg, sh + 1), 0, kGnt);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5), base = pre + gb;
    uint32_t done = 0;
    for (uint32_t c0 = 0; c0 < cr.n; c0 += kGnt) {
        const uint32_t ci = c0 + gl;
        int x = 0, y = 0, z = 0;
        uint32_t cnt = 0;
        if (ci < cr.n) {
            cr.cell(ci, x, y, z);
            cnt = c0 == 0 ? keep0 : c0 == kGnt ? keep1 : nt_cell(a, c, q, x, y, z, nullptr, 0, 0);
        }
        uint32_t chunk;
        const uint32_t before = group_prefix<kGnt>(cnt, chunk);
        if (cnt) nt_cell(a, c, q, x, y, z, c.ntraw, base + done + before, c.cap_nt);
        done += chunk;
    }
    if (tk == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t tot = pre + agg;
        c.st->nnt_raw = tot;
        if (tot > c.cap_nt) atomicOr(&c.st->overflow, 1u);
    }
}

// ... and each capsule of a searching body scans for nodes within its radius (a group of kGcap lanes per capsule)
__device__ __forceinline__ uint32_t nc_cell(const KArgs& a, const CArgs& c, uint4 cp, int isl, f3 pa, f3 pb,
                                            float margin, float extra, int x, int y, int z, uint4* out, uint32_t at,
                                            uint32_t cap, uint32_t ci) {
    uint32_t cnt = 0;
    const uint32_t bk = bucket_of(x, y, z, isl, c.tmask);
    // (the next entry's records read while this one is tested)
    const uint32_t k0 = c.tstart[bk], k1 = c.tstart[bk + 1];
    int4 ncell = make_int4(0, 0, 0, 0);
    float4 npp = make_float4(0, 0, 0, 0);
    if (k0 < k1) ncell = c.scell[k0], npp = c.spos[k0];
    for (uint32_t k = k0; k < k1; k++) {
        const int4 cell = ncell;
        const float4 pp = npp;
        if (k + 1 < k1) ncell = c.scell[k + 1], npp = c.spos[k + 1];
        if (cell.x != x || cell.y != y || cell.z != z) continue;
        const uint32_t gbn = f2u(pp.w);
        if (gbn == cp.z || c.bisl[gbn] != isl) continue;
        if (same_group(a, gbn, cp.z)) continue;
        const f3 p = mk(pp.x, pp.y, pp.z);
        const float r = u2f(cp.w) + a.bodies[gbn].collision_radius + margin + extra;
        const float s = closest_on_segment(p, pa, pb);
        const f3 d = sub(add(pa, mul(sub(pb, pa), s)), p);
        if (dot(d, d) > r * r) continue;
        if (out && at + cnt < cap) out[at + cnt] = make_uint4(gbn, cp.z, (uint32_t)cell.w, ci);
        cnt++;
    }
    if (out) sort_by_node(out, at, cnt, cap);
    return cnt;
}

constexpr int kGcap = 64;
__global__ __launch_bounds__(256) void k_rb_nc(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t qi = (tk * 256 + threadIdx.x) / kGcap, gl = threadIdx.x & (kGcap - 1);
    const bool live = qi < c.n_icap;
    const uint32_t ci = live ? c.icap[qi] : 0u; // (the islands' capsules island by island: the list is island-major)
    uint4 cp = make_uint4(0, 0, 0, 0);
    int isl = -1;
    CellRange cr;
    cr.n = 0;
    f3 pa = mk(0, 0, 0), pb = mk(0, 0, 0);
    float margin = 0, extra = 0;
    if (live) {
        cp = a.cap_data[ci];
        isl = c.bisl[cp.z];
        if (isl >= 0 && c.bquery[cp.z] && c.st->npts > 0) {
            const IslandPar ip = c.ipar[isl];
            margin = ip.margin, extra = c.extra[cp.z];
            pa = ld3(a.pm[c.parity][cp.x]), pb = ld3(a.pm[c.parity][cp.y]);
            const float rq2 = u2f(cp.w) + ip.rmax + ip.margin + extra;
            const f3 mn = sub(vmin(pa, pb), mk(rq2, rq2, rq2)), mx = add(vmax(pa, pb), mk(rq2, rq2, rq2));
            if (near_partner(a, c, cp.z, mn, mx)) cr.set(mn, mx, ip.inv_cell);
        }
    }
    uint32_t total = 0, keep0 = 0; // (a lane's first count is kept for the writes)
    for (uint32_t c0 = 0; c0 < cr.n; c0 += kGcap) {
        const uint32_t cc = c0 + gl;
        uint32_t cnt = 0;
        if (cc < cr.n) {
            int x, y, z;
            cr.cell(cc, x, y, z);
            cnt = nc_cell(a, c, cp, isl, pa, pb, margin, extra, x, y, z, nullptr, 0, 0, ci);
        }
        if (c0 == 0) keep0 = cnt;
        uint32_t chunk;
        (void)group_prefix<kGcap>(cnt, chunk);
        total += chunk;
    }
    uint32_t agg;
    const uint32_t gb = __shfl(wg_prefix(gl == 0 ? total : 0u, agg, sh + 1), 0, kGcap);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5), base = pre + gb;
    uint32_t done = 0;
    for (uint32_t c0 = 0; c0 < cr.n; c0 += kGcap) {
        const uint32_t cc = c0 + gl;
        int x = 0, y = 0, z = 0;
        uint32_t cnt = 0;
        if (cc < cr.n) {
            cr.cell(cc, x, y, z);
            cnt = c0 == 0 ? keep0 : nc_cell(a, c, cp, isl, pa, pb, margin, extra, x, y, z, nullptr, 0, 0, ci);
        }
        uint32_t chunk;
        const uint32_t before = group_prefix<kGcap>(cnt, chunk);
        if (cnt) nc_cell(a, c, cp, isl, pa, pb, margin, extra, x, y, z, c.nc, base + done + before, c.cap_nc, ci);
        done += chunk;
    }
    add_last(c, (live && gl == 0 && isl == (int)c.nisl - 1) ? total : 0u);
    if (tk == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t tot = pre + agg;
        c.st->nnc = tot;
        if (tot > c.cap_nc) atomicOr(&c.st->overflow, 2u);
    }
}

// 3) capsules against triangles: the triangles of the searching bodies into every cell of their grown boxes, each
// entry carrying the triangle's body, group, island and raw bounds
__global__ __launch_bounds__(256) void k_rb_prims(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t ti = tk * 256 + threadIdx.x;
    const uint32_t t = ti < c.n_itri ? c.itri[ti] : 0u;
    // A prim is a triangle whose grown box spans at most 512 cells (the reference enters it into all of them). The
    // capsules keep the triangles whose raw box meets theirs, and such a triangle shares with the capsule a cell of
    // its raw box: entering those cells alone finds the same set.
    CellRange cr;
    cr.n = 0;
    f3 tmn = mk(0, 0, 0), tmx = mk(0, 0, 0);
    uint4 tr = make_uint4(0, 0, 0, 0);
    int isl = -1;
    if (ti < c.n_itri) {
        tr = a.tri_abc[t];
        isl = c.bisl[tr.w];
        if (isl >= 0 && c.bquery[tr.w] && !a.tri_torn[t]) {
            const IslandPar ip = c.ipar[isl];
            const float r = a.bodies[tr.w].collision_radius + ip.margin + c.extra[tr.w];
            const f3 pa = ld3(a.pm[c.parity][tr.x]), pb = ld3(a.pm[c.parity][tr.y]), pc = ld3(a.pm[c.parity][tr.z]);
            tmn = vmin(pa, vmin(pb, pc)), tmx = vmax(pa, vmax(pb, pc));
            const f3 mn = sub(tmn, mk(r, r, r)), mx = add(tmx, mk(r, r, r));
            if (near_partner(a, c, tr.w, mn, mx)) {
                cr.set(mn, mx, ip.inv_cell);
                if (cr.n) cr.set(tmn, tmx, ip.inv_cell);
            }
        }
    }
    uint32_t agg;
    const uint32_t bp = wg_prefix(cr.n, agg, sh + 1);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5);
    if (tk == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t tot = pre + agg;
        c.st->nent = tot;
        if (tot > c.cap_ent) atomicOr(&c.st->overflow, 8u);
    }
    if (cr.n == 0) return;
    const uint32_t e0 = pre + bp;
    const uint4 ra = make_uint4(t, tr.w, (uint32_t)a.bodies[tr.w].group, (uint32_t)isl);
    const float4 rb = make_float4(tmn.x, tmn.y, tmn.z, tmx.x);
    const float2 rc = make_float2(tmx.y, tmx.z);
    for (uint32_t ci = 0; ci < cr.n && e0 + ci < c.cap_ent; ci++) {
        int x, y, z;
        cr.cell(ci, x, y, z);
        const uint32_t b = bucket_of(x, y, z, isl, c.emask);
        c.eb[e0 + ci] = b;
        c.ea[e0 + ci] = ra, c.ebb[e0 + ci] = rb, c.ec[e0 + ci] = rc;
        atomicAdd(&c.ecount[b], 1u);
    }
}

__global__ void k_rb_efill(CArgs c) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = c.st->nent < c.cap_ent ? c.st->nent : c.cap_ent;
    if (e >= n) return;
    const uint32_t b = c.eb[e];
    const uint32_t at = c.estart[b] + atomicSub(&c.ecount[b], 1u) - 1u; // (down to zero, as the node hash)
    c.sa[at] = c.ea[e], c.sb[at] = c.ebb[e], c.sc[at] = c.ec[e];
}

// each capsule of an island collects the triangles of its other bodies whose boxes meet its own, once each, in
// (body, triangle) order: a group of kGcap lanes per capsule, a lane per cell, the finds gathered in LDS, sorted, made
// unique, then written at the capsule's place in the list
constexpr int kCtCap = 256;
__global__ __launch_bounds__(256) void k_rb_ct(KArgs a, CArgs c) {
    __shared__ uint32_t lst[256 / kGcap][kCtCap];
    __shared__ uint32_t uni[256 / kGcap][kCtCap];
    __shared__ uint32_t lcnt[256 / kGcap];
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t qi = (tk * 256 + threadIdx.x) / kGcap, gl = threadIdx.x & (kGcap - 1);
    const uint32_t gi = threadIdx.x / kGcap;
    const bool live = qi < c.n_icap;
    const uint32_t ci = live ? c.icap[qi] : 0u; // (the islands' capsules island by island: the list is island-major)
    uint32_t* L = lst[gi];
    uint32_t* U = uni[gi];
    uint4 cp = make_uint4(0, 0, 0, 0);
    int isl = -1;
    CellRange cr;
    cr.n = 0;
    f3 mn = mk(0, 0, 0), mx = mk(0, 0, 0);
    if (live) {
        cp = a.cap_data[ci];
        isl = c.bisl[cp.z];
        if (isl >= 0) {
            const IslandPar ip = c.ipar[isl];
            const f3 pa = ld3(a.pm[c.parity][cp.x]), pb = ld3(a.pm[c.parity][cp.y]);
            const float rr = u2f(cp.w) + a.bodies[cp.z].collision_radius + ip.margin + c.extra[cp.z];
            mn = sub(vmin(pa, pb), mk(rr, rr, rr)), mx = add(vmax(pa, pb), mk(rr, rr, rr));
            cr.set(mn, mx, ip.inv_cell);
        }
    }
    const uint32_t cg = live ? (uint32_t)a.bodies[cp.z].group : 0u;
    if (gl == 0) lcnt[gi] = 0;
    wave_sync();
    for (uint32_t c0 = 0; c0 < cr.n; c0 += kGcap) {
        const uint32_t cc = c0 + gl;
        if (cc < cr.n) {
            int x, y, z;
            cr.cell(cc, x, y, z);
            const uint32_t bk = bucket_of(x, y, z, isl, c.emask);
            // (an entry's three records are read together, and the next entry's while this one is tested)
            const uint32_t k0 = c.estart[bk], k1 = c.estart[bk + 1];
            uint4 na = make_uint4(0, 0, 0, 0);
            float4 nb = make_float4(0, 0, 0, 0);
            float2 nc2 = make_float2(0, 0);
            if (k0 < k1) na = c.sa[k0], nb = c.sb[k0], nc2 = c.sc[k0];
            for (uint32_t k = k0; k < k1; k++) {
                const uint4 ra = na;
                const float4 rb = nb;
                const float2 rc = nc2;
                if (k + 1 < k1) na = c.sa[k + 1], nb = c.sb[k + 1], nc2 = c.sc[k + 1];
                if ((int)ra.w != isl || ra.y == cp.z || (cg != 0 && cg == ra.z)) continue;
                if (rb.w < mn.x || rc.x < mn.y || rc.y < mn.z || rb.x > mx.x || rb.y > mx.y || rb.z > mx.z) continue;
                const uint32_t at = atomicAdd(&lcnt[gi], 1u);
                if (at < (uint32_t)kCtCap) L[at] = ra.x;
            }
        }
    }
    wave_sync();
    const uint32_t raw = lcnt[gi];
    if (raw > (uint32_t)kCtCap && gl == 0) atomicOr(&c.st->unsupported, 2u);
    const uint32_t n = raw < (uint32_t)kCtCap ? raw : (uint32_t)kCtCap;
    // bitonic sort of the finds (padded to a power of two)
    uint32_t np = 1;
    while (np < n) np <<= 1;
    for (uint32_t k = n + gl; k < np; k += kGcap) L[k] = 0xffffffffu;
    wave_sync();
    for (uint32_t size = 2; size <= np; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t q = gl; q < np / 2; q += kGcap) {
                const uint32_t i = 2 * q - (q & (stride - 1)), j = i + stride;
                const bool up = (i & size) == 0;
                const uint32_t x = L[i], y = L[j];
                if ((x > y) == up) L[i] = y, L[j] = x;
            }
            wave_sync();
        }
    // the distinct ones, in order
    uint32_t total = 0;
    for (uint32_t j0 = 0; j0 < n; j0 += kGcap) {
        const uint32_t j = j0 + gl;
        const bool keep = j < n && (j == 0 || L[j] != L[j - 1]);
        uint32_t chunk;
        const uint32_t before = group_prefix<kGcap>(keep ? 1u : 0u, chunk);
        if (keep) U[total + before] = L[j];
        total += chunk;
    }
    uint32_t agg;
    const uint32_t gb = __shfl(wg_prefix(gl == 0 ? total : 0u, agg, sh + 1), 0, kGcap);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5), base = pre + gb;
    if (live) {
        for (uint32_t j = gl; j < total; j += kGcap) {
            const uint32_t t = U[j];
            if (base + j < c.cap_ct) c.ct[base + j] = make_uint4(cp.z, a.tri_abc[t].w, ci, t);
        }
    }
    add_last(c, (live && gl == 0 && isl == (int)c.nisl - 1) ? total : 0u);
    if (tk == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t tot = pre + agg;
        c.st->nct = tot;
        if (tot > c.cap_ct) atomicOr(&c.st->overflow, 4u);
    }
}

// near_pairs (src/phys/world.cpp:2525-2596): the node-triangle candidates within reach of their relative travel
// until the next rebuild and not receding across the plane, kept in order
__global__ __launch_bounds__(256) void k_rb_near(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t i = tk * 256 + threadIdx.x;
    // (with fast candidates, the list merged island by island)
    const uint32_t nf = min(c.st->nfnt, c.cap_fnt), ns = min(c.st->nnt_raw, c.cap_nt), n = ns + nf;
    const uint4* src = nf ? c.ntm : c.ntraw;
    bool keep = false;
    uint4 pr = make_uint4(0, 0, 0, 0);
    if (i < n) {
        pr = src[i];
        keep = !a.tri_torn[pr.w];
        if (keep) {
            const uint4 tr = a.tri_abc[pr.w];
            const float4* P = a.pm[c.parity];
            const float4* Vv = a.vm[c.parity];
            const f3 xp = ld3(P[pr.z]), xv = ld3(Vv[pr.z]);
            const f3 ap = ld3(P[tr.x]), bp = ld3(P[tr.y]), cp = ld3(P[tr.z]);
            const f3 av = ld3(Vv[tr.x]), bv = ld3(Vv[tr.y]), cv = ld3(Vv[tr.z]);
            f3 bary;
            const f3 q = closest_on_triangle(xp, ap, bp, cp, bary);
            const float cr = a.bodies[pr.x].collision_radius;
            const bool two_sided = !a.tri_hull[pr.w];
            const float r0 = cr + (two_sided ? 0.0f : a.bodies[pr.y].hull_depth);
            const f3 da = sub(xv, av), db = sub(xv, bv), dc = sub(xv, cv);
            const float vrel = fsqrt(smax(dot(da, da), smax(dot(db, db), dot(dc, dc))));
            const float reach = r0 + 1.25f * vrel * c.horizon + 0.02f;
            const f3 xq = sub(xp, q);
            if (dot(xq, xq) >= reach * reach) keep = false;
            else {
                const f3 nn = cross(sub(bp, ap), sub(cp, ap));
                const float nl = length(nn);
                if (nl > 1e-12f) {
                    const f3 n0 = divs(nn, nl);
                    const float s0 = dot(n0, sub(xp, ap));
                    // (a hull's triangle only from in front)
                    if (two_sided || s0 > cr) {
                        const float sg = s0 >= 0 ? 1.0f : -1.0f;
                        const float close = smax(0.0f, smax(-sg * dot(n0, da), smax(-sg * dot(n0, db), -sg * dot(n0, dc))));
                        if (ffabs(s0) - 1.25f * close * c.horizon - 0.02f >= cr) keep = false;
                    }
                }
            }
        }
    }
    uint32_t agg;
    const uint32_t bp = wg_prefix(keep ? 1u : 0u, agg, sh + 1);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5);
    if (keep) c.nt[pre + bp] = pr;
    add_last(c, keep && c.bisl[pr.y] == (int)c.nisl - 1 ? 1u : 0u);
    if (tk == gridDim.x - 1 && threadIdx.x == 0) c.st->nnt = pre + agg;
}

// ---------------------------------------------------------------------------------------------- fast bodies
// World::fast_pairs (src/phys/world.cpp:2310-2522). The margins: a node's own travel until the next refresh (its
// speed x 1.25 intervals), a slow partner's reach (twice the slow bodies' speed x the interval + 1 cm), a fast
// partner's own extra margin (+ 1 cm, + its triangle's corners' travel). (The fast pairs are refreshed with the
// rest, every interval: the fast interval is the slow one.)
struct FastMargins {
    float travel, slow_reach;
};
__device__ __forceinline__ FastMargins fast_margins(const CArgs& c, int isl) {
    FastMargins f;
    f.travel = 1.25f * (float)c.interval * c.dt;
    f.slow_reach = 2.0f * c.ipar[isl].speed * (float)c.interval * c.dt + 0.01f;
    return f;
}
__device__ __forceinline__ f3 body_mn(const KArgs& a, const CArgs& c, uint32_t g) {
    const uint32_t* w = c.body_box + g * BF_N;
    return mk(fkey_inv(w[BF_MN]), fkey_inv(w[BF_MN + 1]), fkey_inv(w[BF_MN + 2]));
}
__device__ __forceinline__ f3 body_mx(const KArgs& a, const CArgs& c, uint32_t g) {
    const uint32_t* w = c.body_box + g * BF_N;
    return mk(fkey_inv(w[BF_MX]), fkey_inv(w[BF_MX + 1]), fkey_inv(w[BF_MX + 2]));
}
__device__ __forceinline__ float body_speed(const KArgs& a, uint32_t g) { return fsqrt(u2f(a.body_frame[g * BF_N + BF_V2])); }

// Each island member's fast queries (a fast body: its partners' triangles, then their capsules; else none) and, for a
// fast body, its grid: cells about its widest reach (at least 15 cm, at most 2 m, at most 48 an axis) over its
// bounds. One workgroup; the members in chunks, the query counts scanned into fqoff.
__global__ __launch_bounds__(1024) void k_rb_fgrid(KArgs a, CArgs c) {
    __shared__ uint32_t wsum[16], carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t m0 = 0; m0 < c.nmem; m0 += 1024) {
        const uint32_t m = m0 + tid;
        uint32_t nq = 0;
        if (m < c.nmem) {
            const uint32_t g = c.imem[m];
            if (c.extra[g] > 0) {
                const int isl = c.bisl[g];
                const FastMargins fm = fast_margins(c, isl);
                const float cr = a.bodies[g].collision_radius;
                const uint32_t k0 = c.poff[m], k1 = min(k0 + c.pcnt[m], c.cap_part);
                float rmax = 0;
                for (uint32_t k = k0; k < k1; k++) {
                    const uint32_t bo = c.plist[k];
                    rmax = smax(rmax, cr + fm.slow_reach + c.extra[bo]);
                    nq += (c.body_tri0[bo + 1] - c.body_tri0[bo]) + (c.body_cap0[bo + 1] - c.body_cap0[bo]);
                }
                const f3 mn = body_mn(a, c, g), mx = body_mx(a, c, g);
                const f3 ext = add(sub(mx, mn), mk(2e-3f, 2e-3f, 2e-3f));
                float cellsz = clampf(2.0f * rmax, 0.15f, 2.0f);
                cellsz = smax(cellsz, smax(ext.x, smax(ext.y, ext.z)) / 48.0f);
                FastGrid G;
                G.inv = 1.0f / cellsz;
                G.ox = mn.x - 1e-3f, G.oy = mn.y - 1e-3f, G.oz = mn.z - 1e-3f;
                G.dx = max(1, min(48, (int)(ext.x * G.inv) + 1));
                G.dy = max(1, min(48, (int)(ext.y * G.inv) + 1));
                G.dz = max(1, min(48, (int)(ext.z * G.inv) + 1));
                G.pad = 0;
                c.fgrid[m] = G;
            }
        }
        // exclusive scan of the counts
        uint32_t x = nq;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        uint32_t pre = carry + x - nq;
        for (uint32_t w = 0; w < wv; w++) pre += wsum[w];
        if (m < c.nmem) c.fqoff[m] = pre;
        __syncthreads();
        if (tid == 1023) carry = pre + nq;
        __syncthreads();
    }
    if (tid == 0) {
        c.fqoff[c.nmem] = carry;
        c.st->nfq = carry;
        if (carry > c.cap_fq) atomicOr(&c.st->overflow, 32u);
    }
}

// The cell of a coordinate along one axis of a fast body's grid, clamped (Grid::cell_index's per axis)
__device__ __forceinline__ int fg_axis(float v, float o, float inv, int d) { return min(max((int)((v - o) * inv), 0), d - 1); }

// A fast body's node test against one query: its cells (between its position -+ its own travel) meet the query's
// cells [lo, hi], and it passes the query's test
struct FastQuery {
    int lo[3], hi[3];
    bool tri, live;
    uint32_t g, bo, prim;   // the fast body, the partner, the triangle or capsule (global index)
    f3 a, b, cc, tmn, tmx, tu;
    float rt, depth, rc0;   // a triangle's reach and hull depth; a capsule's
    bool two_sided;
};

__device__ __forceinline__ bool fast_node(const KArgs& a, const CArgs& c, const FastGrid& G, const FastQuery& q,
                                          float travel, uint32_t ni) {
    if (!(a.ninfo[ni] & BL_NF_CONTACTER)) return false;
    const f3 p = ld3(a.pm[c.parity][ni]), v = ld3(a.vm[c.parity][ni]);
    const float mv = length(v) * travel;
    const float o[3] = {G.ox, G.oy, G.oz};
    const int d[3] = {G.dx, G.dy, G.dz};
    const float pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int l = fg_axis(pc[k] - mv, o[k], G.inv, d[k]), h = fg_axis(pc[k] + mv, o[k], G.inv, d[k]);
        if (h < q.lo[k] || l > q.hi[k]) return false;
    }
    if (q.tri) {
        const float rn = q.rt + length(v) * travel, rb = rn + q.depth;
        if (p.x < q.tmn.x - rb || p.y < q.tmn.y - rb || p.z < q.tmn.z - rb || p.x > q.tmx.x + rb || p.y > q.tmx.y + rb ||
            p.z > q.tmx.z + rb)
            return false;
        return near_triangle(p, q.a, q.b, q.cc, q.tu, rn, q.two_sided, q.depth);
    }
    const float rn = q.rc0 + length(v) * travel;
    const float sg = closest_on_segment(p, q.a, q.b);
    const f3 dd = sub(add(q.a, mul(sub(q.b, q.a), sg)), p);
    return dot(dd, dd) <= rn * rn;
}

// Each fast query scans its fast body's nodes, in order, for the ones found in its cells within reach: a group of
// kGf lanes per query, a lane per node (count, then the group's place in the list, then write). CAPS: the capsule
// queries (node-capsule candidates) instead of the triangle ones.
constexpr int kGf = 16;
template <bool CAPS>
__global__ __launch_bounds__(256) void k_rb_fast(KArgs a, CArgs c) {
    __shared__ uint32_t sh[8];
    const uint32_t tk = wg_ticket(c.lk, sh);
    const uint32_t qi = (tk * 256 + threadIdx.x) / kGf, gl = threadIdx.x & (kGf - 1);
    const uint32_t nq = min(c.st->nfq, c.cap_fq);
    FastQuery q;
    q.live = false;
    FastGrid G{};
    uint32_t n0 = 0, n1 = 0;
    float travel = 0;
    int isl = -1;
    if (qi < nq) {
        // the member: the last whose first query is at most qi; then the partner and the primitive
        uint32_t lo = 0, hi = c.nmem;
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (c.fqoff[mid] <= qi) lo = mid;
            else hi = mid;
        }
        const uint32_t m = lo, g = c.imem[m];
        uint32_t l = qi - c.fqoff[m];
        const uint32_t k0 = c.poff[m], k1 = min(k0 + c.pcnt[m], c.cap_part);
        uint32_t bo = 0;
        bool tri = false, found = false;
        for (uint32_t k = k0; k < k1 && !found; k++) {
            bo = c.plist[k];
            const uint32_t nt = c.body_tri0[bo + 1] - c.body_tri0[bo], ncp = c.body_cap0[bo + 1] - c.body_cap0[bo];
            if (l < nt) tri = true, found = true;
            else if (l < nt + ncp) l -= nt, found = true;
            else l -= nt + ncp;
        }
        isl = c.bisl[g];
        if (found && tri != CAPS) {
            G = c.fgrid[m];
            n0 = a.ch_n0[a.body_c0[g]], n1 = a.ch_n0[a.body_c0[g + 1]];
            const FastMargins fm = fast_margins(c, isl);
            travel = fm.travel;
            const float cr = a.bodies[g].collision_radius;
            const bool o_fast = c.extra[bo] > 0;
            const float fe = c.extra[bo];
            const float r0 = cr + (o_fast ? fe + 0.01f : fm.slow_reach);
            const float r = cr + fm.slow_reach + travel * (body_speed(a, g) + (o_fast ? body_speed(a, bo) : 0.0f)) + (o_fast ? fe : 0.0f);
            const f3 fmn = body_mn(a, c, g), fmx = body_mx(a, c, g);
            q.g = g, q.bo = bo, q.tri = tri;
            f3 smn, smx;
            if (tri) {
                const uint32_t t = c.body_tri0[bo] + l;
                q.prim = t;
                const uint4 tr = a.tri_abc[t];
                const f3 fbmn = sub(fmn, mk(r, r, r)), fbmx = add(fmx, mk(r, r, r));
                if (!a.tri_torn[t]) {
                    q.a = ld3(a.pm[c.parity][tr.x]), q.b = ld3(a.pm[c.parity][tr.y]), q.cc = ld3(a.pm[c.parity][tr.z]);
                    q.tmn = vmin(q.a, vmin(q.b, q.cc)), q.tmx = vmax(q.a, vmax(q.b, q.cc));
                    const f3 &tmn = q.tmn, &tmx = q.tmx;
                    if (!(tmx.x < fbmn.x || tmx.y < fbmn.y || tmx.z < fbmn.z || tmn.x > fbmx.x || tmn.y > fbmx.y || tmn.z > fbmx.z)) {
                        const float r00 = cr + (o_fast ? 0.01f : fm.slow_reach);
                        float rt = r00;
                        if (o_fast) {
                            const f3 va = ld3(a.vm[c.parity][tr.x]), vb = ld3(a.vm[c.parity][tr.y]), vc = ld3(a.vm[c.parity][tr.z]);
                            rt = r00 + travel * fsqrt(smax(dot(va, va), smax(dot(vb, vb), dot(vc, vc))));
                        }
                        q.rt = rt;
                        q.two_sided = !a.tri_hull[t];
                        q.depth = q.two_sided ? 0.0f : a.bodies[bo].hull_depth;
                        const f3 tn = cross(sub(q.b, q.a), sub(q.cc, q.a));
                        const float tn2 = dot(tn, tn);
                        q.tu = tn2 > 1e-20f ? mul(tn, 1.0f / fsqrt(tn2)) : mk(0, 0, 0);
                        const float e = rt + q.depth;
                        smn = sub(tmn, mk(e, e, e)), smx = add(tmx, mk(e, e, e));
                        q.live = true;
                    }
                }
            } else {
                const uint32_t ci = c.body_cap0[bo] + l;
                q.prim = ci;
                const uint4 cp = a.cap_data[ci];
                q.a = ld3(a.pm[c.parity][cp.x]), q.b = ld3(a.pm[c.parity][cp.y]);
                const float rc = u2f(cp.w) + r;
                f3 cmn = sub(vmin(q.a, q.b), mk(rc, rc, rc)), cmx = add(vmax(q.a, q.b), mk(rc, rc, rc));
                if (!(cmx.x < fmn.x || cmx.y < fmn.y || cmx.z < fmn.z || cmn.x > fmx.x || cmn.y > fmx.y || cmn.z > fmx.z)) {
                    q.rc0 = u2f(cp.w) + r0;
                    smn = sub(vmin(q.a, q.b), mk(q.rc0, q.rc0, q.rc0)), smx = add(vmax(q.a, q.b), mk(q.rc0, q.rc0, q.rc0));
                    q.live = true;
                }
            }
            if (q.live) {
                // the query's cells (Grid scan: floor below, truncation above, clamped; none when it misses the grid)
                const float mnv[3] = {smn.x, smn.y, smn.z}, mxv[3] = {smx.x, smx.y, smx.z}, o[3] = {G.ox, G.oy, G.oz};
                const int d[3] = {G.dx, G.dy, G.dz};
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    q.lo[k] = max(0, (int)floorf((mnv[k] - o[k]) * G.inv));
                    q.hi[k] = min(d[k] - 1, (int)((mxv[k] - o[k]) * G.inv));
                    if (q.lo[k] > q.hi[k]) q.live = false;
                }
            }
        }
    }
    uint32_t total = 0;
    if (q.live)
        for (uint32_t i0 = n0; i0 < n1; i0 += kGf) {
            const uint32_t ni = i0 + gl;
            const bool ok = ni < n1 && fast_node(a, c, G, q, travel, ni);
            uint32_t chunk;
            (void)group_prefix<kGf>(ok ? 1u : 0u, chunk);
            total += chunk;
        }
    uint32_t agg;
    const uint32_t gb = __shfl(wg_prefix(gl == 0 ? total : 0u, agg, sh + 1), 0, kGf);
    const uint32_t pre = wg_lookback(c.lk, tk, agg, c.st, sh + 5), base = pre + gb;
    uint4* out = CAPS ? c.fnc : c.fnt;
    const uint32_t cap = CAPS ? c.cap_fnc : c.cap_fnt;
    uint32_t done = 0;
    if (q.live && total)
        for (uint32_t i0 = n0; i0 < n1; i0 += kGf) {
            const uint32_t ni = i0 + gl;
            const bool ok = ni < n1 && fast_node(a, c, G, q, travel, ni);
            uint32_t chunk;
            const uint32_t before = group_prefix<kGf>(ok ? 1u : 0u, chunk);
            if (ok && base + done + before < cap) out[base + done + before] = make_uint4(q.g, q.bo, ni, q.prim);
            done += chunk;
        }
    if (CAPS) add_last(c, (q.live && gl == 0 && isl == (int)c.nisl - 1) ? total : 0u);
    if (tk == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t tot = pre + agg;
        if (CAPS) {
            c.st->nfnc = tot;
            c.st->nnc_slow = c.st->nnc;
            c.st->nnc = c.st->nnc + tot;
            if (tot > c.cap_fnc) atomicOr(&c.st->overflow, 128u);
        } else {
            c.st->nfnt = tot;
            if (tot > c.cap_fnt) atomicOr(&c.st->overflow, 64u);
        }
    }
}

// The first entry of a list (island-major) of pair island isl or later
__device__ __forceinline__ uint32_t isl_lower(const int* bisl, const uint4* L, uint32_t n, int isl) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bisl[L[mid].x] < isl) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// With fast candidates: the node-triangle (raw) and node-capsule lists merged island by island, each island's
// slow candidates, then its fast ones. Entry j of an island's slow run moves past the fast candidates of the earlier
// islands; fast entry j past the slow candidates of the islands up to its own.
__global__ __launch_bounds__(256) void k_rb_splice(KArgs a, CArgs c) {
    (void)a;
    const uint32_t nft = min(c.st->nfnt, c.cap_fnt), nfc = min(c.st->nfnc, c.cap_fnc);
    if (nft == 0 && nfc == 0) return;
    const uint32_t nst = min(c.st->nnt_raw, c.cap_nt), nsc = min(c.st->nnc_slow, c.cap_nc);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t segs[4] = {nft ? nst : 0u, nft, nfc ? nsc : 0u, nfc};
    if (i < segs[0]) {
        const uint4 e = c.ntraw[i];
        c.ntm[i + isl_lower(c.bisl, c.fnt, nft, c.bisl[e.x])] = e;
        return;
    }
    i -= segs[0];
    if (i < segs[1]) {
        const uint4 e = c.fnt[i];
        c.ntm[i + isl_lower(c.bisl, c.ntraw, nst, c.bisl[e.x] + 1)] = e;
        return;
    }
    i -= segs[1];
    if (i < segs[2]) {
        const uint4 e = c.nc[i];
        c.ncm[i + isl_lower(c.bisl, c.fnc, nfc, c.bisl[e.x])] = e;
        return;
    }
    i -= segs[2];
    if (i < segs[3]) {
        const uint4 e = c.fnc[i];
        c.ncm[i + isl_lower(c.bisl, c.nc, nsc, c.bisl[e.x] + 1)] = e;
    }
}

// Each response group's run in the three final lists (NT after the filter, NC, CT): the searches go over the islands'
// triangles and capsules island by island, so an island's candidates - and a group's, its islands being consecutive -
// are one run in each list; its bounds are where the group changes. (A list whose islands are not in order would break
// that: flagged, the rebuild fails.)
__device__ __forceinline__ void island_run(const int* bisl, const int* igroup, const uint4* L, uint32_t n, uint32_t i,
                                           uint32_t* irange, int which, RStat* st) {
    if (i >= n) return;
    auto group = [&](uint32_t k) { const int b = bisl[L[k].x]; return b < 0 ? -1 : igroup[b]; };
    const int isl = group(i);
    const int prev = i > 0 ? group(i - 1) : -1;
    const int next = i + 1 < n ? group(i + 1) : -1;
    if (isl < 0 || prev > isl) atomicOr(&st->unsupported, 8u);
    if (isl < 0) return;
    if (isl != prev) irange[6 * isl + 2 * which] = i;
    if (isl != next) irange[6 * isl + 2 * which + 1] = i + 1;
}

__global__ void k_rb_ranges(KArgs a, CArgs c) {
    (void)a;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) c.st->nisl = c.ngroup;
    const uint32_t nnt = min(c.st->nnt, c.cap_nt + c.cap_fnt);
    const uint32_t nnc = min(c.st->nnc, c.st->nfnc ? c.cap_nc + c.cap_fnc : c.cap_nc);
    const uint32_t nct = min(c.st->nct, c.cap_ct);
    island_run(c.bisl, c.igroup, c.nt, nnt, i, c.irange, 0, c.st);
    island_run(c.bisl, c.igroup, c.st->nfnc ? c.ncm : c.nc, nnc, i, c.irange, 1, c.st);
    island_run(c.bisl, c.igroup, c.ct, nct, i, c.irange, 2, c.st);
}

// ---------------------------------------------------------------------------------------------- narrow phase
// pair_detect (src/phys/world.cpp:2807-2962) on the concatenated candidates (node-triangle, node-capsule,
// capsule-triangle), 256 per block: each block's hits packed to the front of its 256 slots, in candidate order,
// and its count
__global__ __launch_bounds__(256) void k_narrow(KArgs a, PArgs p, int parity) { narrow_block(a, p, parity, blockIdx.x); }


// ---------------------------------------------------------------------------------------------- response
// The nodes and weights of a contact (collide_pairs, src/phys/world.cpp:2997-3076): side A then side B.
struct ContactNodes {
    uint32_t n[5];
    float w[5];
    int na, nb;
    float fric;
    bool hull; // a node behind a hull triangle's face, deeper than its contact radius: hull_push_out
};

// Every branch sets all five slots (the unused ones to node 0, weight 0): stores of a slot that only some branches
// make are merged by the compiler into one store through a selected address, which puts the slots in scratch memory.
__device__ __forceinline__ void set_slots(ContactNodes& cn, uint32_t n0, uint32_t n1, uint32_t n2, uint32_t n3,
                                          uint32_t n4, float w0, float w1, float w2, float w3, float w4) {
    cn.n[0] = n0, cn.n[1] = n1, cn.n[2] = n2, cn.n[3] = n3, cn.n[4] = n4;
    cn.w[0] = w0, cn.w[1] = w1, cn.w[2] = w2, cn.w[3] = w3, cn.w[4] = w4;
}

// (from the hit alone: the narrow phase wrote the candidate's nodes, friction and shape into it)
__device__ __forceinline__ void contact_nodes(const Hit& h, ContactNodes& cn) {
    cn.na = (int)(h.shape & 15u), cn.nb = (int)((h.shape >> 4) & 15u), cn.hull = (h.shape >> 8) & 1u;
    cn.fric = h.fric;
    if (cn.na == 1 && cn.nb == 3) // node-triangle
        set_slots(cn, h.n[0], h.n[1], h.n[2], h.n[3], 0u, 1.0f, h.bary.x, h.bary.y, h.bary.z, 0.0f);
    else if (cn.na == 1) // node-capsule
        set_slots(cn, h.n[0], h.n[1], h.n[2], 0u, 0u, 1.0f, 1 - h.s, h.s, 0.0f, 0.0f);
    else // capsule-triangle
        set_slots(cn, h.n[0], h.n[1], h.n[2], h.n[3], h.n[4], 1 - h.s, h.s, h.bary.x, h.bary.y, h.bary.z);
}

// the concrete model of the inter-body contacts, as compile-time constants
constexpr ground_model kGmAll[BL_SURF_N] = BL_GROUND_MODELS;
constexpr ground_model kGmContact = kGmAll[BL_SURF_CONCRETE];
static_assert(kGmContact.soft_depth == 0, "the contact model has no fluid layer");

// contact_force<NA, NB> (src/phys/world.cpp:59-91, primitive_collision src/phys/static_world.cpp:8-51) and
// hull_push_out<3> (src/phys/world.cpp:93-120), both sides movable, no force snapshots, split at the forces: the terms
// that read only velocities, inverse masses and the geometry (which the response does not change) are computed once
// per hit, before the window's rounds; each round then does what reads the forces accumulated so far, the same
// operations in the same order. The side a nodes come first (slots 0 .. na - 1), then side b's.
struct CfPre {
    float m, Vn, roff, sv, stop, e_static, g_slide;
    float t_push; // hull_push_out: (target - vr) m / dt
    f3 vel, slip;
    float wim[5]; // w_i x inverse mass_i
    bool ok;
};

__device__ __forceinline__ void contact_force_pre(const ContactNodes& cn, const float* im, const f3* vel, f3 nrm,
                                                  float pen, float dt, CfPre& q) {
    constexpr ground_model gm = kGmContact;
    const int na = cn.na, nn = cn.na + cn.nb;
    // (hull_push_out's W - its node's inverse mass, then the triangle's weighted - is this sum with w_0 = 1)
    float W = 0;
#pragma unroll
    for (int i = 0; i < 5; i++)
        if (i < nn) W += cn.w[i] * cn.w[i] * im[i];
    q.ok = !(W <= 1e-12f);
    q.m = 1.0f / W;
    f3 va = mk(0, 0, 0), vb = mk(0, 0, 0);
#pragma unroll
    for (int i = 0; i < 5; i++) {
        q.wim[i] = cn.w[i] * im[i];
        if (i < na) va = add(va, mul(vel[i], cn.w[i]));
        else if (i < nn) vb = add(vb, mul(vel[i], cn.w[i]));
    }
    // hull_push_out: the push-out speed (a fifth of the depth a step, at most 2 m/s) against the relative speed
    const float vr = dot(sub(vel[0], vb), nrm), target = smin(0.2f * pen / dt, 2.0f);
    q.t_push = (target - vr) * q.m / dt;
    // primitive_collision(f_rel, va - vb, m, nrm, dt, concrete, pen, fric, push_max 2, slop 0, restitution -1)
    const f3 v = sub(va, vb);
    q.vel = v;
    q.Vn = dot(v, nrm);
    const float sgl = gm.soft_depth, take = 0.8f;
    q.roff = (take * q.Vn - smin(0.2f * smax(0.0f, pen - sgl - 0.0f) / dt, 2.0f)) * q.m / dt;
    f3 slip = sub(v, mul(nrm, q.Vn));
    q.sv = length(slip);
    q.slip = q.sv > 1e-9f ? divs(slip, q.sv) : mk(0, 0, 0);
    q.stop = q.m * q.sv / dt;
    q.e_static = 1.0f - bl_expf(-q.sv / gm.va);
    const float g = gm.mc + (gm.ms - gm.mc) * bl_expf(-pow_alpha(q.sv / gm.vs, gm.alpha));
    q.g_slide = g + smin(gm.t2 * q.sv, 5.0f);
}

// the LDS cell of the slots a contact does not use (slots of other contacts are 0 .. 5 kRespW - 1)
constexpr uint32_t kRespDummy = 5 * 512;

__device__ __forceinline__ float selp(bool c, float x, float y) { return c ? x : y; }
__device__ __forceinline__ f3 selp(bool c, f3 x, f3 y) { return mk(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z); }

// The rest of contact_force (or hull_push_out) on the forces in LDS, without branches: slot i of the contact reads
// and writes fs[3 root[i] ..], the one LDS copy of its node's force (an unused slot the dummy cell, its result not
// taken). The forces are all read first, as contact_force reads them all before it writes; a slot after another of
// the same node in the contact (a ball's capsule of no length: dup[i], else -1) starts its write from that one's
// result.
__device__ __forceinline__ void contact_force_apply(const ContactNodes& cn, const CfPre& q, const uint16_t* root,
                                                    const int* dup, float* fs, f3 nrm, float pen) {
    constexpr ground_model gm = kGmContact;
    const int na = cn.na, nn = cn.na + cn.nb;
    f3 Fi[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const float* F = fs + 3 * root[i];
        Fi[i] = mk(F[0], F[1], F[2]);
    }
    f3 acc = mk(0, 0, 0);
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const f3 x = mul(Fi[i], q.wim[i]);
        acc = selp(i < na, add(acc, x), selp(i < nn, sub(acc, x), acc));
    }
    const f3 F = mul(acc, q.m);
    const float Fn = dot(F, nrm);
    float R = -Fn;
    R = selp(q.Vn < 0, R - q.roff, R);
    const f3 slipF = sub(F, mul(nrm, Fn));
    const float G = R * gm.strength * cn.fric;
    const bool stat = q.sv < gm.va && G > 0 && length(slipF) <= gm.ms * G;
    const f3 o_stat = sub(sub(mul(nrm, R), mul(q.slip, smin(gm.ms * G * q.e_static, q.stop))), slipF);
    const f3 o_slide = sub(mul(nrm, R), mul(q.slip, smin(q.g_slide * G, smax(0.0f, q.stop + dot(slipF, q.slip)))));
    const f3 zero = mk(0, 0, 0);
    f3 out = selp(pen >= gm.soft_depth && R > 0, add(zero, selp(stat, o_stat, o_slide)), zero);
    // hull_push_out: the normal force the push-out needs besides what the node has, applied when positive (its
    // force on its node is f, f x the weight 1 here)
    const float fn = q.t_push - dot(acc, nrm) * q.m;
    out = selp(cn.hull, mul(nrm, fn), out);
    const bool write = !cn.hull || fn > 0;
    f3 nw[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        f3 base = Fi[i];
#pragma unroll
        for (int j = 0; j < i; j++) base = selp(dup[i] == j, nw[j], base);
        const f3 d = mul(out, cn.w[i]);
        nw[i] = selp(write, selp(i < na, add(base, d), sub(base, d)), base);
        float* Fs = fs + 3 * root[i];
        Fs[0] = nw[i].x, Fs[1] = nw[i].y, Fs[2] = nw[i].z;
    }
}

// One workgroup per response group (consecutive pair islands; a grid-stride loop over them; groups share no node, so
// they run side by side and the result is the serial one). A group's hits are those of its runs in the three lists, in
// list order (within an island the reference's order; its islands' hits are independent); they go in windows of at
// most 512: the most consecutive narrow-phase blocks of a run (up to 256) whose hits in the run fit. In a window each
// hit's predecessors (the latest earlier hit of the window on each of its nodes) give its level, the longest chain to
// it; the levels run one after another, the hits of a level (no shared node) in parallel: in round r the hits whose
// predecessors all ran in earlier rounds. The window's node forces are held in LDS meanwhile, one copy per node at the
// node's first slot in the window (its root), read from F before the first round and written back after the last.
constexpr int kRespW = 512, kRespKeys = 4096, kRespBlk = 256, kRespH = 4096;
constexpr uint32_t kRespSpin = 1u << 22; // a wave's polls for its hits' predecessors before the response gives up
// a node's bucket among the window's kRespH (a multiplicative hash: consecutive nodes spread)
__device__ __forceinline__ uint32_t resp_bucket(uint32_t n) { return (n * 0x9E3779B1u) >> (32 - 12); }
static_assert(kRespH == 1 << 12 && kRespH % kRespW == 0, "resp_bucket takes the hash's top 12 bits");
static_assert(kRespDummy == 5 * kRespW && 2 * kRespKeys >= 3 * (kRespDummy + 1), "the dummy slot follows the window's slots");
__global__ __launch_bounds__(kRespW) void k_response(KArgs a, PArgs pin, int parity) {
    __shared__ uint64_t keys[kRespKeys];
    __shared__ uint32_t bstart[kRespH + 1];
    __shared__ uint16_t sroot[kRespW * 5]; // per slot: the first slot of its node in the window
    __shared__ int16_t spred[kRespW * 5];  // per slot: the latest earlier hit with a slot of its node, or -1
    __shared__ uint32_t dn[kRespW + 1];    // per hit: 1 once it has run; [kRespW]: always 1 (no predecessor)
    __shared__ uint32_t sred[kRespW / 64];
    __shared__ uint16_t wfirst[kRespBlk];
    __shared__ uint32_t wblk[kRespBlk];
    __shared__ uint32_t wpre[kRespBlk + 1], wtot[kRespBlk / 64], wnb;
    __shared__ uint4 wcut[8]; // the window's blocks that a run's start or end cuts (at most two a run): block, start, end, hits
    __shared__ uint2 wcutr[8];
    __shared__ uint32_t wcut_n;
    PArgs p = pin;
    p.nnt = pin.st->nnt, p.nnc = pin.st->nnc, p.nct = pin.st->nct;
    if (pin.st->nfnc) p.nc = pin.ncm; // (the node-capsule list merged with the fast bodies')
    const uint32_t nisl = pin.st->nisl;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t contacts = 0;
    for (uint32_t isl = blockIdx.x; isl < nisl; isl += gridDim.x) { // (a group)
        // its runs in the concatenated candidates (node-triangle, node-capsule, capsule-triangle) as one sequence of
        // (virtual) blocks: each run's blocks, a block shared by two runs once for each
        uint32_t slo[3], shi[3], sb0[3], sv0[3], nv = 0;
        for (int L = 0; L < 3; L++) {
            const uint32_t base = L == 0 ? 0u : (L == 1 ? p.nnt : p.nnt + p.nnc);
            slo[L] = base + p.irange[6 * isl + 2 * L], shi[L] = base + p.irange[6 * isl + 2 * L + 1];
            sb0[L] = slo[L] / 256, sv0[L] = nv;
            if (slo[L] < shi[L]) nv += (shi[L] - 1) / 256 - sb0[L] + 1;
        }
        {
            for (uint32_t wb = 0; wb < nv;) {
                // the next blocks' hits within [lo, hi) (hits are in candidate order in a block), their prefix, and how
                // many of the blocks fit one window (at least one)
                // (a block a run's start or end cuts - at most two a run - has its hits before the cut counted by a
                // wave, all of its hits' candidates read at once: one round of loads, not a binary search's chain)
                uint32_t cntv = 0, f = 0, e = 0, b = 0;
                int cut = -1;
                if (tid == 0) wcut_n = 0;
                __syncthreads();
                if (tid < (uint32_t)kRespBlk) {
                    const uint32_t v = wb + tid;
                    const int L = v < sv0[1] ? 0 : (v < sv0[2] ? 1 : 2);
                    const uint32_t lo = slo[L], hi = shi[L];
                    b = sb0[L] + (v - sv0[L]);
                    if (v < nv) {
                        e = p.bcnt[b];
                        if (b * 256 < lo || b * 256 + 256 > hi) {
                            cut = (int)atomicAdd(&wcut_n, 1u);
                            wcut[cut] = make_uint4(b, b * 256 < lo ? lo : 0u, b * 256 + 256 > hi ? hi : 0xffffffffu, e);
                        }
                    }
                }
                __syncthreads();
                for (uint32_t k = wv; k < wcut_n; k += kRespW / 64) {
                    const uint4 x = wcut[k];
                    uint32_t pr[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const uint32_t h = lane + 64 * j;
                        pr[j] = h < x.w ? p.hraw[(size_t)x.x * 256 + h].pair : 0xffffffffu;
                    }
                    uint32_t cl = 0, ch = 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        cl += (uint32_t)__builtin_popcountll(__ballot(pr[j] < x.y));
                        ch += (uint32_t)__builtin_popcountll(__ballot(pr[j] < x.z));
                    }
                    if (lane == 0) wcutr[k] = make_uint2(cl, ch);
                }
                __syncthreads();
                if (tid < (uint32_t)kRespBlk) {
                    if (cut >= 0) f = wcutr[cut].x, e = wcutr[cut].y;
                    wfirst[tid] = (uint16_t)f, wblk[tid] = b, cntv = e - f;
                }
                uint32_t x = cntv;
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t y = __shfl_up(x, off);
                    if (lane >= (uint32_t)off) x += y;
                }
                if (tid < (uint32_t)kRespBlk && lane == 63) wtot[wv] = x;
                if (tid == 0) wpre[0] = 0, wnb = 0;
                __syncthreads();
                if (tid < (uint32_t)kRespBlk) {
                    for (uint32_t w = 0; w < wv; w++) x += wtot[w];
                    wpre[tid + 1] = x;
                }
                __syncthreads();
                if (tid < (uint32_t)kRespBlk && wb + tid < nv && (tid == 0 || wpre[tid + 1] <= (uint32_t)kRespW))
                    atomicMax(&wnb, tid + 1);
                __syncthreads();
                const uint32_t nbw = wnb, m = wpre[nbw];
                const bool valid = tid < m;
                Hit hit;
                ContactNodes cn;
                cn.na = cn.nb = 0;
                if (valid) {
                    // the block holding the window's hit tid: the last one whose prefix is at most tid
                    uint32_t l2 = 0, h2 = nbw - 1;
                    while (l2 < h2) {
                        const uint32_t mid = (l2 + h2 + 1) >> 1;
                        if (wpre[mid] <= tid) l2 = mid;
                        else h2 = mid - 1;
                    }
                    hit = p.hraw[(size_t)wblk[l2] * 256 + wfirst[l2] + (tid - wpre[l2])];
                    contact_nodes(hit, cn);
                }
                if (m > 0) {
                    // the window's slots hashed by node into buckets (count, scan, scatter): a slot finds the other
                    // slots of its node in its bucket
                    const int nn = cn.na + cn.nb;
                    // the nodes' inverse masses, velocities and forces (read now, used after the bucketing)
                    float im[5];
                    f3 vel[5], f0[5];
                    {
                        const float4* P = a.pm[parity];
                        const float4* Vv = a.vm[parity];
#pragma unroll
                        for (int j = 0; j < 5; j++) {
                            im[j] = 0, vel[j] = mk(0, 0, 0), f0[j] = mk(0, 0, 0);
                            if (j < nn) {
                                const uint32_t n = cn.n[j];
                                im[j] = P[n].w, vel[j] = ld3(Vv[n]), f0[j] = ld3(a.F[n]);
                            }
                        }
                    }
                    for (uint32_t k = tid; k < (uint32_t)kRespH; k += kRespW) bstart[k] = 0;
                    __syncthreads();
                    uint32_t pos[5];
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        if (j < nn) pos[j] = atomicAdd(&bstart[resp_bucket(cn.n[j])], 1u);
                    __syncthreads();
                    {
                        constexpr int per = kRespH / kRespW;
                        uint32_t v[per], tot = 0;
#pragma unroll
                        for (int i = 0; i < per; i++) v[i] = bstart[tid * per + i], tot += v[i];
                        uint32_t x = tot;
                        for (int off = 1; off < 64; off <<= 1) {
                            const uint32_t y = __shfl_up(x, off);
                            if (lane >= (uint32_t)off) x += y;
                        }
                        if (lane == 63) sred[wv] = x;
                        __syncthreads();
                        uint32_t run = x - tot;
                        for (uint32_t w = 0; w < wv; w++) run += sred[w];
#pragma unroll
                        for (int i = 0; i < per; i++) bstart[tid * per + i] = run, run += v[i];
                        if (tid == kRespW - 1) bstart[kRespH] = run;
                    }
                    __syncthreads();
                    uint64_t* ent = keys;
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        if (j < nn) ent[bstart[resp_bucket(cn.n[j])] + pos[j]] = ((uint64_t)cn.n[j] << 32) | (tid * 5 + j);
                    __syncthreads();
                    // each bucket record's root (the least slot of its node in the bucket) and predecessor (the
                    // greatest earlier hit with a slot of the node), a thread per record, into LDS by slot
                    const uint32_t nent = bstart[kRespH];
                    for (uint32_t k = tid; k < nent; k += kRespW) {
                        const uint64_t me = ent[k];
                        const uint32_t node = (uint32_t)(me >> 32), t0 = (uint32_t)me, h0 = t0 / 5;
                        const uint32_t b = resp_bucket(node), lo = bstart[b], hi = bstart[b + 1];
                        uint32_t root = t0;
                        int pq = -1;
                        for (uint32_t i = lo; i < hi; i += 4) {
                            uint64_t e[4];
#pragma unroll
                            for (int u = 0; u < 4; u++) e[u] = i + u < hi ? ent[i + u] : ~0ull;
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                if ((uint32_t)(e[u] >> 32) != node) continue;
                                const uint32_t t = (uint32_t)e[u], h = t / 5;
                                root = min(root, t);
                                if (h < h0) pq = max(pq, (int)h);
                            }
                        }
                        sroot[t0] = (uint16_t)root, spred[t0] = (int16_t)pq;
                    }
                    __syncthreads();
                    // (a slot the contact does not use: the dummy cell, no predecessor - the always-set flag)
                    uint16_t rt[5], px[5];
                    int dup[5];
#pragma unroll
                    for (int j = 0; j < 5; j++) {
                        rt[j] = (uint16_t)kRespDummy, px[j] = (uint16_t)kRespW, dup[j] = -1;
                        if (j < nn) {
                            const int pq = spred[tid * 5 + j];
                            rt[j] = sroot[tid * 5 + j], px[j] = (uint16_t)(pq >= 0 ? pq : kRespW);
                        }
#pragma unroll
                        for (int i = 0; i < j; i++)
                            if (j < nn && rt[i] == rt[j]) dup[j] = i;
                    }
                    // the nodes' forces into LDS at their roots (the bucket records are read by now), the
                    // velocity-only terms of each hit
                    __syncthreads();
                    float* fs = (float*)keys;
                    CfPre q;
                    q.ok = false;
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        if (j < nn && rt[j] == tid * 5 + j)
                            fs[3 * rt[j]] = f0[j].x, fs[3 * rt[j] + 1] = f0[j].y, fs[3 * rt[j] + 2] = f0[j].z;
                    if (valid) contact_force_pre(cn, im, vel, hit.nrm, hit.pen, a.dt, q);
                    // dataflow: a hit runs as soon as every predecessor has run (a flag per hit in LDS, written
                    // with release and read with acquire at workgroup scope), so a chain's link costs one contact's
                    // own latency, not a barrier round. Each wave loops until all its hits have run: a hit's
                    // predecessors are earlier hits, so the earliest hit not run is always ready. The polls are
                    // bounded all the same: a wait that outlasts kRespSpin passes raises the error word and ends.
                    if (tid == 0) dn[kRespW] = 1u;
                    dn[tid] = valid ? 0u : 1u;
                    bool done = !valid;
                    __syncthreads();
                    for (uint32_t spins = 0; !__all(done); spins++) {
                        if (!done) {
                            uint32_t f[5];
#pragma unroll
                            for (int j = 0; j < 5; j++)
                                f[j] = __hip_atomic_load(&dn[px[j]], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (f[0] & f[1] & f[2] & f[3] & f[4]) {
                                if (q.ok) contact_force_apply(cn, q, rt, dup, fs, hit.nrm, hit.pen), contacts++;
                                __hip_atomic_store(&dn[tid], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                                done = true;
                            } else if (spins >= kRespSpin) {
                                if (p.err) atomicOr(p.err, 2);
                                done = true;
                            }
                        }
                    }
                    __syncthreads();
                    // the forces back to F
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        if (j < nn && rt[j] == tid * 5 + j) {
                            float* F = (float*)(a.F + cn.n[j]);
                            F[0] = fs[3 * rt[j]], F[1] = fs[3 * rt[j] + 1], F[2] = fs[3 * rt[j] + 2];
                        }
                }
                __syncthreads();
                wb += nbw;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) contacts += __shfl_xor(contacts, off);
    if (lane == 0 && contacts) atomicAdd(&a.counters[CNT_CONTACTS], (unsigned long long)contacts);
}

// ---------------------------------------------------------------------------------------------- exclusive scan
// One launch, one pass (the look-back of the one-pass compaction), 2 048 values a workgroup.
constexpr uint32_t kScanPer = 8, kScanTile = 256 * kScanPer;

__global__ __launch_bounds__(256) void k_scan(const uint32_t* in, uint32_t* out, uint32_t n, Lookback lk, RStat* st) {
    __shared__ uint32_t sh[8];
    const uint32_t t = wg_ticket(lk, sh);
    const size_t i0 = (size_t)t * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer], run = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++) {
        const uint32_t x = i0 + k < n ? in[i0 + k] : 0u;
        v[k] = run;
        run += x;
    }
    uint32_t agg;
    const uint32_t bp = wg_prefix(run, agg, sh + 1);
    const uint32_t off = wg_lookback(lk, t, agg, st, sh + 5) + bp;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++)
        if (i0 + k < n) out[i0 + k] = off + v[k];
}

// ---------------------------------------------------------------------------------------------- host side
namespace {

#define CHK(x)                                                                                                      \
    do {                                                                                                            \
        if ((x) != hipSuccess) return BL_ERR_HIP;                                                                   \
    } while (0)
#define RC(x)                                                                                                       \
    do {                                                                                                            \
        const int r_ = (x);                                                                                         \
        if (r_) return r_;                                                                                          \
    } while (0)

inline dim3 grid_for(size_t n, int block = 256) { return dim3((unsigned)((n + block - 1) / block)); }

inline uint32_t pow2_at_least(size_t n) {
    uint32_t v = 64;
    while (v < n) v <<= 1;
    return v;
}

} // namespace

// The look-back words and tickets of a one-pass launch of nb workgroups.
int ContactEngine::lookback(Lookback& k, size_t nb, hipStream_t st) {
    if (!lk_words.p || nb > lk_words.cap || lk_epoch_ >= 0x3fffffffu) {
        // fresh words (and after the epoch wraps): cleared, so that no stale word carries the current epoch
        RC(lk_words.ensure(nb));
        RC(lk_ticket.ensure(1));
        hipLaunchKernelGGL(k_rb_clear, grid_for(lk_words.cap * 2), dim3(256), 0, st, (uint32_t*)lk_words.p,
                           (uint32_t)(lk_words.cap * 2));
        hipLaunchKernelGGL(k_rb_clear, dim3(1), dim3(64), 0, st, lk_ticket.p, 1u);
        lk_epoch_ = 0, lk_base_ = 0;
    }
    k.words = lk_words.p, k.ticket = lk_tickets_ ? lk_ticket.p : nullptr, k.base = lk_base_, k.epoch = ++lk_epoch_;
    if (lk_tickets_) lk_base_ += (uint32_t)nb;
    return BL_OK;
}

int ContactEngine::scan(const uint32_t* in, uint32_t* out, size_t n, hipStream_t st) {
    if (n == 0) return BL_OK;
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    Lookback k;
    RC(lookback(k, nb, st));
    hipLaunchKernelGGL(k_scan, dim3((unsigned)nb), dim3(256), 0, st, in, out, (uint32_t)n, k, stat.p);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

void ContactEngine::release() {
    up_valid_ = false;
    lk_words.release(), lk_ticket.release();
    if (hstat) (void)hipHostFree(hstat);
    hstat = nullptr;
    lk_epoch_ = 0, lk_base_ = 0;
    bisl.release(), bmem.release(), pcnt.release(), poff.release(), plist.release();
    nmem = cap_part = 0;
    for (Buf<uint32_t>* b : {&imem_off, &imem, &bquery, &body_tri0, &sample_base, &icount, &pbucket, &tcount, &tstart,
                             &eb, &ecount, &estart, &bcnt, &icap, &irange})
        b->release();
    igroup.release(), ngroup_ = 0;
    bbox.release(), samples.release(), extra.release(), ipar.release(), stat.release(), pcell.release(), ppos.release();
    itri.release(), n_itri = 0, n_icap = 0, gresp_ = 0;
    scell.release(), spos.release(), ea.release(), ebb.release(), ec.release(), sa.release(), sb.release(), sc.release();
    for (Buf<uint4>* b : {&ntraw, &nc, &ct, &nt}) b->release();
    hraw.release();
    nnt = nnc = nct = 0;
    cap_nt = cap_nc = cap_ct = cap_ent = esize = 0;
    cap_fq = cap_fnt = cap_fnc = 0;
    for (Buf<uint4>* b : {&fnt, &fnc, &ntm, &ncm}) b->release();
    body_cap0.release(), fqoff.release(), fgrid.release();
}

// The buffers whose sizes follow the candidate counts (grown when a rebuild reports an overflow).
int ContactEngine::ensure_lists() {
    if (!cap_fq) cap_fq = 256, cap_fnt = 1024, cap_fnc = 1024;
    RC(ntraw.ensure(cap_nt));
    RC(nt.ensure((size_t)cap_nt + cap_fnt));
    RC(nc.ensure(cap_nc));
    RC(fnt.ensure(cap_fnt));
    RC(fnc.ensure(cap_fnc));
    RC(ntm.ensure((size_t)cap_nt + cap_fnt));
    RC(ncm.ensure((size_t)cap_nc + cap_fnc));
    RC(ct.ensure(cap_ct));
    RC(eb.ensure(cap_ent));
    RC(ea.ensure(cap_ent));
    RC(ebb.ensure(cap_ent));
    RC(ec.ensure(cap_ent));
    RC(sa.ensure(cap_ent));
    RC(sb.ensure(cap_ent));
    RC(sc.ensure(cap_ent));
    esize = pow2_at_least(2 * (size_t)cap_ent);
    RC(ecount.ensure((size_t)esize + 1));
    RC(estart.ensure((size_t)esize + 1));
    counts_dirty_ = true;
    return BL_OK;
}

// The per-body tables (first triangle, first size sample), the node hash and the first list sizes.
int ContactEngine::setup(const HostView& hv, uint32_t N, uint32_t T, uint32_t C) {
    up_valid_ = false;
    nnodes = N, ntris = T, ncaps = C, nbodies = (uint32_t)hv.bodies.size();
    std::vector<uint32_t> tri0(nbodies + 1, 0), s0(nbodies + 1, 0);
    for (uint32_t i = 0; i < nbodies; i++) {
        tri0[i + 1] = tri0[i] + hv.bodies[i].ntri;
        s0[i + 1] = s0[i] + (hv.bodies[i].ntri + 6) / 7;
    }
    nsamples = s0[nbodies];
    h_tri0_ = tri0;
    h_cap0_.assign(nbodies + 1, 0);
    for (uint32_t i = 0; i < nbodies; i++) h_cap0_[i + 1] = h_cap0_[i] + hv.bodies[i].ncap;
    h_nnode_.resize(nbodies);
    for (uint32_t i = 0; i < nbodies; i++) h_nnode_[i] = hv.bodies[i].nnode;
    RC(body_tri0.ensure(nbodies + 1));
    RC(sample_base.ensure(nbodies + 1));
    RC(samples.ensure(nsamples + 1));
    CHK(hipMemcpy(body_tri0.p, tri0.data(), 4 * (nbodies + 1), hipMemcpyHostToDevice));
    RC(body_cap0.ensure(nbodies + 1));
    CHK(hipMemcpy(body_cap0.p, h_cap0_.data(), 4 * (nbodies + 1), hipMemcpyHostToDevice));
    CHK(hipMemcpy(sample_base.p, s0.data(), 4 * (nbodies + 1), hipMemcpyHostToDevice));
    RC(bisl.ensure(nbodies + 1));
    RC(bmem.ensure(nbodies + 1));
    RC(bbox.ensure(6 * (size_t)nbodies + 6));
    RC(extra.ensure((size_t)nbodies + 1));
    RC(bquery.ensure(nbodies + 1));
    RC(icount.ensure(nbodies + 1));
    RC(ipar.ensure(nbodies + 1));
    RC(stat.ensure(1));
    RC(pbucket.ensure(N + 1));
    RC(pcell.ensure(N + 1));
    RC(ppos.ensure(N + 1));
    RC(scell.ensure(N + 1));
    RC(spos.ensure(N + 1));
    tsize = pow2_at_least(4 * (size_t)N);
    RC(tcount.ensure((size_t)tsize + 1));
    counts_dirty_ = true;
    RC(tstart.ensure((size_t)tsize + 1));
    // (first guesses: a list that outgrows its buffer grows it, and the rebuild runs again once; the filter's
    // launch covers the node-triangle buffer, so a large guess costs every rebuild)
    cap_nt = std::max<uint32_t>(4096, T / 8);
    cap_nc = std::max<uint32_t>(4096, 4 * C);
    cap_ct = std::max<uint32_t>(4096, 4 * C);
    cap_ent = std::max<uint32_t>(4096, 8 * T);
    RC(ensure_lists());
    nnt = nnc = nct = 0;
    return BL_OK;
}

// World::build_islands (src/phys/world.cpp:636-762) for awake node/beam bodies: union-find over the expanded bounds
// of bodies that can touch (contacters against triangles or capsules, capsules against triangles); an island's
// bodies in body order, islands by their first body. Islands of several bodies, or of one self-colliding body,
// search for pairs. The candidate pairs of bodies come from the 3-D grid of bl_islands.h (the union order does not
// change the islands). The scratch vectors live across frames.
void ContactEngine::build_islands(const HostView& hv, float frame_time) {
    const size_t nb = hv.bodies.size();
    std::vector<int>& parent = parent_;
    parent.resize(nb);
    for (size_t i = 0; i < nb; i++) parent[i] = (int)i;
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::vector<float>& bx = bx_;
    bx.resize(nb * 6);
    for (size_t i = 0; i < nb; i++) {
        const float m = smin(hv.max_speed[i], 200.0f) * frame_time * 2.0f + 0.3f;
        const float r = m + hv.bodies[i].hull_reach;
        const float mn[3] = {hv.mn[i].x, hv.mn[i].y, hv.mn[i].z}, mx[3] = {hv.mx[i].x, hv.mx[i].y, hv.mx[i].z};
        for (int k = 0; k < 3; k++) bx[6 * i + k] = mn[k] - r, bx[6 * i + 3 + k] = mx[k] + r;
    }
    // (i < j, their boxes overlapping)
    auto consider = [&](size_t i, size_t j) {
        const BodyView &bi = hv.bodies[i], &bj = hv.bodies[j];
        if (bi.group != 0 && bi.group == bj.group) return;
        const bool ti = bi.ntri > 0, tj = bj.ntri > 0, ci = bi.ncap > 0, cj = bj.ncap > 0;
        const bool can = (bi.contacters && (tj || cj)) || (bj.contacters && (ti || ci)) || (ci && tj) || (cj && ti);
        if (!can) return;
        const int x = find((int)i), y = find((int)j);
        if (x != y) parent[x] = y;
    };
    if (hv.inter_body) overlapping_pairs(bx.data(), nb, pair_ws_, [&](uint32_t i, uint32_t j) { consider(i, j); });
    // islands numbered by their first body; the pair islands' members in body order
    root_isl_.assign(nb, -1);
    body_isl_.resize(nb);
    isl_size_.clear(), isl_first_.clear();
    for (size_t i = 0; i < nb; i++) {
        const int r = find((int)i);
        if (root_isl_[r] < 0) root_isl_[r] = (int)isl_size_.size(), isl_size_.push_back(0), isl_first_.push_back((int)i);
        body_isl_[i] = root_isl_[r];
        isl_size_[root_isl_[r]]++;
    }
    const size_t ni = isl_size_.size();
    pair_id_.assign(ni, -1);
    uint32_t np = 0;
    for (size_t k = 0; k < ni; k++)
        if (hv.inter_body && (isl_size_[k] > 1 || hv.bodies[isl_first_[k]].self)) pair_id_[k] = (int)np++;
    pair_off.assign(np + 1, 0);
    for (size_t k = 0; k < ni; k++)
        if (pair_id_[k] >= 0) pair_off[pair_id_[k] + 1] = isl_size_[k];
    for (uint32_t k = 0; k < np; k++) pair_off[k + 1] += pair_off[k];
    pair_mem.resize(pair_off[np]);
    cursor_.assign(pair_off.begin(), pair_off.end() - 1);
    for (size_t i = 0; i < nb; i++) {
        const int p = pair_id_[body_isl_[i]];
        if (p >= 0) pair_mem[cursor_[p]++] = (uint32_t)i;
    }
}

int ContactEngine::begin_frame(hipStream_t st) {
    const size_t np = npair_islands();
    const std::vector<uint32_t>& off = pair_off;
    const std::vector<uint32_t>& mem = pair_mem;
    std::vector<int>& isl = isl_;
    std::vector<int>& mi = root_isl_; // (the island building is done with it)
    isl.assign(nbodies, -1), mi.assign(nbodies, -1);
    for (size_t k = 0; k < np; k++)
        for (uint32_t p = off[k]; p < off[k + 1]; p++) isl[mem[p]] = (int)k, mi[mem[p]] = (int)p;
    nmem = (uint32_t)mem.size();
    // the triangles and capsules the searches run over: the pair islands' bodies', island by island, each island's
    // in body order. The candidate lists follow this order, so each island's candidates are one run in each list (an
    // island's own order is the reference's; islands share no node, so their relative order changes nothing); no
    // other triangle or capsule has candidates. And the node hash's size (only when the island bodies changed).
    if (!up_valid_ || isl != up_isl_) {
        itri_h_.clear(), icap_h_.clear();
        size_t inodes = 0;
        for (size_t k = 0; k < np; k++)
            for (uint32_t q = off[k]; q < off[k + 1]; q++) {
                const uint32_t g = mem[q];
                for (uint32_t t = h_tri0_[g]; t < h_tri0_[g + 1]; t++) itri_h_.push_back(t);
                for (uint32_t t = h_cap0_[g]; t < h_cap0_[g + 1]; t++) icap_h_.push_back(t);
                inodes += h_nnode_[g];
            }
        n_itri = (uint32_t)itri_h_.size(), n_icap = (uint32_t)icap_h_.size();
        tsize_f = std::min<uint32_t>(tsize, pow2_at_least(4 * inodes));
    }
    const void* was[4] = {imem_off.p, imem.p, itri.p, icap.p};
    RC(imem_off.ensure(np + 1));
    RC(imem.ensure(mem.size() + 1));
    RC(fqoff.ensure(mem.size() + 1));
    RC(fgrid.ensure(mem.size() + 1));
    RC(itri.ensure((size_t)n_itri + 1));
    RC(icap.ensure((size_t)n_icap + 1));
    // the response's groups: consecutive pair islands balanced by size (nodes and triangles), at most kRespGroups, so
    // that a group's hits are one run in each list; a workgroup per group, in powers of two (a captured segment keeps
    // its launch until the count outgrows it)
    constexpr uint32_t kRespGroups = 256;
    {
        igroup_h_.assign(np + 1, 0);
        double total = 0;
        std::vector<double> wk(np, 0.0);
        for (size_t k = 0; k < np; k++) {
            for (uint32_t q = off[k]; q < off[k + 1]; q++) {
                const uint32_t g = mem[q];
                wk[k] += (double)h_nnode_[g] + (double)(h_tri0_[g + 1] - h_tri0_[g]);
            }
            total += wk[k];
        }
        const uint32_t G = (uint32_t)std::min<size_t>(np, kRespGroups);
        uint32_t grp = 0;
        double acc = 0;
        for (size_t k = 0; k < np; k++) {
            igroup_h_[k] = (int)grp;
            acc += wk[k];
            if (grp + 1 < G && acc >= total * (double)(grp + 1) / (double)G) grp++;
        }
        ngroup_ = np ? std::min<uint32_t>(grp + 1, G) : 0;
        if (np && igroup_h_[np - 1] + 1 < (int)ngroup_) ngroup_ = (uint32_t)igroup_h_[np - 1] + 1;
    }
    if (ngroup_ > gresp_) gresp_ = pow2_at_least(ngroup_);
    RC(irange.ensure(6 * (size_t)std::max<uint32_t>(gresp_, 1)));
    const void* was_g = igroup.p;
    RC(igroup.ensure(np + 1));
    RC(pcnt.ensure(nmem + 1));
    RC(poff.ensure(nmem + 1));
    cap_part = std::max<uint32_t>(cap_part, std::max<uint32_t>(4096, 8 * nmem));
    RC(plist.ensure(cap_part));
    // (the same tables as the last frame's, in the same buffers: nothing to upload)
    const bool same = up_valid_ && was[0] == imem_off.p && was[1] == imem.p && was[2] == itri.p && was[3] == icap.p &&
                      was_g == igroup.p && igroup_h_ == up_group_ && isl == up_isl_ &&
                      mi == up_mi_ && off == up_off_ && mem == up_mem_;
    if (!same) {
        const std::vector<uint32_t> off0(1, 0u);
        const std::vector<uint32_t>& offu = np ? off : off0;
        if (nbodies) CHK(hipMemcpyAsync(bisl.p, isl.data(), sizeof(int) * nbodies, hipMemcpyHostToDevice, st));
        if (nbodies) CHK(hipMemcpyAsync(bmem.p, mi.data(), sizeof(int) * nbodies, hipMemcpyHostToDevice, st));
        CHK(hipMemcpyAsync(imem_off.p, offu.data(), 4 * offu.size(), hipMemcpyHostToDevice, st));
        if (!mem.empty()) CHK(hipMemcpyAsync(imem.p, mem.data(), 4 * mem.size(), hipMemcpyHostToDevice, st));
        if (n_itri) CHK(hipMemcpyAsync(itri.p, itri_h_.data(), 4 * (size_t)n_itri, hipMemcpyHostToDevice, st));
        if (n_icap) CHK(hipMemcpyAsync(icap.p, icap_h_.data(), 4 * (size_t)n_icap, hipMemcpyHostToDevice, st));
        if (np) CHK(hipMemcpyAsync(igroup.p, igroup_h_.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
        CHK(hipStreamSynchronize(st)); // (off0 goes out of scope)
        up_isl_ = isl, up_mi_ = mi, up_off_ = off, up_mem_ = mem, up_group_ = igroup_h_;
        up_valid_ = true;
    }
    nnt = nnc = nct = 0;
    return BL_OK;
}

// The rebuild's launches: no host wait between them; the counts stay on the device.
int ContactEngine::enqueue_rebuild(const KArgs& a, CArgs& c, hipStream_t st) {
    const uint32_t N = nnodes, T = ntris, C = ncaps, NB = nbodies;
    c.cap_nt = cap_nt, c.cap_nc = cap_nc, c.cap_ct = cap_ct, c.cap_ent = cap_ent, c.cap_part = cap_part;
    c.cap_fq = cap_fq, c.cap_fnt = cap_fnt, c.cap_fnc = cap_fnc;
    c.fnt = fnt.p, c.fnc = fnc.p, c.ntm = ntm.p, c.ncm = ncm.p;
    c.plist = plist.p;
    c.ntraw = ntraw.p, c.nc = nc.p, c.ct = ct.p, c.nt = nt.p;
    c.eb = eb.p, c.ea = ea.p, c.ebb = ebb.p, c.ec = ec.p, c.sa = sa.p, c.sb = sb.p, c.sc = sc.p;
    c.ecount = ecount.p, c.estart = estart.p;
    c.emask = esize - 1;
    // the one-pass stages: a launch of nb workgroups each, with its own look-back epoch
    auto stage = [&](auto kern, size_t nb) -> int {
        if (nb == 0) return BL_OK;
        RC(lookback(c.lk, nb, st));
        hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, st, a, c);
        return BL_OK;
    };
    hipLaunchKernelGGL(k_rb_clear, dim3(1), dim3(64), 0, st, (uint32_t*)stat.p, (uint32_t)(sizeof(RStat) / 4));
    // the bucket counts are counted down to zero by every rebuild that runs to its end; cleared after a fresh
    // allocation or a rebuild that did not
    if (counts_dirty_) {
        hipLaunchKernelGGL(k_rb_clear, grid_for((size_t)tsize + 1), dim3(256), 0, st, tcount.p, tsize + 1);
        if (esize) hipLaunchKernelGGL(k_rb_clear, grid_for((size_t)esize + 1), dim3(256), 0, st, ecount.p, esize + 1);
    }
    counts_dirty_ = true;
    if (T) hipLaunchKernelGGL(k_rb_torn, grid_for(T), dim3(256), 0, st, a, c);
    hipLaunchKernelGGL(k_rb_island, dim3(c.nisl), dim3(256), 0, st, a, c);
    hipLaunchKernelGGL(k_rb_boxes, grid_for(NB), dim3(256), 0, st, a, c);
    RC(stage(k_rb_partners, ((size_t)nmem + 3) / 4));
    // 1) the node hash
    RC(stage(k_rb_pts, ((size_t)N + 255) / 256));
    RC(scan(tcount.p, tstart.p, (size_t)tsize_f + 1, st));
    hipLaunchKernelGGL(k_rb_fill, grid_for(N), dim3(256), 0, st, a, c);
    // 2) triangles, then capsules, against the node hash
    if (n_itri) RC(stage(k_rb_nt, ((size_t)n_itri * kGnt + 255) / 256));
    if (n_icap) RC(stage(k_rb_nc, ((size_t)n_icap * kGcap + 255) / 256));
    // 2b) the fast bodies' nodes against their partners' triangles and capsules (their grids, queries, candidates),
    // then the lists merged island by island where there are any
    hipLaunchKernelGGL(k_rb_fgrid, dim3(1), dim3(1024), 0, st, a, c);
    RC(stage(k_rb_fast<false>, ((size_t)cap_fq * kGf + 255) / 256));
    RC(stage(k_rb_fast<true>, ((size_t)cap_fq * kGf + 255) / 256));
    hipLaunchKernelGGL(k_rb_splice, grid_for((size_t)cap_nt + cap_fnt + cap_nc + cap_fnc), dim3(256), 0, st, a, c);
    // 3) capsules against the triangles' cells
    if (C && T) {
        RC(stage(k_rb_prims, ((size_t)n_itri + 255) / 256));
        RC(scan(ecount.p, estart.p, (size_t)esize + 1, st));
        hipLaunchKernelGGL(k_rb_efill, grid_for(cap_ent), dim3(256), 0, st, c);
        if (n_icap) RC(stage(k_rb_ct, ((size_t)n_icap * kGcap + 255) / 256));
    }
    // the near-pair filter on the node-triangle candidates
    if (T) RC(stage(k_rb_near, ((size_t)cap_nt + cap_fnt + 255) / 256));
    // each island's runs in the three lists
    hipLaunchKernelGGL(k_rb_clear, grid_for(6 * (size_t)c.ngroup), dim3(256), 0, st, c.irange, 6 * c.ngroup);
    const size_t nr = std::max<size_t>(1, std::max<size_t>((size_t)cap_nt + cap_fnt, std::max<size_t>((size_t)cap_nc + cap_fnc, cap_ct)));
    hipLaunchKernelGGL(k_rb_ranges, grid_for(nr), dim3(256), 0, st, a, c);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

// The frame's pair lists anew (its first substep and every `interval` after), all pair islands at once. One wait
// for the device at the end; a list that outgrew its buffer grows it and the rebuild runs again (the rebuild only
// reads the state, so a second run finds the same lists).
int ContactEngine::rebuild(const KArgs& a, const uint32_t* body_box, hipStream_t st, int parity) {
    nnt = nnc = nct = 0;
    if (npair_islands() == 0) return BL_OK;
    CArgs c{};
    c.parity = parity;
    c.horizon = (float)(interval > 1 ? interval : 1) * a.dt;
    c.interval = interval, c.dt = a.dt;
    c.bisl = bisl.p, c.imem_off = imem_off.p, c.imem = imem.p, c.nisl = (uint32_t)npair_islands();
    c.nmem = nmem, c.bmem = bmem.p, c.pcnt = pcnt.p, c.poff = poff.p;
    c.ipar = ipar.p, c.bbox = bbox.p, c.extra = extra.p, c.bquery = bquery.p, c.body_box = body_box;
    c.body_tri0 = body_tri0.p, c.body_cap0 = body_cap0.p, c.sample_base = sample_base.p, c.samples = samples.p;
    c.fgrid = fgrid.p, c.fqoff = fqoff.p;
    c.nsamples = nsamples;
    c.itri = itri.p, c.n_itri = n_itri, c.icap = icap.p, c.n_icap = n_icap, c.irange = irange.p;
    c.igroup = igroup.p, c.ngroup = ngroup_;
    c.st = stat.p, c.icount = icount.p;
    c.pbucket = pbucket.p, c.pcell = pcell.p, c.ppos = ppos.p;
    c.tmask = tsize_f - 1, c.tcount = tcount.p, c.tstart = tstart.p;
    c.scell = scell.p, c.spos = spos.p;
    for (int attempt = 0;; attempt++) {
        RC(enqueue_rebuild(a, c, st));
        if (!hstat && hipHostMalloc((void**)&hstat, sizeof(RStat), hipHostMallocDefault) != hipSuccess) {
            hstat = nullptr;
            return BL_ERR_HIP;
        }
        CHK(hipMemcpyAsync(hstat, stat.p, sizeof(RStat), hipMemcpyDeviceToHost, st));
        CHK(hipStreamSynchronize(st));
        last = *hstat;
        // (a rebuild that timed out or overflowed may not have counted the buckets back down: cleared next time)
        counts_dirty_ = (last.unsupported & 4) != 0 || last.overflow != 0;
        if (last.unsupported & 8) {
            blg::set_last_error("internal: a pair island's candidates are not one run in a list");
            counts_dirty_ = true;
            return BL_ERR_STATE;
        }
        if ((last.unsupported & 4) && !lk_tickets_) {
            // a look-back wait timed out with workgroups ordered by index: tickets from now on, and again
            lk_tickets_ = true;
            counts_dirty_ = true;
            continue;
        }
        if (!last.overflow) break;
        if (attempt > 4) return BL_ERR_HIP;
        if (last.overflow & 1) cap_nt = std::max<uint32_t>(cap_nt, last.nnt_raw + last.nnt_raw / 2 + 1);
        if (last.overflow & 2) cap_nc = std::max<uint32_t>(cap_nc, last.nnc_slow + last.nnc_slow / 2 + 1);
        if (last.overflow & 32) cap_fq = std::max<uint32_t>(cap_fq, last.nfq + last.nfq / 2 + 1);
        if (last.overflow & 64) cap_fnt = std::max<uint32_t>(cap_fnt, last.nfnt + last.nfnt / 2 + 1);
        if (last.overflow & 128) cap_fnc = std::max<uint32_t>(cap_fnc, last.nfnc + last.nfnc / 2 + 1);
        if (last.overflow & 4) cap_ct = std::max<uint32_t>(cap_ct, last.nct + last.nct / 2 + 1);
        if (last.overflow & 8) cap_ent = std::max<uint32_t>(cap_ent, last.nent + last.nent / 2 + 1);
        if (last.overflow & 16) {
            cap_part = std::max<uint32_t>(cap_part, last.npart + last.npart / 2 + 1);
            RC(plist.ensure(cap_part));
        }
        RC(ensure_lists());
    }
    nnt = last.nnt, nnc = last.nnc, nct = last.nct;
    pair_rebuilds += (int64_t)npair_islands();
    contact_pairs = last.last_pairs;
    return BL_OK;
}

// The substep's contacts: the narrow phase over the frame's candidates, then the response in candidate order.
int ContactEngine::narrow_args(PArgs& p, uint32_t& nblk) {
    const size_t n = nnt + nnc + nct;
    nblk = (uint32_t)((n + 255) / 256);
    p = PArgs{};
    if (n == 0) return BL_OK;
    RC(hraw.ensure((size_t)nblk * 256));
    RC(bcnt.ensure(nblk));
    p.nt = nt.p, p.nc = nc.p, p.ct = ct.p, p.ncm = ncm.p;
    p.nnt = (uint32_t)nnt, p.nnc = (uint32_t)nnc, p.nct = (uint32_t)nct, p.st = stat.p;
    p.hraw = hraw.p, p.bcnt = bcnt.p, p.irange = irange.p, p.gresp = std::max<uint32_t>(gresp_, 1);
    p.err = err_word;
    return BL_OK;
}

int ContactEngine::collide(const KArgs& a, hipStream_t st, int parity) {
    count_tests(1);
    PArgs p;
    uint32_t nblk;
    RC(narrow_args(p, nblk));
    if (nblk == 0) return BL_OK;
    return launch_collide(a, p, nblk, st, parity);
}

int ContactEngine::launch_response(const KArgs& a, const PArgs& p, uint32_t nblk, hipStream_t st, int parity) {
    (void)nblk;
    hipLaunchKernelGGL(k_response, dim3(p.gresp), dim3(kRespW), 0, st, a, p, parity);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

int ContactEngine::collide_fixed(PArgs& p, uint32_t& nblk) {
    const size_t cap = (size_t)cap_nt + cap_fnt + cap_nc + cap_fnc + cap_ct;
    nblk = (uint32_t)((cap + 255) / 256);
    if (nblk == 0) nblk = 1;
    RC(hraw.ensure((size_t)nblk * 256));
    RC(bcnt.ensure(nblk));
    p = PArgs{};
    p.nt = nt.p, p.nc = nc.p, p.ct = ct.p, p.ncm = ncm.p, p.st = stat.p;
    p.hraw = hraw.p, p.bcnt = bcnt.p, p.irange = irange.p, p.gresp = std::max<uint32_t>(gresp_, 1);
    p.err = err_word;
    return BL_OK;
}

int ContactEngine::launch_collide(const KArgs& a, const PArgs& p, uint32_t nblk, hipStream_t st, int parity) {
    hipLaunchKernelGGL(k_narrow, dim3(nblk), dim3(256), 0, st, a, p, parity);
    hipLaunchKernelGGL(k_response, dim3(p.gresp), dim3(kRespW), 0, st, a, p, parity);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

} // namespace blg
