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
// built in the reference's order - candidates by triangle (body, then index), then hash cell (z, y, x), then
// bucket position; a bucket holds its nodes by insertion (body, then node): a stable radix sort of (bucket,
// insertion index) keys gives the same buckets; all compactions are order-preserving scans. The response runs in
// one workgroup: contacts that share no node are independent, so each contact is placed at the dependency level
// one past the latest earlier contact sharing a node with it (a sort of (node, contact) keys in LDS) and each level
// runs in parallel - the same arithmetic, in the same per-node order, as the serial loop.
#include "bl_contacts.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#pragma clang fp contract(off)

namespace blg {

__constant__ ground_model k_gm_c[BL_SURF_N] = BL_GROUND_MODELS;

// ---------------------------------------------------------------------------------------------- geometry
__device__ __forceinline__ f3 ld3(float4 v) { return mk(v.x, v.y, v.z); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ f3 normalize(f3 a) {
    const float l = length(a);
    return l > 1e-12f ? mul(a, 1.0f / l) : mk(0, 0, 0);
}
__device__ __forceinline__ float max_edge2(f3 a, f3 b, f3 c) {
    const f3 ba = sub(b, a), cb = sub(c, b), ac = sub(a, c);
    return smax(dot(ba, ba), smax(dot(cb, cb), dot(ac, ac)));
}
// cell_hash (src/phys/world.cpp:23-25): the int products wrap as unsigned ones do
__device__ __forceinline__ uint32_t cell_hash(int x, int y, int z) {
    return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}
__device__ __forceinline__ int cell_of(float v, float inv_cell) { return (int)floorf(v * inv_cell); }

// closest_on_triangle (src/phys/softbody.h:158-190, Ericson 5.1.5)
__device__ f3 closest_on_triangle(f3 p, f3 a, f3 b, f3 c, f3& bary) {
    const f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { bary = mk(1, 0, 0); return a; }
    const f3 bp = sub(p, b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { bary = mk(0, 1, 0); return b; }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const float v = d1 / (d1 - d3);
        bary = mk(1 - v, v, 0);
        return add(a, mul(ab, v));
    }
    const f3 cp = sub(p, c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { bary = mk(0, 0, 1); return c; }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const float w = d2 / (d2 - d6);
        bary = mk(1 - w, 0, w);
        return add(a, mul(ac, w));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary = mk(0, 1 - w, w);
        return add(b, mul(sub(c, b), w));
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    bary = mk(1 - v - w, v, w);
    return add(add(a, mul(ab, v)), mul(ac, w));
}

// closest_on_segment, closest_seg_seg (src/phys/world.cpp:27-47)
__device__ __forceinline__ float closest_on_segment(f3 p, f3 a, f3 b) {
    const f3 ab = sub(b, a);
    const float l2 = dot(ab, ab);
    if (l2 < 1e-12f) return 0;
    return clampf(dot(sub(p, a), ab) / l2, 0, 1);
}
__device__ void closest_seg_seg(f3 p1, f3 q1, f3 p2, f3 q2, float& s, float& t) {
    const f3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    if (a <= 1e-12f && e <= 1e-12f) { s = t = 0; return; }
    if (a <= 1e-12f) { s = 0; t = clampf(f / e, 0, 1); return; }
    const float c = dot(d1, r);
    if (e <= 1e-12f) { t = 0; s = clampf(-c / a, 0, 1); return; }
    const float b = dot(d1, d2), denom = a * e - b * b;
    s = denom != 0 ? clampf((b * f - c * e) / denom, 0, 1) : 0;
    t = (b * s + f) / e;
    if (t < 0) { t = 0; s = clampf(-c / a, 0, 1); }
    else if (t > 1) { t = 1; s = clampf((b - c) / a, 0, 1); }
}

// near_triangle (src/phys/world.cpp:1135-1142) for a two-sided triangle
__device__ __forceinline__ bool near_triangle(f3 p, f3 a, f3 b, f3 c, f3 tu, float r) {
    const float s = dot(sub(p, a), tu);
    f3 bary;
    const f3 q = closest_on_triangle(p, a, b, c, bary);
    const f3 d = sub(q, p);
    return ffabs(s) <= r && dot(d, d) <= r * r;
}

__device__ __forceinline__ bool near_partner(const CArgs& c, uint32_t lb, f3 mn, f3 mx) {
    if (c.self[lb]) return true;
    for (uint32_t k = c.part_off[lb]; k < c.part_off[lb + 1]; k++) {
        const float* q = c.boxes + 6 * c.part[k];
        if (mn.x <= q[3] && mx.x >= q[0] && mn.y <= q[4] && mx.y >= q[1] && mn.z <= q[5] && mx.z >= q[2]) return true;
    }
    return false;
}

__device__ __forceinline__ bool same_group(const KArgs& a, uint32_t ga, uint32_t gb) {
    const int x = a.bodies[ga].group, y = a.bodies[gb].group;
    return x != 0 && x == y;
}

// ---------------------------------------------------------------------------------------------- rebuild stages
// Triangles stretched past 2.5 times their longest rest edge are torn (src/phys/world.cpp:1874-1877); and every
// 7th triangle's size sample for the cell size (:1916-1929), in the island's order.
__global__ void k_tris_torn(KArgs a, CArgs c) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntris) return;
    const uint4 tr = a.tri_abc[t];
    if (c.island_local[tr.w] < 0) return;
    const f3 pa = ld3(a.pm[c.parity][tr.x]), pb = ld3(a.pm[c.parity][tr.y]), pc = ld3(a.pm[c.parity][tr.z]);
    const float e2 = max_edge2(pa, pb, pc);
    if (!a.tri_torn[t] && a.tri_rest2[t] > 0 && e2 > 2.5f * 2.5f * a.tri_rest2[t]) a.tri_torn[t] = 1;
    const uint32_t k = t - c.body_tri0[tr.w];
    if (k % 7 == 0) c.samples[c.sample_base[tr.w] + k / 7] = fsqrt(e2);
}

// 1) the contacter nodes of the island's querying bodies near a partner: flag, then (after a scan) their records
__global__ void k_pts(KArgs a, CArgs c, int write) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nnodes) return;
    const uint32_t info = a.ninfo[i], gb = info >> 16;
    const int lb = c.island_local[gb];
    bool ok = lb >= 0 && c.query[lb] && (info & BL_NF_CONTACTER);
    f3 p = mk(0, 0, 0);
    if (ok) {
        p = ld3(a.pm[c.parity][i]);
        ok = near_partner(c, (uint32_t)lb, p, p);
    }
    if (!write) {
        c.flag[i] = ok;
        return;
    }
    if (!ok) return;
    const uint32_t k = c.scan[i];
    const int x = cell_of(p.x, c.inv_cell), y = cell_of(p.y, c.inv_cell), z = cell_of(p.z, c.inv_cell);
    c.pt_cell[k] = make_int4(x, y, z, (int)i);
    c.pt_pos[k] = make_float4(p.x, p.y, p.z, u2f((uint32_t)lb));
    c.keys[k] = ((uint64_t)(cell_hash(x, y, z) & c.mask) << 32) | k;
}

// buckets from the sorted (bucket, index) keys: bucket_start, the items, and run_end (the end of the run of one
// body's nodes in a bucket: a querying body skips its own nodes in one step)
__global__ void k_buckets(CArgs c, uint32_t n, uint32_t nbk, const uint64_t* sorted, uint32_t* bstart, uint32_t* items,
                          const float4* pos, uint32_t* run_end) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    const uint32_t cur = k < n ? (uint32_t)(sorted[k] >> 32) : nbk;
    const uint32_t prev = k > 0 ? (uint32_t)(sorted[k - 1] >> 32) : 0xffffffffu;
    if (k == 0) {
        for (uint32_t b = 0; b <= cur; b++) bstart[b] = 0;
    } else if (cur != prev) {
        for (uint32_t b = prev + 1; b <= cur; b++) bstart[b] = k;
    }
    if (k < n) items[k] = (uint32_t)sorted[k];
    if (pos && k < n) {
        const uint32_t body = f2u(pos[(uint32_t)sorted[k]].w);
        uint32_t e = k + 1;
        while (e < n && (uint32_t)(sorted[e] >> 32) == cur && f2u(pos[(uint32_t)sorted[e]].w) == body) e++;
        run_end[k] = e;
    }
}

// 2) each triangle of the querying bodies scans its cells for nodes in reach (count, then write in order)
__global__ void k_query_tris(KArgs a, CArgs c, int write) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntris) return;
    const uint4 tr = a.tri_abc[t];
    const int lbt = c.island_local[tr.w];
    uint32_t cnt = 0, out = write ? c.scan[t] : 0;
    if (lbt >= 0 && c.query[lbt] && !a.tri_torn[t] && c.npts > 0) {
        const BodyParam& O = a.bodies[tr.w];
        const float rq = c.rmax + c.margin + 0.0f + 0.0f;
        const f3 pa = ld3(a.pm[c.parity][tr.x]), pb = ld3(a.pm[c.parity][tr.y]), pc = ld3(a.pm[c.parity][tr.z]);
        const f3 tmn = vmin(pa, vmin(pb, pc)), tmx = vmax(pa, vmax(pb, pc));
        if (near_partner(c, lbt, sub(tmn, mk(rq, rq, rq)), add(tmx, mk(rq, rq, rq)))) {
            const f3 tn = cross(sub(pb, pa), sub(pc, pa));
            const float tn2 = dot(tn, tn);
            const f3 tu = tn2 > 1e-20f ? mul(tn, 1.0f / fsqrt(tn2)) : mk(0, 0, 0);
            const int x0 = cell_of(tmn.x - rq, c.inv_cell), x1 = cell_of(tmx.x + rq, c.inv_cell);
            const int y0 = cell_of(tmn.y - rq, c.inv_cell), y1 = cell_of(tmx.y + rq, c.inv_cell);
            const int z0 = cell_of(tmn.z - rq, c.inv_cell), z1 = cell_of(tmx.z + rq, c.inv_cell);
            if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) <= 512) {
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            const uint32_t bk = cell_hash(x, y, z) & c.mask;
                            for (uint32_t k = c.bstart[bk]; k < c.bstart[bk + 1]; k++) {
                                const uint32_t pi = c.items[k];
                                const int4 cell = c.pt_cell[pi];
                                const float4 pp = c.pt_pos[pi];
                                const uint32_t lbn = f2u(pp.w);
                                if (lbn == (uint32_t)lbt && !O.self_coll) { k = c.run_end[k] - 1; continue; }
                                if (cell.x != x || cell.y != y || cell.z != z) continue;
                                if (pp.x < tmn.x - rq || pp.y < tmn.y - rq || pp.z < tmn.z - rq || pp.x > tmx.x + rq ||
                                    pp.y > tmx.y + rq || pp.z > tmx.z + rq)
                                    continue;
                                const uint32_t node = (uint32_t)cell.w, gbn = c.island_body[lbn];
                                const BodyParam& B = a.bodies[gbn];
                                const bool self = lbn == (uint32_t)lbt;
                                if (self && (!B.self_coll || tr.x == node || tr.y == node || tr.z == node)) continue;
                                if (!self && same_group(a, gbn, tr.w)) continue;
                                const f3 p = mk(pp.x, pp.y, pp.z);
                                const float r = B.collision_radius + c.margin + 0.0f;
                                const float rb = r + 0.0f;
                                if (p.x < tmn.x - rb || p.y < tmn.y - rb || p.z < tmn.z - rb || p.x > tmx.x + rb || p.y > tmx.y + rb || p.z > tmx.z + rb) continue;
                                if (!near_triangle(p, pa, pb, pc, tu, r)) continue;
                                if (write) c.nt[out++] = make_uint4(gbn, tr.w, node, t);
                                cnt++;
                            }
                        }
            }
        }
    }
    if (!write) c.count[t] = cnt;
}

// ... and each capsule of those bodies scans for nodes within its radius
__global__ void k_query_caps(KArgs a, CArgs c, int write) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= a.ncaps) return;
    const uint4 cp = a.cap_data[ci];
    const int lbt = c.island_local[cp.z];
    uint32_t cnt = 0, out = write ? c.scan[ci] : 0;
    if (lbt >= 0 && c.query[lbt] && c.npts > 0) {
        const f3 pa = ld3(a.pm[c.parity][cp.x]), pb = ld3(a.pm[c.parity][cp.y]);
        const float rq2 = u2f(cp.w) + c.rmax + c.margin + 0.0f;
        const f3 mn = sub(vmin(pa, pb), mk(rq2, rq2, rq2)), mx = add(vmax(pa, pb), mk(rq2, rq2, rq2));
        if (near_partner(c, lbt, mn, mx)) {
            const int x0 = cell_of(mn.x, c.inv_cell), x1 = cell_of(mx.x, c.inv_cell), y0 = cell_of(mn.y, c.inv_cell);
            const int y1 = cell_of(mx.y, c.inv_cell), z0 = cell_of(mn.z, c.inv_cell), z1 = cell_of(mx.z, c.inv_cell);
            if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) <= 512)
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            const uint32_t bk = cell_hash(x, y, z) & c.mask;
                            for (uint32_t k = c.bstart[bk]; k < c.bstart[bk + 1]; k++) {
                                const uint32_t pi = c.items[k];
                                const int4 cell = c.pt_cell[pi];
                                const float4 pp = c.pt_pos[pi];
                                const uint32_t lbn = f2u(pp.w);
                                if (lbn == (uint32_t)lbt) { k = c.run_end[k] - 1; continue; }
                                if (cell.x != x || cell.y != y || cell.z != z) continue;
                                const uint32_t gbn = c.island_body[lbn];
                                if (same_group(a, gbn, cp.z)) continue;
                                const f3 p = mk(pp.x, pp.y, pp.z);
                                const float r = u2f(cp.w) + a.bodies[gbn].collision_radius + c.margin + 0.0f;
                                const float s = closest_on_segment(p, pa, pb);
                                const f3 d = sub(add(pa, mul(sub(pb, pa), s)), p);
                                if (dot(d, d) > r * r) continue;
                                if (write) c.nc[out++] = make_uint4(gbn, cp.z, (uint32_t)cell.w, ci);
                                cnt++;
                            }
                        }
        }
    }
    if (!write) c.count[ci] = cnt;
}

// 3) capsules against triangles: the triangles as primitives in every cell of their (radius-expanded) boxes
__global__ void k_prims(KArgs a, CArgs c, int write) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntris) return;
    const uint4 tr = a.tri_abc[t];
    const int lb = c.island_local[tr.w];
    uint32_t ncell = 0;
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1, z0 = 0, z1 = -1;
    if (lb >= 0 && c.query[lb] && !a.tri_torn[t]) {
        const float r = a.bodies[tr.w].collision_radius + c.margin + 0.0f;
        const f3 pa = ld3(a.pm[c.parity][tr.x]), pb = ld3(a.pm[c.parity][tr.y]), pc = ld3(a.pm[c.parity][tr.z]);
        const f3 mn = sub(vmin(pa, vmin(pb, pc)), mk(r, r, r)), mx = add(vmax(pa, vmax(pb, pc)), mk(r, r, r));
        if (near_partner(c, lb, mn, mx)) {
            x0 = cell_of(mn.x, c.inv_cell), x1 = cell_of(mx.x, c.inv_cell), y0 = cell_of(mn.y, c.inv_cell);
            y1 = cell_of(mx.y, c.inv_cell), z0 = cell_of(mn.z, c.inv_cell), z1 = cell_of(mx.z, c.inv_cell);
            const int nc = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
            if (nc <= 512) ncell = (uint32_t)nc;
        }
    }
    if (write == 0) {
        c.flag[t] = ncell > 0;
        c.count[t] = ncell;
        return;
    }
    if (!ncell) return;
    const uint32_t pid = c.scan[t], e0 = c.scan2[t];
    c.prim_tri[pid] = t;
    uint32_t e = e0;
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) {
                c.entry_pid[e] = pid;
                c.keys[e] = ((uint64_t)(cell_hash(x, y, z) & c.mask) << 32) | e;
                e++;
            }
}

// each capsule of the island collects the triangles of the other bodies whose boxes meet its own, once each, sorted
// by (body, triangle) (count, then write)
constexpr int kCtMax = 128;
__global__ void k_query_ct(KArgs a, CArgs c, int write) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= a.ncaps) return;
    const uint4 cp = a.cap_data[ci];
    const int lbc = c.island_local[cp.z];
    uint32_t found[kCtMax];
    int nf = 0;
    if (lbc >= 0 && c.nprims > 0) {
        const f3 pa = ld3(a.pm[c.parity][cp.x]), pb = ld3(a.pm[c.parity][cp.y]);
        const float rr = u2f(cp.w) + a.bodies[cp.z].collision_radius + c.margin + 0.0f;
        const f3 mn = sub(vmin(pa, pb), mk(rr, rr, rr)), mx = add(vmax(pa, pb), mk(rr, rr, rr));
        const int x0 = cell_of(mn.x, c.inv_cell), x1 = cell_of(mx.x, c.inv_cell), y0 = cell_of(mn.y, c.inv_cell);
        const int y1 = cell_of(mx.y, c.inv_cell), z0 = cell_of(mn.z, c.inv_cell), z1 = cell_of(mx.z, c.inv_cell);
        if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) <= 512)
            for (int z = z0; z <= z1; z++)
                for (int y = y0; y <= y1; y++)
                    for (int x = x0; x <= x1; x++) {
                        const uint32_t bk = cell_hash(x, y, z) & c.mask;
                        for (uint32_t k = c.bstart[bk]; k < c.bstart[bk + 1]; k++) {
                            const uint32_t t = c.prim_tri[c.entry_pid[c.items[k]]];
                            const uint4 tr = a.tri_abc[t];
                            if (tr.w == cp.z || same_group(a, cp.z, tr.w)) continue;
                            const f3 ta = ld3(a.pm[c.parity][tr.x]), tb = ld3(a.pm[c.parity][tr.y]), tc = ld3(a.pm[c.parity][tr.z]);
                            const f3 tmn = vmin(ta, vmin(tb, tc)), tmx = vmax(ta, vmax(tb, tc));
                            if (tmx.x < mn.x || tmx.y < mn.y || tmx.z < mn.z || tmn.x > mx.x || tmn.y > mx.y || tmn.z > mx.z) continue;
                            // (global triangle order is (body, triangle) order: keep the distinct ones sorted)
                            int pos = nf;
                            while (pos > 0 && found[pos - 1] > t) pos--;
                            if (pos > 0 && found[pos - 1] == t) continue;
                            if (nf == kCtMax) {
                                atomicOr(c.overflow, 1u);
                                continue;
                            }
                            for (int q = nf; q > pos; q--) found[q] = found[q - 1];
                            found[pos] = t;
                            nf++;
                        }
                    }
    }
    if (!write) {
        c.count[ci] = (uint32_t)nf;
        return;
    }
    uint32_t out = c.scan[ci];
    for (int q = 0; q < nf; q++) c.ct[out + q] = make_uint4(cp.z, a.tri_abc[found[q]].w, ci, found[q]);
}

// near_pairs (src/phys/world.cpp:2525-2596): the node-triangle candidates within reach of their relative travel
// until the next rebuild and not receding across the plane (flag; compaction by scan)
__global__ void k_near(KArgs a, CArgs c, const uint4* nt, uint32_t n, int write, uint4* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 pr = nt[i];
    if (write) {
        if (c.flag[i]) out[c.scan[i]] = pr;
        return;
    }
    bool keep = !a.tri_torn[pr.w];
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
        const float r0 = cr + 0.0f;
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
                const float sg = s0 >= 0 ? 1.0f : -1.0f;
                const float close = smax(0.0f, smax(-sg * dot(n0, da), smax(-sg * dot(n0, db), -sg * dot(n0, dc))));
                if (ffabs(s0) - 1.25f * close * c.horizon - 0.02f >= cr) keep = false;
            }
        }
    }
    c.flag[i] = keep;
}

// ---------------------------------------------------------------------------------------------- narrow phase
// pair_detect (src/phys/world.cpp:2807-2962) on the concatenated candidates (node-triangle, node-capsule,
// capsule-triangle): a hit flag and its record (compacted by scan, in pair order)
__global__ void k_narrow(KArgs a, PArgs p, int parity) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = p.nnt + p.nnc + p.nct;
    if (i >= n) return;
    const float4* P = a.pm[parity];
    const float4* Vv = a.vm[parity];
    Hit h;
    h.pair = i, h.s = 0, h.bary = mk(0, 0, 0), h.nrm = mk(0, 0, 0), h.pen = 0;
    bool hit = false;
    if (i < p.nnt) {
        const uint4 pr = p.nt[i];
        if (!a.tri_torn[pr.w]) {
            const uint4 tr = a.tri_abc[pr.w];
            const float r = a.bodies[pr.x].collision_radius;
            const f3 np = ld3(P[pr.z]), ap = ld3(P[tr.x]), bp = ld3(P[tr.y]), cp3 = ld3(P[tr.z]);
            const f3 lo = sub(vmin(ap, vmin(bp, cp3)), mk(r, r, r)), hi = add(vmax(ap, vmax(bp, cp3)), mk(r, r, r));
            if (!(np.x < lo.x || np.y < lo.y || np.z < lo.z || np.x > hi.x || np.y > hi.y || np.z > hi.z)) {
                f3 bary;
                const f3 cp = closest_on_triangle(np, ap, bp, cp3, bary);
                const f3 d = sub(np, cp);
                const float dist2 = dot(d, d);
                if (dist2 < r * r) {
                    const float dist = fsqrt(dist2);
                    f3 nrm;
                    if (dist > 1e-5f) nrm = divs(d, dist);
                    else {
                        nrm = normalize(cross(sub(bp, ap), sub(cp3, ap)));
                        const f3 vt = add(add(mul(ld3(Vv[tr.x]), bary.x), mul(ld3(Vv[tr.y]), bary.y)), mul(ld3(Vv[tr.z]), bary.z));
                        if (dot(sub(ld3(Vv[pr.z]), vt), nrm) > 0) nrm = mk(-nrm.x, -nrm.y, -nrm.z);
                    }
                    h.bary = bary, h.nrm = nrm, h.pen = r - dist;
                    hit = true;
                }
            }
        }
    } else if (i < p.nnt + p.nnc) {
        const uint4 pr = p.nc[i - p.nnt];
        const uint4 cap = a.cap_data[pr.w];
        const f3 np = ld3(P[pr.z]), ap = ld3(P[cap.x]), bp = ld3(P[cap.y]);
        const float t = closest_on_segment(np, ap, bp);
        const f3 cp = add(ap, mul(sub(bp, ap), t));
        const f3 d = sub(np, cp);
        const float r = a.bodies[pr.x].collision_radius + u2f(cap.w);
        const float dist2 = dot(d, d);
        if (dist2 < r * r) {
            const float dist = fsqrt(dist2);
            h.s = t;
            h.nrm = dist > 1e-5f ? divs(d, dist) : mk(0, 1, 0);
            h.pen = r - dist;
            hit = true;
        }
    } else {
        const uint4 pr = p.ct[i - p.nnt - p.nnc];
        const uint4 cap = a.cap_data[pr.z];
        const uint4 tri = a.tri_abc[pr.w];
        const f3 p0 = ld3(P[cap.x]), p1 = ld3(P[cap.y]);
        const f3 ta = ld3(P[tri.x]), tb = ld3(P[tri.y]), tc = ld3(P[tri.z]);
        float best = 1e30f, best_s = 0;
        f3 best_bary = mk(0, 0, 0), best_seg = mk(0, 0, 0), best_tri = mk(0, 0, 0);
        auto consider = [&](float s, f3 sp, f3 tp, f3 bary) {
            const f3 dd = sub(sp, tp);
            const float d2 = dot(dd, dd);
            if (d2 < best) best = d2, best_s = s, best_seg = sp, best_tri = tp, best_bary = bary;
        };
        f3 bary;
        f3 q = closest_on_triangle(p0, ta, tb, tc, bary);
        consider(0.0f, p0, q, bary);
        q = closest_on_triangle(p1, ta, tb, tc, bary);
        consider(1.0f, p1, q, bary);
        const f3 tv[3] = {ta, tb, tc};
        for (int e = 0; e < 3; e++) {
            float s, u;
            closest_seg_seg(p0, p1, tv[e], tv[(e + 1) % 3], s, u);
            const f3 sp = add(p0, mul(sub(p1, p0), s));
            const f3 tp = add(tv[e], mul(sub(tv[(e + 1) % 3], tv[e]), u));
            float bb[3] = {0, 0, 0};
            bb[e] = 1 - u;
            bb[(e + 1) % 3] = u;
            consider(s, sp, tp, mk(bb[0], bb[1], bb[2]));
        }
        const f3 tn = cross(sub(tb, ta), sub(tc, ta));
        const float d0 = dot(sub(p0, ta), tn), d1 = dot(sub(p1, ta), tn);
        bool crossing = false;
        if ((d0 < 0) != (d1 < 0) && ffabs(d0 - d1) > 1e-12f) {
            const float s = d0 / (d0 - d1);
            const f3 ip = add(p0, mul(sub(p1, p0), s));
            q = closest_on_triangle(ip, ta, tb, tc, bary);
            const f3 qi = sub(q, ip);
            if (dot(qi, qi) < 1e-8f) {
                crossing = true;
                best = 0, best_s = s, best_seg = ip, best_tri = q, best_bary = bary;
            }
        }
        const float r = u2f(cap.w) + a.bodies[pr.x].collision_radius;
        if (best < r * r) {
            const float dist = fsqrt(best);
            f3 nrm;
            float pen;
            if (!crossing && dist > 1e-5f) {
                nrm = divs(sub(best_seg, best_tri), dist);
                pen = r - dist;
            } else {
                nrm = normalize(tn);
                const f3 mid = mul(add(p0, p1), 0.5f);
                if (dot(sub(mid, ta), nrm) < 0) nrm = mk(-nrm.x, -nrm.y, -nrm.z);
                const float dm = smin(ffabs(d0), ffabs(d1)) / smax(1e-6f, length(tn));
                pen = r + smin(dm, 0.2f);
            }
            h.s = best_s, h.bary = best_bary, h.nrm = nrm, h.pen = pen;
            hit = true;
        }
    }
    p.hflag[i] = hit;
    p.hraw[i] = h;
}

__global__ void k_hits_compact(PArgs p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = p.nnt + p.nnc + p.nct;
    if (i >= n) return;
    if (p.hflag[i]) p.hits[p.hscan[i]] = p.hraw[i];
    if (i == n - 1) *p.nhits = p.hscan[i] + p.hflag[i];
}

// ---------------------------------------------------------------------------------------------- response
// The nodes and weights of a contact (collide_pairs, src/phys/world.cpp:2997-3076): side A then side B.
struct ContactNodes {
    uint32_t n[5];
    float w[5];
    int na, nb;
    float fric;
};

__device__ __forceinline__ void contact_nodes(const KArgs& a, const PArgs& p, const Hit& h, ContactNodes& cn) {
    if (h.pair < p.nnt) {
        const uint4 pr = p.nt[h.pair];
        const uint4 t = a.tri_abc[pr.w];
        cn.na = 1, cn.nb = 3;
        cn.n[0] = pr.z, cn.w[0] = 1.0f;
        cn.n[1] = t.x, cn.n[2] = t.y, cn.n[3] = t.z;
        cn.w[1] = h.bary.x, cn.w[2] = h.bary.y, cn.w[3] = h.bary.z;
        // (a tyre node would take its own friction: tyre nodes are refused when a body is added)
        cn.fric = 0.8f * smin(a.bodies[pr.x].contact_friction, a.bodies[pr.y].contact_friction);
    } else if (h.pair < p.nnt + p.nnc) {
        const uint4 pr = p.nc[h.pair - p.nnt];
        const uint4 cap = a.cap_data[pr.w];
        cn.na = 1, cn.nb = 2;
        cn.n[0] = pr.z, cn.w[0] = 1.0f;
        cn.n[1] = cap.x, cn.n[2] = cap.y;
        cn.w[1] = 1 - h.s, cn.w[2] = h.s;
        cn.fric = 0.8f * smin(a.bodies[pr.x].contact_friction, a.bodies[pr.y].contact_friction);
    } else {
        const uint4 pr = p.ct[h.pair - p.nnt - p.nnc];
        const uint4 cap = a.cap_data[pr.z];
        const uint4 t = a.tri_abc[pr.w];
        cn.na = 2, cn.nb = 3;
        cn.n[0] = cap.x, cn.n[1] = cap.y;
        cn.w[0] = 1 - h.s, cn.w[1] = h.s;
        cn.n[2] = t.x, cn.n[3] = t.y, cn.n[4] = t.z;
        cn.w[2] = h.bary.x, cn.w[3] = h.bary.y, cn.w[4] = h.bary.z;
        cn.fric = 0.8f * smin(a.bodies[pr.x].contact_friction, a.bodies[pr.y].contact_friction);
    }
}

// contact_force<NA, NB> (src/phys/world.cpp:59-91), both sides movable, no force snapshots
__device__ bool contact_force(const KArgs& a, int parity, const ContactNodes& cn, f3 nrm, float pen) {
    const float4* P = a.pm[parity];
    const float4* Vv = a.vm[parity];
    float W = 0;
    for (int i = 0; i < cn.na + cn.nb; i++) W += cn.w[i] * cn.w[i] * P[cn.n[i]].w;
    if (W <= 1e-12f) return false;
    const float m = 1.0f / W;
    f3 va = mk(0, 0, 0), vb = mk(0, 0, 0), acc = mk(0, 0, 0);
    for (int i = 0; i < cn.na; i++) {
        va = add(va, mul(ld3(Vv[cn.n[i]]), cn.w[i]));
        acc = add(acc, mul(ld3(a.F[cn.n[i]]), cn.w[i] * P[cn.n[i]].w));
    }
    for (int i = cn.na; i < cn.na + cn.nb; i++) {
        vb = add(vb, mul(ld3(Vv[cn.n[i]]), cn.w[i]));
        acc = sub(acc, mul(ld3(a.F[cn.n[i]]), cn.w[i] * P[cn.n[i]].w));
    }
    const f3 f_rel = mul(acc, m);
    const f3 f = primitive_collision(f_rel, sub(va, vb), m, nrm, a.dt, k_gm_c[BL_SURF_CONCRETE], pen, cn.fric, 2.0f, 0.0f, -1.0f);
    for (int i = 0; i < cn.na; i++) {
        const float4 F = a.F[cn.n[i]];
        a.F[cn.n[i]] = make_float4(F.x + f.x * cn.w[i], F.y + f.y * cn.w[i], F.z + f.z * cn.w[i], F.w);
    }
    for (int i = cn.na; i < cn.na + cn.nb; i++) {
        const float4 F = a.F[cn.n[i]];
        a.F[cn.n[i]] = make_float4(F.x - f.x * cn.w[i], F.y - f.y * cn.w[i], F.z - f.z * cn.w[i], F.w);
    }
    return true;
}

// One workgroup: windows of 1024 contacts; per window the (node, contact) keys sorted in LDS give each contact its
// predecessors (the latest earlier contact of the window on each of its nodes), the longest chain to it its level;
// the levels run one after another, the contacts of a level (no shared node) in parallel.
constexpr int kRespW = 1024, kRespKeys = 8192;
__global__ __launch_bounds__(kRespW) void k_response(KArgs a, PArgs p, int parity) {
    __shared__ uint64_t keys[kRespKeys];
    __shared__ int16_t preds[kRespW * 5];
    __shared__ uint16_t level[kRespW];
    __shared__ uint32_t sred[kRespW / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t nh = *p.nhits;
    uint32_t contacts = 0;
    for (uint32_t w0 = 0; w0 < nh; w0 += kRespW) {
        const uint32_t h = w0 + tid;
        const bool valid = h < nh;
        Hit hit;
        ContactNodes cn;
        cn.na = cn.nb = 0;
        if (valid) {
            hit = p.hits[h];
            contact_nodes(a, p, hit, cn);
        }
        for (int j = 0; j < 5; j++) {
            keys[tid * 5 + j] = (valid && j < cn.na + cn.nb) ? (((uint64_t)cn.n[j] << 32) | (tid << 3) | (uint32_t)j) : ~0ull;
            preds[tid * 5 + j] = -1;
        }
        for (uint32_t k = kRespW * 5 + tid; k < kRespKeys; k += kRespW) keys[k] = ~0ull;
        __syncthreads();
        // bitonic sort of the keys
        for (uint32_t size = 2; size <= kRespKeys; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t q = tid; q < kRespKeys / 2; q += kRespW) {
                    const uint32_t i = 2 * q - (q & (stride - 1));
                    const uint32_t j = i + stride;
                    const bool up = (i & size) == 0;
                    const uint64_t x = keys[i], y = keys[j];
                    if ((x > y) == up) keys[i] = y, keys[j] = x;
                }
                __syncthreads();
            }
        // predecessors: the previous key of the same node from another (earlier) contact
        for (uint32_t k = tid; k < kRespKeys; k += kRespW) {
            const uint64_t key = keys[k];
            if (key == ~0ull) continue;
            const uint32_t node = (uint32_t)(key >> 32), me = (uint32_t)(key >> 3) & (kRespW - 1);
            int q = (int)k - 1;
            while (q >= 0 && (uint32_t)(keys[q] >> 32) == node && ((uint32_t)(keys[q] >> 3) & (kRespW - 1)) == me) q--;
            if (q >= 0 && (uint32_t)(keys[q] >> 32) == node) preds[me * 5 + (key & 7)] = (int16_t)((keys[q] >> 3) & (kRespW - 1));
        }
        level[tid] = 0;
        __syncthreads();
        // levels: one past the deepest predecessor (a few rounds: the depth of the chains)
        int changed = 1;
        while (__syncthreads_or(changed)) {
            changed = 0;
            uint32_t L = level[tid];
            for (int j = 0; j < 5; j++) {
                const int pq = preds[tid * 5 + j];
                if (pq >= 0 && level[pq] + 1u > L) L = level[pq] + 1u;
            }
            __syncthreads();
            if (L != level[tid]) level[tid] = (uint16_t)L, changed = 1;
        }
        uint32_t my = valid ? level[tid] : 0, top = my;
        for (int off = 32; off > 0; off >>= 1) top = max(top, (uint32_t)__shfl_xor((int)top, off));
        if ((tid & 63) == 0) sred[tid >> 6] = top;
        __syncthreads();
        uint32_t maxl = 0;
        for (int w = 0; w < kRespW / 64; w++) maxl = max(maxl, sred[w]);
        for (uint32_t L = 0; L <= maxl; L++) {
            if (valid && my == L && contact_force(a, parity, cn, hit.nrm, hit.pen)) contacts++;
            __syncthreads();
        }
    }
    for (int off = 32; off > 0; off >>= 1) contacts += __shfl_xor(contacts, off);
    if ((tid & 63) == 0 && contacts) atomicAdd(&a.counters[CNT_CONTACTS], (unsigned long long)contacts);
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

inline int bits_for(uint32_t nbk) {
    int b = 0;
    while ((1u << b) < nbk) b++;
    return b;
}

// build_buckets' table size (src/phys/world.cpp:1985-1995): a power of two, at least 64 and twice the entries
inline uint32_t bucket_count(size_t entries) {
    uint32_t nbk = 64;
    while (nbk < entries * 2) nbk <<= 1;
    return nbk;
}

} // namespace

int ContactEngine::scan(const uint32_t* in, uint32_t* out, size_t n, hipStream_t st) {
    if (n == 0) return BL_OK;
    size_t bytes = 0;
    CHK(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st));
    RC(temp.ensure(bytes));
    CHK(rocprim::exclusive_scan(temp.p, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st));
    return BL_OK;
}

int ContactEngine::sort(const uint64_t* in, uint64_t* out, size_t n, int end_bit, hipStream_t st) {
    if (n == 0) return BL_OK;
    size_t bytes = 0;
    CHK(rocprim::radix_sort_keys(nullptr, bytes, in, out, n, 0u, (unsigned)end_bit, st));
    RC(temp.ensure(bytes));
    CHK(rocprim::radix_sort_keys(temp.p, bytes, in, out, n, 0u, (unsigned)end_bit, st));
    return BL_OK;
}

// The total of an exclusive scan over n flags or counts (waits for the stream).
int ContactEngine::total(const uint32_t* scan_out, const uint32_t* in, size_t n, uint32_t& tot, hipStream_t st) {
    tot = 0;
    if (n == 0) return BL_OK;
    uint32_t x[2] = {0, 0};
    CHK(hipMemcpyAsync(&x[0], scan_out + n - 1, 4, hipMemcpyDeviceToHost, st));
    CHK(hipMemcpyAsync(&x[1], in + n - 1, 4, hipMemcpyDeviceToHost, st));
    CHK(hipStreamSynchronize(st));
    tot = x[0] + x[1];
    return BL_OK;
}

void ContactEngine::release() {
    temp.release();
    for (Buf<uint32_t>* b : {&flag, &scanv, &scan2, &count, &bstart, &items, &run_end, &entry_pid, &prim_tri, &hflag, &hscan,
                             &nhits, &overflow, &isl_query, &isl_self, &isl_body, &part_off, &part, &sample_base, &body_tri0})
        b->release();
    isl_local.release(), boxes.release(), samples.release(), keys.release(), sorted.release();
    pt_cell.release(), pt_pos.release();
    for (Buf<uint4>* b : {&nt, &nt2, &nc, &ct, &all_nt, &all_nc, &all_ct}) b->release();
    hraw.release(), hits.release();
    nnt = nnc = nct = 0;
}

// The per-body tables the rebuilds read (first triangle, first size sample); the flags cleared.
int ContactEngine::setup(const HostView& hv) {
    const size_t nb = hv.bodies.size();
    std::vector<uint32_t> tri0(nb), s0(nb);
    uint32_t acc = 0;
    for (size_t i = 0; i < nb; i++) {
        tri0[i] = hv.bodies[i].tri0;
        s0[i] = acc;
        acc += (hv.bodies[i].ntri + 6) / 7;
    }
    nsamples = acc;
    RC(body_tri0.ensure(nb));
    RC(sample_base.ensure(nb));
    RC(samples.ensure(acc ? acc : 1));
    RC(overflow.ensure(1));
    RC(nhits.ensure(1));
    if (nb) {
        CHK(hipMemcpy(body_tri0.p, tri0.data(), 4 * nb, hipMemcpyHostToDevice));
        CHK(hipMemcpy(sample_base.p, s0.data(), 4 * nb, hipMemcpyHostToDevice));
    }
    CHK(hipMemset(overflow.p, 0, 4));
    CHK(hipMemset(nhits.p, 0, 4));
    nnt = nnc = nct = 0;
    unsupported = false;
    return BL_OK;
}

// World::build_islands (src/phys/world.cpp:636-762) for awake node/beam bodies: union-find over the expanded bounds
// of bodies that can touch (contacters against triangles or capsules, capsules against triangles); an island's
// bodies in body order, islands by their first body. Islands of several bodies, or of one self-colliding body,
// search for pairs.
void ContactEngine::build_islands(const HostView& hv, float frame_time) {
    const size_t nb = hv.bodies.size();
    std::vector<int> parent(nb);
    for (size_t i = 0; i < nb; i++) parent[i] = (int)i;
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::vector<float> bx(nb * 6);
    for (size_t i = 0; i < nb; i++) {
        const float m = smin(hv.max_speed[i], 200.0f) * frame_time * 2.0f + 0.3f;
        const float r = m + 0.0f; // (no hull triangles: no hull reach)
        const float mn[3] = {hv.mn[i].x, hv.mn[i].y, hv.mn[i].z}, mx[3] = {hv.mx[i].x, hv.mx[i].y, hv.mx[i].z};
        for (int k = 0; k < 3; k++) bx[6 * i + k] = mn[k] - r, bx[6 * i + 3 + k] = mx[k] + r;
    }
    if (hv.inter_body)
        for (size_t i = 0; i < nb; i++)
            for (size_t j = i + 1; j < nb; j++) {
                const float *p = &bx[6 * i], *q = &bx[6 * j];
                if (!(p[0] <= q[3] && p[3] >= q[0] && p[1] <= q[4] && p[4] >= q[1] && p[2] <= q[5] && p[5] >= q[2])) continue;
                const BodyView &bi = hv.bodies[i], &bj = hv.bodies[j];
                if (bi.group != 0 && bi.group == bj.group) continue;
                const bool ti = bi.ntri > 0, tj = bj.ntri > 0, ci = bi.ncap > 0, cj = bj.ncap > 0;
                const bool can = (bi.contacters && (tj || cj)) || (bj.contacters && (ti || ci)) || (ci && tj) || (cj && ti);
                if (!can) continue;
                const int x = find((int)i), y = find((int)j);
                if (x != y) parent[x] = y;
            }
    islands.clear();
    std::vector<int> island_of(nb, -1);
    for (size_t i = 0; i < nb; i++) {
        const int r = find((int)i);
        if (island_of[r] < 0) island_of[r] = (int)islands.size(), islands.emplace_back();
        islands[island_of[r]].push_back((uint32_t)i);
    }
    pair_islands.clear();
    for (auto& is : islands)
        if (hv.inter_body && (is.size() > 1 || hv.bodies[is[0]].self)) pair_islands.push_back(is);
}

// One island's pair lists (rebuild_pairs, then near_pairs), appended to the frame's.
int ContactEngine::rebuild_island(const KArgs& a, const HostView& hv, const std::vector<uint32_t>& isl,
                                  const std::vector<float>& body_v2, const std::vector<float>& body_box,
                                  const std::vector<float>& smp, hipStream_t st, int parity, int64_t& island_pairs) {
    const uint32_t nb = (uint32_t)isl.size();
    float rmax = 0, slow_speed = 0;
    bool any_caps = false, any_tris = false;
    for (uint32_t g : isl) {
        const BodyView& b = hv.bodies[g];
        rmax = smax(rmax, b.collision_radius);
        any_caps |= b.ncap > 0;
        any_tris |= b.ntri > 0;
    }
    for (uint32_t g : isl) {
        const float ms = std::sqrt(body_v2[g]);
        if (ms <= 8.0f) slow_speed = smax(slow_speed, ms);
        else unsupported = true; // a fast body: its own pair schedule (fast_pairs) is not stepped here
    }
    const float speed = slow_speed + 0.5f;
    const float margin = 2.0f * speed * (float)interval * hv.dt + 0.02f;
    // the cell size from every 7th triangle's longest edge (a double sum in body, triangle order)
    double tri_size = 0;
    int nsmp = 0;
    for (uint32_t g : isl) {
        const BodyView& b = hv.bodies[g];
        uint32_t base = 0;
        for (uint32_t q = 0; q < g; q++) base += (hv.bodies[q].ntri + 6) / 7;
        for (uint32_t t = 0; t < b.ntri; t += 7) tri_size += smp[base + t / 7], nsmp++;
    }
    const float mean = nsmp ? (float)(tri_size / nsmp) : 1.0f;
    const float inv_cell = 1.0f / clampf(mean + 2.0f * (rmax + margin), 0.25f, 1.0f);
    const float reach = rmax + margin;
    // partners: island bodies of other groups whose expanded boxes meet
    std::vector<float> bxs(nb * 6);
    for (uint32_t k = 0; k < nb; k++)
        for (int q = 0; q < 3; q++) {
            bxs[6 * k + q] = body_box[6 * isl[k] + q] - (reach + 0.0f + 0.0f);
            bxs[6 * k + 3 + q] = body_box[6 * isl[k] + 3 + q] + (reach + 0.0f + 0.0f);
        }
    std::vector<std::vector<uint32_t>> partners(nb);
    for (uint32_t x = 0; x < nb; x++)
        for (uint32_t y = x + 1; y < nb; y++) {
            const BodyView &bx = hv.bodies[isl[x]], &by = hv.bodies[isl[y]];
            if (bx.group != 0 && bx.group == by.group) continue;
            const float *p = &bxs[6 * x], *q = &bxs[6 * y];
            if (!(p[0] <= q[3] && p[3] >= q[0] && p[1] <= q[4] && p[4] >= q[1] && p[2] <= q[5] && p[5] >= q[2])) continue;
            partners[x].push_back(y), partners[y].push_back(x);
        }
    std::vector<uint32_t> poff(nb + 1, 0), plist, query(nb), self(nb);
    for (uint32_t k = 0; k < nb; k++) {
        plist.insert(plist.end(), partners[k].begin(), partners[k].end());
        poff[k + 1] = (uint32_t)plist.size();
        self[k] = hv.bodies[isl[k]].self;
        query[k] = !partners[k].empty() || self[k];
    }
    std::vector<int> local(hv.bodies.size(), -1);
    for (uint32_t k = 0; k < nb; k++) local[isl[k]] = (int)k;
    RC(isl_local.ensure(local.size()));
    RC(isl_query.ensure(nb));
    RC(isl_self.ensure(nb));
    RC(isl_body.ensure(nb));
    RC(boxes.ensure(bxs.size()));
    RC(part_off.ensure(nb + 1));
    RC(part.ensure(plist.size() + 1));
    CHK(hipMemcpyAsync(isl_local.p, local.data(), sizeof(int) * local.size(), hipMemcpyHostToDevice, st));
    CHK(hipMemcpyAsync(isl_query.p, query.data(), 4 * nb, hipMemcpyHostToDevice, st));
    CHK(hipMemcpyAsync(isl_self.p, self.data(), 4 * nb, hipMemcpyHostToDevice, st));
    CHK(hipMemcpyAsync(isl_body.p, isl.data(), 4 * nb, hipMemcpyHostToDevice, st));
    CHK(hipMemcpyAsync(boxes.p, bxs.data(), sizeof(float) * bxs.size(), hipMemcpyHostToDevice, st));
    CHK(hipMemcpyAsync(part_off.p, poff.data(), 4 * (nb + 1), hipMemcpyHostToDevice, st));
    if (!plist.empty()) CHK(hipMemcpyAsync(part.p, plist.data(), 4 * plist.size(), hipMemcpyHostToDevice, st));
    const size_t N = a.nnodes, T = a.ntris, C = a.ncaps;
    const size_t big = std::max(N, std::max(T, C)) + 1;
    RC(flag.ensure(big));
    RC(scanv.ensure(big));
    RC(scan2.ensure(big));
    RC(count.ensure(big));
    CArgs c{};
    c.parity = parity, c.margin = margin, c.inv_cell = inv_cell, c.rmax = rmax;
    c.horizon = (float)(interval > 1 ? interval : 1) * hv.dt;
    c.island_local = isl_local.p, c.query = isl_query.p, c.self = isl_self.p, c.island_body = isl_body.p;
    c.boxes = boxes.p, c.part_off = part_off.p, c.part = part.p;
    c.body_tri0 = body_tri0.p, c.sample_base = sample_base.p, c.samples = samples.p, c.overflow = overflow.p;
    c.flag = flag.p, c.scan = scanv.p, c.scan2 = scan2.p, c.count = count.p;
    uint32_t nnt_i = 0, nnc_i = 0, nct_i = 0;
    // 1) the contacter nodes near a partner into the hash
    hipLaunchKernelGGL(k_pts, grid_for(N), dim3(256), 0, st, a, c, 0);
    RC(scan(flag.p, scanv.p, N, st));
    RC(total(scanv.p, flag.p, N, c.npts, st));
    if (c.npts && (any_tris || any_caps)) {
        const uint32_t nbk = bucket_count(c.npts);
        c.mask = nbk - 1;
        RC(keys.ensure(c.npts));
        RC(sorted.ensure(c.npts));
        RC(pt_cell.ensure(c.npts));
        RC(pt_pos.ensure(c.npts));
        RC(bstart.ensure((size_t)nbk + 1));
        RC(items.ensure(c.npts));
        RC(run_end.ensure(c.npts));
        c.keys = keys.p, c.pt_cell = pt_cell.p, c.pt_pos = pt_pos.p;
        hipLaunchKernelGGL(k_pts, grid_for(N), dim3(256), 0, st, a, c, 1);
        RC(sort(keys.p, sorted.p, c.npts, 32 + bits_for(nbk), st));
        hipLaunchKernelGGL(k_buckets, grid_for((size_t)c.npts + 1), dim3(256), 0, st, c, c.npts, nbk, sorted.p, bstart.p,
                           items.p, (const float4*)pt_pos.p, run_end.p);
        c.bstart = bstart.p, c.items = items.p, c.run_end = run_end.p;
        // 2) the triangles, then the capsules, of the querying bodies against the hash (count, scan, write)
        if (T) {
            hipLaunchKernelGGL(k_query_tris, grid_for(T), dim3(256), 0, st, a, c, 0);
            RC(scan(count.p, scanv.p, T, st));
            RC(total(scanv.p, count.p, T, nnt_i, st));
            if (nnt_i) {
                RC(nt.ensure(nnt_i));
                c.nt = nt.p;
                hipLaunchKernelGGL(k_query_tris, grid_for(T), dim3(256), 0, st, a, c, 1);
            }
        }
        if (C) {
            hipLaunchKernelGGL(k_query_caps, grid_for(C), dim3(256), 0, st, a, c, 0);
            RC(scan(count.p, scanv.p, C, st));
            RC(total(scanv.p, count.p, C, nnc_i, st));
            if (nnc_i) {
                RC(nc.ensure(nnc_i));
                c.nc = nc.p;
                hipLaunchKernelGGL(k_query_caps, grid_for(C), dim3(256), 0, st, a, c, 1);
            }
        }
    }
    // 3) capsules against triangles: the triangles into every cell of their boxes, the capsules query
    if (any_caps && any_tris) {
        hipLaunchKernelGGL(k_prims, grid_for(T), dim3(256), 0, st, a, c, 0);
        uint32_t nprims = 0, nentries = 0;
        RC(scan(flag.p, scanv.p, T, st));
        RC(total(scanv.p, flag.p, T, nprims, st));
        RC(scan(count.p, scan2.p, T, st));
        RC(total(scan2.p, count.p, T, nentries, st));
        c.nprims = nprims;
        if (nprims) {
            const uint32_t nbk = bucket_count(nentries);
            c.mask = nbk - 1;
            RC(keys.ensure(nentries));
            RC(sorted.ensure(nentries));
            RC(entry_pid.ensure(nentries));
            RC(prim_tri.ensure(nprims));
            RC(bstart.ensure((size_t)nbk + 1));
            RC(items.ensure(nentries));
            c.keys = keys.p, c.entry_pid = entry_pid.p, c.prim_tri = prim_tri.p;
            hipLaunchKernelGGL(k_prims, grid_for(T), dim3(256), 0, st, a, c, 1);
            RC(sort(keys.p, sorted.p, nentries, 32 + bits_for(nbk), st));
            hipLaunchKernelGGL(k_buckets, grid_for((size_t)nentries + 1), dim3(256), 0, st, c, nentries, nbk, sorted.p,
                               bstart.p, items.p, (const float4*)nullptr, (uint32_t*)nullptr);
            c.bstart = bstart.p, c.items = items.p;
            if (C) {
                hipLaunchKernelGGL(k_query_ct, grid_for(C), dim3(256), 0, st, a, c, 0);
                RC(scan(count.p, scanv.p, C, st));
                RC(total(scanv.p, count.p, C, nct_i, st));
                if (nct_i) {
                    RC(ct.ensure(nct_i));
                    c.ct = ct.p;
                    hipLaunchKernelGGL(k_query_ct, grid_for(C), dim3(256), 0, st, a, c, 1);
                }
            }
        }
    }
    // near_pairs on the node-triangle candidates, then the island's lists onto the frame's
    uint32_t nkeep = 0;
    if (nnt_i) {
        RC(nt2.ensure(nnt_i));
        RC(flag.ensure(nnt_i));
        RC(scanv.ensure(nnt_i));
        c.flag = flag.p, c.scan = scanv.p;
        hipLaunchKernelGGL(k_near, grid_for(nnt_i), dim3(256), 0, st, a, c, (const uint4*)nt.p, nnt_i, 0, nt2.p);
        RC(scan(flag.p, scanv.p, nnt_i, st));
        RC(total(scanv.p, flag.p, nnt_i, nkeep, st));
        hipLaunchKernelGGL(k_near, grid_for(nnt_i), dim3(256), 0, st, a, c, (const uint4*)nt.p, nnt_i, 1, nt2.p);
    }
    if (nkeep || nnc_i || nct_i) {
        // (the lists grow: keep what earlier islands wrote)
        auto grow = [&](Buf<uint4>& b, size_t have, size_t need) -> int {
            if (b.p && need <= b.cap) return BL_OK;
            Buf<uint4> nb2;
            RC(nb2.ensure(need));
            if (have) CHK(hipMemcpyAsync(nb2.p, b.p, sizeof(uint4) * have, hipMemcpyDeviceToDevice, st));
            CHK(hipStreamSynchronize(st));
            b.release();
            b = nb2;
            nb2.p = nullptr;
            return BL_OK;
        };
        RC(grow(all_nt, nnt, nnt + nkeep));
        RC(grow(all_nc, nnc, nnc + nnc_i));
        RC(grow(all_ct, nct, nct + nct_i));
        if (nkeep) CHK(hipMemcpyAsync(all_nt.p + nnt, nt2.p, sizeof(uint4) * nkeep, hipMemcpyDeviceToDevice, st));
        if (nnc_i) CHK(hipMemcpyAsync(all_nc.p + nnc, nc.p, sizeof(uint4) * nnc_i, hipMemcpyDeviceToDevice, st));
        if (nct_i) CHK(hipMemcpyAsync(all_ct.p + nct, ct.p, sizeof(uint4) * nct_i, hipMemcpyDeviceToDevice, st));
    }
    nnt += nkeep, nnc += nnc_i, nct += nct_i;
    island_pairs = (int64_t)nkeep + nnc_i + nct_i;
    CHK(hipStreamSynchronize(st));
    CHK(hipGetLastError());
    uint32_t ov = 0;
    CHK(hipMemcpy(&ov, overflow.p, 4, hipMemcpyDeviceToHost));
    if (ov) unsupported = true;
    return BL_OK;
}

// The frame's pair lists anew (its first substep and every `interval` after): torn triangles and size samples
// first, then each pair island's rebuild, in island order.
int ContactEngine::rebuild(const KArgs& a, const HostView& hv, const std::vector<float>& body_v2,
                           const std::vector<float>& body_box, hipStream_t st, int parity) {
    nnt = nnc = nct = 0;
    if (pair_islands.empty()) return BL_OK;
    std::vector<int> local(hv.bodies.size(), -1);
    for (auto& is : pair_islands)
        for (uint32_t g : is) local[g] = 0;
    RC(isl_local.ensure(local.size()));
    CHK(hipMemcpyAsync(isl_local.p, local.data(), sizeof(int) * local.size(), hipMemcpyHostToDevice, st));
    CArgs c{};
    c.parity = parity, c.island_local = isl_local.p, c.samples = samples.p, c.sample_base = sample_base.p;
    c.body_tri0 = body_tri0.p;
    if (a.ntris) hipLaunchKernelGGL(k_tris_torn, grid_for(a.ntris), dim3(256), 0, st, a, c);
    std::vector<float> smp(nsamples ? nsamples : 1, 0.0f);
    if (nsamples) CHK(hipMemcpyAsync(smp.data(), samples.p, 4 * nsamples, hipMemcpyDeviceToHost, st));
    CHK(hipStreamSynchronize(st));
    for (auto& is : pair_islands) {
        int64_t pairs = 0;
        RC(rebuild_island(a, hv, is, body_v2, body_box, smp, st, parity, pairs));
        pair_rebuilds++;
        contact_pairs = pairs; // (the last island's count, as the reference's counter keeps)
    }
    return BL_OK;
}

// The substep's contacts: narrow phase over the frame's pairs, order-preserving compaction, the response.
int ContactEngine::collide(const KArgs& a, hipStream_t st, int parity) {
    const size_t n = nnt + nnc + nct;
    narrow_tests += (int64_t)n;
    if (n == 0) return BL_OK;
    RC(hflag.ensure(n));
    RC(hscan.ensure(n));
    RC(hraw.ensure(n));
    RC(hits.ensure(n));
    PArgs p{};
    p.nt = all_nt.p, p.nc = all_nc.p, p.ct = all_ct.p;
    p.nnt = (uint32_t)nnt, p.nnc = (uint32_t)nnc, p.nct = (uint32_t)nct;
    p.hflag = hflag.p, p.hscan = hscan.p, p.hraw = hraw.p, p.hits = hits.p, p.nhits = nhits.p;
    hipLaunchKernelGGL(k_narrow, grid_for(n), dim3(256), 0, st, a, p, parity);
    RC(scan(hflag.p, hscan.p, n, st));
    hipLaunchKernelGGL(k_hits_compact, grid_for(n), dim3(256), 0, st, p);
    hipLaunchKernelGGL(k_response, dim3(1), dim3(kRespW), 0, st, a, p, parity);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

} // namespace blg
