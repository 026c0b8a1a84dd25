// The narrow phase of the contacts, shared by its own kernel (k_narrow) and the force phase's launch that runs it
// alongside (k_forces_narrow): pair_detect (src/phys/world.cpp:2807-2962) on the concatenated candidates
// (node-triangle, node-capsule, capsule-triangle), 256 per block: each block's hits packed to the front of its 256
// slots, in candidate order, and its count. With the geometry it needs (src/phys/softbody.h:158-190,
// src/phys/world.cpp:27-47).
#pragma once

#include "bl_contacts.h"

namespace blg {

#pragma clang fp contract(off)

__device__ __forceinline__ f3 ld3(float4 v) { return mk(v.x, v.y, v.z); }

__device__ __forceinline__ f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

__device__ __forceinline__ f3 normalize(f3 a) {
    const float l = length(a);
    return l > 1e-12f ? mul(a, 1.0f / l) : mk(0, 0, 0);
}

// closest_on_triangle (src/phys/softbody.h:158-190, Ericson 5.1.5)
__device__ inline f3 closest_on_triangle(f3 p, f3 a, f3 b, f3 c, f3& bary) {
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

__device__ inline void closest_seg_seg(f3 p1, f3 q1, f3 p2, f3 q2, float& s, float& t) {
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


__device__ __forceinline__ void narrow_block(const KArgs& a, const PArgs& pin, int parity, uint32_t blk) {
    __shared__ uint32_t wcnt[4];
    PArgs p = pin;
    p.nnt = pin.st->nnt, p.nnc = pin.st->nnc, p.nct = pin.st->nct;
    if (pin.st->nfnc) p.nc = pin.ncm; // (the node-capsule list merged with the fast bodies')
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t n = p.nnt + p.nnc + p.nct;
    const float4* P = a.pm[parity];
    const float4* Vv = a.vm[parity];
    Hit h;
    h.pair = i, h.s = 0, h.bary = mk(0, 0, 0), h.nrm = mk(0, 0, 0), h.pen = 0;
    h.n[0] = h.n[1] = h.n[2] = h.n[3] = h.n[4] = 0, h.fric = 0, h.shape = 0;
    bool hit = false;
    if (i >= n) {
        // (past the end: no hit; the block's compaction below needs every thread)
    } else if (i < p.nnt) {
        const uint4 pr = p.nt[i];
        if (!a.tri_torn[pr.w]) {
            const uint4 tr = a.tri_abc[pr.w];
            const BodyParam& Bt = a.bodies[pr.y];
            const bool two_sided = !a.tri_hull[pr.w];
            // (collide_pairs' contact of the candidate: src/phys/world.cpp:2997-3017; a tyre node would take its own
            // friction - tyre nodes are refused when a body is added)
            h.n[0] = pr.z, h.n[1] = tr.x, h.n[2] = tr.y, h.n[3] = tr.z;
            h.fric = 0.8f * smin(a.bodies[pr.x].contact_friction, Bt.contact_friction);
            const float r = Bt.faces_only && !two_sided ? Bt.face_skin : a.bodies[pr.x].collision_radius;
            const f3 np = ld3(P[pr.z]), ap = ld3(P[tr.x]), bp = ld3(P[tr.y]), cp3 = ld3(P[tr.z]);
            bool go;
            if (!two_sided) {
                // a hull triangle: over its face the node is pushed out along the face's normal, from as deep as
                // the hull's depth behind it; off the face, outside, the ordinary contact with its edges
                const f3 fn = normalize(cross(sub(bp, ap), sub(cp3, ap)));
                const float sf = dot(sub(np, ap), fn);
                go = !(sf >= r || sf < -Bt.hull_depth);
                if (go) {
                    f3 bary;
                    const f3 q = sub(np, mul(fn, sf));
                    const f3 cq = sub(closest_on_triangle(q, ap, bp, cp3, bary), q);
                    if (dot(cq, cq) <= 1e-6f) {
                        h.bary = bary, h.nrm = fn, h.pen = r - sf;
                        hit = true, go = false;
                    } else if (sf < 0 || Bt.faces_only) {
                        go = false;
                    }
                }
            } else {
                const f3 lo = sub(vmin(ap, vmin(bp, cp3)), mk(r, r, r)), hi = add(vmax(ap, vmax(bp, cp3)), mk(r, r, r));
                go = !(np.x < lo.x || np.y < lo.y || np.z < lo.z || np.x > hi.x || np.y > hi.y || np.z > hi.z);
            }
            if (go) {
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
            // (a node behind a hull face deeper than its contact radius is pushed out: hull_push_out)
            h.shape = 1u | 3u << 4 | (!two_sided && h.pen > a.bodies[pr.x].collision_radius ? 1u << 8 : 0u);
        }
    } else if (i < p.nnt + p.nnc) {
        const uint4 pr = p.nc[i - p.nnt];
        const uint4 cap = a.cap_data[pr.w];
        h.n[0] = pr.z, h.n[1] = cap.x, h.n[2] = cap.y, h.shape = 1u | 2u << 4;
        h.fric = 0.8f * smin(a.bodies[pr.x].contact_friction, a.bodies[pr.y].contact_friction);
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
        h.n[0] = cap.x, h.n[1] = cap.y, h.n[2] = tri.x, h.n[3] = tri.y, h.n[4] = tri.z, h.shape = 2u | 3u << 4;
        h.fric = 0.8f * smin(a.bodies[pr.x].contact_friction, a.bodies[pr.y].contact_friction);
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
    // the block's hits to the front of its slot range, in candidate order
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t m = __ballot(hit);
    if (lane == 0) wcnt[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < 4; w++) off += w < wv ? wcnt[w] : 0u, tot += wcnt[w];
    if (hit) p.hraw[blk * 256 + off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = h;
    if (threadIdx.x == 0) p.bcnt[blk] = tot;
}

} // namespace blg
