// The soft-body substep on gfx950: beam spring-damper forces with yield and break, their atomic-free gather onto
// the nodes, the static (ground) contacts and the semi-implicit integration.
//
// Reference semantics: SoftBody::clear_forces / compute_beam_forces / integrate_nodes (src/phys/softbody.cpp:
// 330-381, 412-470, 769-795) and World::collide_static (src/phys/world.cpp:781-869) with
// StaticWorld::collide_point (src/phys/static_world.cpp:364-454), in the phase order of World::simulate_island
// (src/phys/world.cpp:3326-3591) for bodies without inter-body contact pairs.
//
// Layout in HBM (all bodies concatenated; see bl_world.h):
//   nodes      pm[2][N] float4 {p, inv_mass}, vm[2][N] float4 {v, mass}  (double-buffered: a substep reads one
//              buffer and writes the other, so no node is read after it was integrated)
//   incidences one record per (beam, end node), node-major and in beam-index order within a node (CSR noff[]):
//              iA float4 {L, k, d, strength}, iM float4 {max_pos, max_neg, beam index, flags}, iOth the other end.
//              Each end keeps its own copy of the beam's mutable state; both copies see the same inputs and run
//              the same arithmetic, so they stay identical and no end ever reads what the other writes.
//   chunks     contiguous node ranges of one body whose incidences fit in LDS (ch_n0[]): one workgroup each.
//
// A chunk's workgroup evaluates its incidences in parallel (one thread per incidence, forces into LDS), then one
// thread per node adds them up in beam-index order - the order in which the reference's serial loop scatters
// f[a] += fv, f[b] -= fv - so that every node force is bit-identical without atomics.
//
// Three launch forms share that chunk step:
//   k_substep          one launch per substep, one workgroup per chunk (captured in a hipGraph by the host);
//   k_body_persistent  one workgroup per body, many substeps per launch (small bodies);
//   k_resident         one workgroup per chunk, many substeps per launch, the incidences' state and the chunk's
//                      nodes held in registers across substeps, a grid barrier between substeps (cooperative
//                      launch, so every workgroup is resident).
#include "bl_world.h"
#include "bl_narrow.h"

#pragma clang fp contract(off)

namespace blg {

__constant__ ground_model k_gm[BL_SURF_N] = BL_GROUND_MODELS;

// ---------------------------------------------------------------------------------------------- one beam end
enum : uint32_t { EV_CHANGED = 1, EV_BROKE = 2 };

// One incidence of compute_beam_forces (src/phys/softbody.cpp:416-468) on this end's copy of the beam state
// (A = L, k, d, strength; M = max_pos, max_neg, beam, flags). Returns true with the force signed for this end, or
// false when the beam gives none this substep; `ev` reports a state change (yield, break) and a break.
__device__ __forceinline__ bool beam_eval(float4& A, float4& M, const float4 Pa, const float4 Pb, const float4 Va,
                                          const float4 Vb, const float4* __restrict__ bcold, f3& out, float& slen_out,
                                          uint32_t& ev) {
    ev = 0;
    uint32_t tf = f2u(M.w);
    if (tf & (TF_BROKEN | TF_SKIP)) return false;
    const f3 dis = sub(mk(Pa.x, Pa.y, Pa.z), mk(Pb.x, Pb.y, Pb.z));
    const float len2 = dot(dis, dis);
    if (len2 < 1e-12f) return false;
    const float inv_len = 1.0f / fsqrt(len2);
    const float len = len2 * inv_len;
    float L = A.x, strength = A.w, max_pos = M.x, max_neg = M.y;
    const float diff = len - L;
    const f3 dv = sub(mk(Va.x, Va.y, Va.z), mk(Vb.x, Vb.y, Vb.z));
    float k = A.y, d = A.z;
    const uint32_t type = (tf >> TF_TYPE_SHIFT) & 3u;
    const bool brk = (tf & TF_BRK) != 0;
    const uint32_t beam = f2u(M.z);
    if (type == BL_BT_ROPE) {
        if (diff < 0) { k = 0; d *= 0.1f; }
    } else if (type == BL_BT_SUPPORT) {
        if (diff > 0) {
            k = 0;
            d *= 0.1f;
            if (brk && diff > L * bcold[beam].z) {
                M.w = u2f(tf | TF_BROKEN);
                ev = EV_CHANGED | EV_BROKE;
                return false;
            }
        }
    }
    float slen = -k * diff - d * dot(dv, dis) * inv_len;
    if (!(tf & TF_NODEFORM) && k != 0) {
        if (slen > max_pos && diff < 0) {
            const float4 cold = bcold[beam]; // plastic, L0, support_limit
            const float yield_len = max_pos / k;
            const float def = diff + yield_len * (1.0f - cold.x);
            const float Lold = L;
            L = smax(smin(0.1f, 0.3f * cold.y), L + def);
            slen = slen - (slen - max_pos) * 0.5f;
            if (L < Lold) max_pos *= Lold / L;
            ev = EV_CHANGED;
        } else if (slen < max_neg && diff > 0) {
            const float4 cold = bcold[beam];
            const float yield_len = max_neg / k;
            const float def = diff + yield_len * (1.0f - cold.x);
            const float Lold = L;
            L = L + def;
            slen = slen - (slen - max_neg) * 0.5f;
            if (Lold > 0 && L > Lold) max_neg *= L / Lold;
            if (!(tf & TF_DUCTILE)) strength -= def * k;
            ev = EV_CHANGED;
        }
    }
    if (ev) {
        A.x = L, A.w = strength;
        M.x = max_pos, M.y = max_neg;
    }
    if (brk && ffabs(slen) > strength && !(tf & TF_NOBREAK)) {
        M.w = u2f(tf | TF_BROKEN);
        ev = EV_CHANGED | EV_BROKE;
        return false;
    }
    slen_out = slen;
    const f3 fv = mul(dis, slen * inv_len);
    // f[b] -= fv is f[b] + (-fv) to the bit: the b end stores the negated force and the gather only adds
    out = (tf & TF_ISB) ? mk(-fv.x, -fv.y, -fv.z) : fv;
    return true;
}

// ---------------------------------------------------------------------------------------------- node phase
// Heightfield::sample (src/phys/static_world.cpp:189-205) on the dense heights
__device__ __forceinline__ bool terrain_sample(const KArgs& a, float x, float z, float& out_h, f3& n) {
    const float lx = (x - a.tox) / a.tcell, lz = (z - a.toz) / a.tcell;
    if (!(lx >= 0 && lz >= 0 && lx < a.tnx - 1 && lz < a.tnz - 1)) return false;
    const int ix = (int)lx, iz = (int)lz;
    const float fx = lx - ix, fz = lz - iz;
    const size_t r0 = (size_t)iz * a.tnx, r1 = r0 + a.tnx;
    const float h00 = a.th[r0 + ix], h10 = a.th[r0 + ix + 1], h01 = a.th[r1 + ix], h11 = a.th[r1 + ix + 1];
    if (smin(smin(h00, h10), smin(h01, h11)) < -1e29f) return false;
    if (fx > fz) {
        out_h = h00 + (h10 - h00) * fx + (h11 - h10) * fz;
        n = normalize(mk(-(h10 - h00) / a.tcell, 1.0f, -(h11 - h10) / a.tcell));
    } else {
        out_h = h00 + (h11 - h01) * fx + (h01 - h00) * fz;
        n = normalize(mk(-(h11 - h01) / a.tcell, 1.0f, -(h01 - h00) / a.tcell));
    }
    return true;
}

// Heightfield::surface_at (src/phys/static_world.cpp:213-218): the nearest vertex's (lround: halves away from zero)
__device__ __forceinline__ uint32_t terrain_surface(const KArgs& a, float x, float z) {
    int ix = (int)roundf((x - a.tox) / a.tcell), iz = (int)roundf((z - a.toz) / a.tcell);
    ix = ix < 0 ? 0 : (ix > a.tnx - 1 ? a.tnx - 1 : ix);
    iz = iz < 0 ? 0 : (iz > a.tnz - 1 ? a.tnz - 1 : iz);
    return a.ts[(size_t)iz * a.tnx + ix];
}

// StaticWorld::collide_point (src/phys/static_world.cpp:364-474): the terrain when near it (no road or paving), the
// boxes in index order, the vertical cylinders (the deepest contact; a cylinder leaves max_force as it was).
__device__ __forceinline__ bool collide_point(const KArgs& a, f3 p, float r, bool near_terrain, f3& normal,
                                              float& depth_out, uint32_t& surface, float& max_force) {
    bool hit = false;
    float out_depth = 0;
    max_force = 0;
    if (near_terrain) {
        float th;
        f3 n;
        if (terrain_sample(a, p.x, p.z, th, n)) {
            const float d = (th - p.y) * n.y + r;
            if (d > 0) {
                out_depth = d;
                normal = n;
                surface = terrain_surface(a, p.x, p.z);
                hit = true;
            }
        }
    }
    for (int k = 0; k < a.nbox; k++) {
        const bl_box& b = a.boxes[k];
        if (p.x < b.aabb_mn[0] - r || p.x > b.aabb_mx[0] + r || p.y < b.aabb_mn[1] - r || p.y > b.aabb_mx[1] + r ||
            p.z < b.aabb_mn[2] - r || p.z > b.aabb_mx[2] + r)
            continue;
        const f3 c0 = mk(b.rot[0], b.rot[1], b.rot[2]), c1 = mk(b.rot[3], b.rot[4], b.rot[5]),
                 c2 = mk(b.rot[6], b.rot[7], b.rot[8]);
        const f3 rel = sub(p, mk(b.center[0], b.center[1], b.center[2]));
        const f3 l = mk(dot(rel, c0), dot(rel, c1), dot(rel, c2));
        const f3 q = mk(clampf(l.x, -b.half[0], b.half[0]), clampf(l.y, -b.half[1], b.half[1]),
                        clampf(l.z, -b.half[2], b.half[2]));
        const f3 dd = sub(l, q);
        const float dist2 = dot(dd, dd);
        float depth;
        f3 nl;
        if (dist2 > 1e-12f) {
            if (dist2 >= r * r) continue;
            const float dist = fsqrt(dist2);
            depth = r - dist;
            nl = divs(dd, dist);
        } else {
            const f3 pen = sub(mk(b.half[0], b.half[1], b.half[2]), mk(ffabs(l.x), ffabs(l.y), ffabs(l.z)));
            if (pen.x <= pen.y && pen.x <= pen.z) { depth = pen.x; nl = mk(signf(l.x), 0, 0); }
            else if (pen.y <= pen.z) { depth = pen.y; nl = mk(0, signf(l.y), 0); }
            else { depth = pen.z; nl = mk(0, 0, signf(l.z)); }
            depth += r;
        }
        if (depth > out_depth) {
            out_depth = depth;
            normal = add(add(mul(c0, nl.x), mul(c1, nl.y)), mul(c2, nl.z));
            surface = b.surface;
            max_force = b.max_force;
            hit = true;
        }
    }
    for (int k = 0; k < a.ncyl; k++) {
        const bl_cylinder& c = a.cyls[k];
        const f3 rel = sub(p, mk(c.base[0], c.base[1], c.base[2]));
        if (rel.y < -r || rel.y > c.height + r) continue;
        if (ffabs(rel.x) >= c.radius + r || ffabs(rel.z) >= c.radius + r) continue;
        const float dxz = fsqrt(rel.x * rel.x + rel.z * rel.z);
        const float side = c.radius + r - dxz;
        const float top = c.height + r - rel.y;
        if (side <= 0) continue;
        float depth;
        f3 n;
        if (top < side && rel.y > c.height * 0.5f) depth = top, n = mk(0, 1, 0);
        else depth = side, n = dxz > 1e-6f ? mk(rel.x / dxz, 0, rel.z / dxz) : mk(1, 0, 0);
        if (depth > out_depth) {
            out_depth = depth;
            normal = n;
            surface = c.surface;
            hit = true;
        }
    }
    depth_out = out_depth;
    return hit;
}

// World::collide_static's terrain test (src/phys/world.cpp:812): a node within 5 cm of its body's terrain top
__device__ __forceinline__ bool near_terrain(const KArgs& a, float py, float r, uint32_t body) {
    return a.tmax && py - r <= a.tmax[body] + 0.05f;
}

struct NodeCounts {
    uint32_t contacts, clamps, nans;
    float max_v2; // top |v|^2 of the integrated nodes (integrate_nodes' max_v2)
};

// One node after its force sum: World::collide_static (src/phys/world.cpp:799-869, boxes only) then
// SoftBody::integrate_nodes (src/phys/softbody.cpp:773-791). P, V are updated in place.
__device__ __forceinline__ void node_finish(const KArgs& a, uint32_t info, float2 aux, float4& P, float4& V, f3 f,
                                            NodeCounts& cnt) {
    f3 p = mk(P.x, P.y, P.z), v = mk(V.x, V.y, V.z);
    float inv_mass = P.w;
    const float mass = V.w;
    if ((info & BL_NF_GROUND) && inv_mass > 0) {
        if (p.y < a.kill_y) {
            v = mk(0, 0, 0);
            inv_mass = 0;
        } else {
            const bool nt = near_terrain(a, p.y, aux.x, info >> 16);
            f3 n;
            float depth, max_force;
            uint32_t surf;
            if ((nt || a.nbox > 0 || a.ncyl > 0) && collide_point(a, p, aux.x, nt, n, depth, surf, max_force)) {
                cnt.contacts++;
                const ground_model& gm = k_gm[surf < BL_SURF_N ? surf : 0];
                const BodyParam& bp = a.bodies[info >> 16];
                if (max_force > 0) {
                    f3 fc = primitive_collision(f, v, mass, n, a.dt, gm, depth, aux.y, bp.push_max, bp.slop, -1.0f);
                    const float fl = length(fc);
                    if (fl > max_force) fc = mul(fc, max_force / fl);
                    f = add(f, fc);
                } else if (info & BL_NF_TYRE) {
                    // (a tyre node: its contact force is in f already, from the tyre pass)
                } else {
                    f = add(f, primitive_collision(f, v, mass, n, a.dt, gm, depth, aux.y, bp.push_max, bp.slop, bp.bounce));
                }
            }
        }
    }
    if (inv_mass > 0) {
        v = add(v, mul(f, inv_mass * a.dt));
        float v2 = dot(v, v);
        const float kmax = 400.0f;
        // (the counts as additions of the tests: an increment in each branch compiles to one store through a
        // selected address, which puts the counts in scratch memory)
        const bool clamp = v2 > kmax * kmax, nan = !clamp && !(v2 >= 0.0f);
        if (clamp) {
            v = mul(v, kmax / fsqrt(v2));
            v2 = kmax * kmax;
        } else if (nan) {
            v = mk(0, 0, 0);
            v2 = 0;
        }
        cnt.clamps += clamp ? 1u : 0u;
        cnt.nans += nan ? 1u : 0u;
        p = add(p, mul(v, a.dt));
        cnt.max_v2 = v2 > cnt.max_v2 ? v2 : cnt.max_v2;
    }
    P = make_float4(p.x, p.y, p.z, inv_mass);
    V = make_float4(v.x, v.y, v.z, mass);
}

// The force sum of one node from the chunk's incidence forces in LDS, in beam-index order; starts from
// clear_forces' f = g m (src/phys/softbody.cpp:381).
__device__ __forceinline__ f3 node_force(const KArgs& a, float mass, uint32_t b0, uint32_t b1, const float* sfx,
                                         const float* sfy, const float* sfz, const uint8_t* sok) {
    f3 f = mul(a.g, mass);
    for (uint32_t s = b0; s < b1; s++)
        if (sok[s]) f = add(f, mk(sfx[s], sfy[s], sfz[s]));
    return f;
}

// ---------------------------------------------------------------------------------------------- streamed chunk
// One chunk's substep with the incidence records streamed from memory: all of a thread's loads are issued before
// its arithmetic (PER incidences per thread).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt(const float4* p) {
    const nt_f4 v = __builtin_nontemporal_load((const nt_f4*)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t ld_nt(const uint32_t* p) { return __builtin_nontemporal_load(p); }
// an incidence's beam ends (a, b) from its own node (the chunk's first node n0 plus the offset in its flags tf) and
// the other end
__device__ __forceinline__ uint2 inc_ends(uint32_t n0, uint32_t tf, uint32_t other) {
    const uint32_t own = n0 + (tf >> TF_OWN_SHIFT);
    return (tf & TF_ISB) ? make_uint2(other, own) : make_uint2(own, other);
}

template <int T, int PER, bool FORCES_ONLY = false>
__device__ __forceinline__ void chunk_substep(const uint4 cr, const KArgs& a, int parity, float* sfx, float* sfy,
                                              float* sfz, uint8_t* sok, NodeCounts& cnt) {
    const float4* __restrict__ pin = a.pm[parity];
    const float4* __restrict__ vin = a.vm[parity];
    float4* __restrict__ pout = a.pm[parity ^ 1];
    float4* __restrict__ vout = a.vm[parity ^ 1];
    const uint32_t n0 = cr.x, n1 = cr.y, e0 = cr.z, e1 = cr.w;
    float4 A[PER], M[PER];
    uint32_t oth[PER];
    uint2 ab[PER];
    // (the incidence records stream through once a substep: loaded non-temporal, so that they do not push the
    // nodes, read by many workgroups, out of the L2)
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + threadIdx.x + k * T;
        if (e < e1) A[k] = ld_nt(a.iA + e), M[k] = ld_nt(a.iM + e), oth[k] = ld_nt(a.iOth + e);
    }
    // the node phase's own loads (a chunk holds at most T nodes: one a thread), issued with the beam phase's
    const uint32_t i = n0 + threadIdx.x;
    const bool own = i < n1;
    float4 P = make_float4(0, 0, 0, 0), V = make_float4(0, 0, 0, 0);
    uint32_t nb0 = 0, nb1 = 0, info = 0;
    float2 aux = make_float2(0, 0);
    if (own) {
        V = vin[i], nb0 = a.noff[i], nb1 = a.noff[i + 1];
        if (!FORCES_ONLY) P = pin[i], info = a.ninfo[i], aux = a.naux[i];
    }
    float4 Pa[PER], Pb[PER], Va[PER], Vb[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + threadIdx.x + k * T;
        if (e < e1) {
            ab[k] = inc_ends(n0, f2u(M[k].w), oth[k]);
            Pa[k] = pin[ab[k].x], Pb[k] = pin[ab[k].y], Va[k] = vin[ab[k].x], Vb[k] = vin[ab[k].y];
        }
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + threadIdx.x + k * T;
        if (e >= e1) continue;
        f3 fv = mk(0, 0, 0);
        float slen = 0;
        uint32_t ev;
        const bool ok = beam_eval(A[k], M[k], Pa[k], Pb[k], Va[k], Vb[k], a.bcold, fv, slen, ev);
        const uint32_t tf = f2u(M[k].w);
        if (ev & EV_CHANGED) a.iA[e] = A[k], a.iM[e] = M[k];
        if ((ev & EV_BROKE) && !(tf & TF_ISB)) atomicAdd(&a.broken[a.ninfo[ab[k].x] >> 16], 1ull);
        if (ok && !(tf & TF_ISB)) a.stress[f2u(M[k].z)] = slen;
        const uint32_t s = threadIdx.x + k * T;
        sfx[s] = fv.x, sfy[s] = fv.y, sfz[s] = fv.z;
        sok[s] = ok;
    }
    __syncthreads();
    if (own) {
        const f3 f = node_force(a, V.w, nb0 - e0, nb1 - e0, sfx, sfy, sfz, sok);
        if (FORCES_ONLY) {
            a.F[i] = make_float4(f.x, f.y, f.z, 0.0f);
        } else {
            node_finish(a, info, aux, P, V, f, cnt);
            pout[i] = P, vout[i] = V;
        }
    }
}

// The workgroup's counts onto its chunk's slot (a plain read-modify-write: one workgroup owns a chunk at a time;
// the slots are folded into the counters once a frame, so the substeps need no atomics).
template <int T>
// (pre: the slot's value, read by thread 0 ahead of time; null: read here)
__device__ __forceinline__ void flush_chunk(const KArgs& a, NodeCounts cnt, uint32_t c, uint4* sred,
                                            const uint4* pre = nullptr) {
    uint32_t c0 = cnt.contacts, c1 = cnt.clamps, c2 = cnt.nans;
    float v2 = cnt.max_v2;
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off);
        c1 += __shfl_xor(c1, off);
        c2 += __shfl_xor(c2, off);
        const float o = __shfl_xor(v2, off);
        v2 = o > v2 ? o : v2;
    }
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = make_uint4(c0, c1, c2, f2u(v2));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 t = pre ? *pre : a.cstat[c];
        for (int w = 0; w < T / 64; w++) {
            const uint4 x = sred[w];
            t.x += x.x, t.y += x.y, t.z += x.z;
            t.w = u2f(x.w) > u2f(t.w) ? x.w : t.w;
        }
        a.cstat[c] = t;
    }
}

// The hardware deals workgroup b to XCD b % 8; the host's table gives each XCD an equal share of every large body
// as one contiguous range of chunks (neighbours share an L2) and deals small bodies round the XCDs.
__device__ __forceinline__ uint32_t xcd_chunk(const KArgs& a, uint32_t b) { return a.wg_chunk ? a.wg_chunk[b] : b; }

template <int T, int PER>
__global__ __launch_bounds__(T) void k_substep(KArgs a, int parity) {
    __shared__ float sfx[T * PER], sfy[T * PER], sfz[T * PER];
    __shared__ uint8_t sok[T * PER];
    __shared__ uint4 sred[T / 64];
    NodeCounts cnt = {0, 0, 0, 0.0f};
    // (the chunk's count slot read at the start, so that the workgroup's tail does not wait for it)
    const uint32_t c = xcd_chunk(a, blockIdx.x);
    uint4 pre = make_uint4(0, 0, 0, 0);
    if (threadIdx.x == 0) pre = a.cstat[c];
    chunk_substep<T, PER>(a.wg_rng[blockIdx.x], a, parity, sfx, sfy, sfz, sok, cnt);
    flush_chunk<T>(a, cnt, c, sred, &pre);
}

// Many substeps per launch, one workgroup per body: the body's chunks one after another, a workgroup barrier
// between phases (bodies are independent without inter-body contacts).
template <int T, int PER>
__global__ __launch_bounds__(T) void k_body_persistent(KArgs a, int parity0, int substeps) {
    __shared__ float sfx[T * PER], sfy[T * PER], sfz[T * PER];
    __shared__ uint8_t sok[T * PER];
    __shared__ uint4 sred[T / 64];
    const uint32_t body = blockIdx.x;
    const uint32_t c0 = a.body_c0[body], c1 = a.body_c0[body + 1];
    NodeCounts cnt = {0, 0, 0, 0.0f};
    for (int s = 0; s < substeps; s++) {
        const int parity = parity0 ^ (s & 1);
        for (uint32_t c = c0; c < c1; c++) {
            chunk_substep<T, PER>(a.ch_rng[c], a, parity, sfx, sfy, sfz, sok, cnt);
            __syncthreads();
        }
    }
    if (c1 > c0) flush_chunk<T>(a, cnt, c0, sred);
}

// ---------------------------------------------------------------------------------------------- resident form
// Grid-wide barrier of a cooperative launch: a monotonic arrival counter (cleared before the launch). Bounded:
// after `spin_limit` polls, or once another workgroup has given up, it raises the error word and the caller ends
// its loop, so the grid always drains.
__device__ __forceinline__ bool grid_barrier(unsigned int* count, int* err, unsigned int target, uint32_t spin_limit,
                                             int* sflag) {
    __builtin_amdgcn_s_waitcnt(0); // this wave's stores of the substep acknowledged
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence(); // the workgroup's stores out to the device scope
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0;
        int bad = 0;
        while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { bad = 1; break; }
            if (++spins > spin_limit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bad = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        *sflag = bad;
    }
    __syncthreads();
    return *sflag == 0;
}

// Many substeps per launch over the whole chip: each workgroup owns one chunk for the launch, keeps its
// incidences' state (PER per thread) and its nodes (at most one per thread) in registers, and reads the other
// chunks' nodes from the previous substep's buffer after the grid barrier. Writes the state back at the end.
template <int T, int PER, int WAVES>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void k_resident(
    KArgs a, int parity0, int substeps, unsigned int* bar, int* err, uint32_t spin_limit) {
    __shared__ float sfx[T * PER], sfy[T * PER], sfz[T * PER];
    __shared__ uint8_t sok[T * PER];
    __shared__ uint4 sred[T / 64];
    __shared__ int sflag;
    const uint32_t c = xcd_chunk(a, blockIdx.x);
    const uint32_t n0 = a.ch_n0[c], n1 = a.ch_n0[c + 1];
    const uint32_t e0 = a.noff[n0], e1 = a.noff[n1];
    float4 A[PER], M[PER];
    uint2 ab[PER];
    float stress[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + threadIdx.x + k * T;
        A[k] = make_float4(0, 0, 0, 0), M[k] = make_float4(0, 0, 0, u2f(TF_SKIP)), ab[k] = make_uint2(0, 0);
        stress[k] = 0;
        if (e < e1) {
            A[k] = a.iA[e], M[k] = a.iM[e], ab[k] = inc_ends(n0, f2u(M[k].w), a.iOth[e]);
            stress[k] = (f2u(M[k].w) & TF_ISB) ? 0.0f : a.stress[f2u(M[k].z)];
        }
    }
    const uint32_t i = n0 + threadIdx.x;
    const bool has_node = i < n1;
    float4 P = make_float4(0, 0, 0, 0), V = make_float4(0, 0, 0, 0);
    uint32_t info = 0, b0 = 0, b1 = 0;
    float2 aux = make_float2(0, 0);
    if (has_node) {
        P = a.pm[parity0][i], V = a.vm[parity0][i];
        info = a.ninfo[i], aux = a.naux[i];
        b0 = a.noff[i] - e0, b1 = a.noff[i + 1] - e0;
    }
    NodeCounts cnt = {0, 0, 0, 0.0f};
    const unsigned int nb = gridDim.x;
    for (int s = 0; s < substeps; s++) {
        const int parity = parity0 ^ (s & 1);
        const float4* __restrict__ pin = a.pm[parity];
        const float4* __restrict__ vin = a.vm[parity];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const uint32_t e = e0 + threadIdx.x + k * T;
            if (e >= e1) continue;
            const float4 Pa = pin[ab[k].x], Pb = pin[ab[k].y], Va = vin[ab[k].x], Vb = vin[ab[k].y];
            f3 fv = mk(0, 0, 0);
            float slen = 0;
            uint32_t ev;
            const bool ok = beam_eval(A[k], M[k], Pa, Pb, Va, Vb, a.bcold, fv, slen, ev);
            if ((ev & EV_BROKE) && !(f2u(M[k].w) & TF_ISB)) atomicAdd(&a.broken[a.ninfo[ab[k].x] >> 16], 1ull);
            if (ok) stress[k] = slen;
            const uint32_t sl = threadIdx.x + k * T;
            sfx[sl] = fv.x, sfy[sl] = fv.y, sfz[sl] = fv.z;
            sok[sl] = ok;
        }
        __syncthreads();
        if (has_node) {
            const f3 f = node_force(a, V.w, b0, b1, sfx, sfy, sfz, sok);
            node_finish(a, info, aux, P, V, f, cnt);
            a.pm[parity ^ 1][i] = P, a.vm[parity ^ 1][i] = V;
        }
        if (!grid_barrier(bar, err, (unsigned int)(s + 1) * nb, spin_limit, &sflag)) break;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + threadIdx.x + k * T;
        if (e < e1) {
            a.iA[e] = A[k], a.iM[e] = M[k];
            if (!(f2u(M[k].w) & TF_ISB)) a.stress[f2u(M[k].z)] = stress[k];
        }
    }
    flush_chunk<T>(a, cnt, c, sred);
}


// ---------------------------------------------------------------------------------------------- halo form
// Device-coherent 16-byte load / store (two 64-bit agent-scope accesses): they bypass the XCD-private caches, so a
// node written by a workgroup on another XCD is seen without cache maintenance.
__device__ __forceinline__ float4 ld_dev(const float4* p) {
    uint64_t* q = reinterpret_cast<uint64_t*>(const_cast<float4*>(p));
    const uint64_t lo = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t hi = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(u2f((uint32_t)lo), u2f((uint32_t)(lo >> 32)), u2f((uint32_t)hi), u2f((uint32_t)(hi >> 32)));
}
__device__ __forceinline__ void st_dev(float4* p, float4 v) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    __hip_atomic_store(q, (uint64_t)f2u(v.x) | ((uint64_t)f2u(v.y) << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, (uint64_t)f2u(v.z) | ((uint64_t)f2u(v.w) << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Many substeps per launch; each workgroup owns one chunk, keeps its incidences' state in registers and its nodes
// in LDS, and per substep: waits until the chunks it reads from (and that read from it) have finished the previous
// substep (their progress words), copies its halo - the outside nodes its beams reach - into LDS, evaluates its
// incidences from LDS, integrates its nodes, publishes them and its progress. No grid-wide barrier; every wait is
// bounded (spin_limit polls, or another workgroup's error) so the grid always drains.
template <int T, int PER, int HCAP, int WAVES>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void k_halo(
    KArgs a, int parity0, int substeps, unsigned int* flags, int* err, uint32_t spin_limit) {
    __shared__ float4 hp[HCAP], hv[HCAP];
    __shared__ float sfx[T * PER], sfy[T * PER], sfz[T * PER];
    __shared__ uint8_t sok[T * PER];
    __shared__ uint4 sred[T / 64];
    __shared__ int sbad;
    const uint32_t tid = threadIdx.x;
    const uint32_t c = xcd_chunk(a, blockIdx.x);
    const uint32_t n0 = a.ch_n0[c], n1 = a.ch_n0[c + 1], nown = n1 - n0;
    const uint32_t e0 = a.noff[n0], e1 = a.noff[n1];
    const uint32_t h0 = a.halo_off[c], next = a.halo_off[c + 1] - h0;
    const uint32_t q0 = a.nbr_off[c], nnbr = a.nbr_off[c + 1] - q0;
    const uint32_t body = a.chunk_body[c];
    if (tid == 0) sbad = 0;
    float4 A[PER], M[PER];
    uint32_t loc[PER];
    float stress[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + tid + k * T;
        A[k] = make_float4(0, 0, 0, 0), M[k] = make_float4(0, 0, 0, u2f(TF_SKIP)), loc[k] = 0, stress[k] = 0;
        if (e < e1) {
            A[k] = a.iA[e], M[k] = a.iM[e], loc[k] = a.iloc[e];
            stress[k] = (f2u(M[k].w) & TF_ISB) ? 0.0f : a.stress[f2u(M[k].z)];
        }
    }
    const uint32_t i = n0 + tid;
    const bool has_node = tid < nown;
    float4 P = make_float4(0, 0, 0, 0), V = make_float4(0, 0, 0, 0);
    uint32_t info = 0, b0 = 0, b1 = 0;
    float2 aux = make_float2(0, 0);
    if (has_node) {
        P = a.pm[parity0][i], V = a.vm[parity0][i];
        info = a.ninfo[i], aux = a.naux[i];
        b0 = a.noff[i] - e0, b1 = a.noff[i + 1] - e0;
        hp[tid] = P, hv[tid] = V;
    }
    NodeCounts cnt = {0, 0, 0, 0.0f};
    for (int s = 0; s < substeps; s++) {
        const int parity = parity0 ^ (s & 1);
        const float4* __restrict__ pin = a.pm[parity];
        const float4* __restrict__ vin = a.vm[parity];
        if (s > 0) {
            for (uint32_t j = tid; j < nnbr; j += T) {
                const uint32_t nb = a.nbr[q0 + j];
                uint32_t spins = 0;
                while (__hip_atomic_load(&flags[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < (unsigned int)s) {
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > spin_limit) {
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        sbad = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads(); // (the halo is read only after every neighbour has published)
        for (uint32_t j = tid; j < next; j += T) {
            const uint32_t g = a.halo[h0 + j];
            hp[nown + j] = ld_dev(&pin[g]), hv[nown + j] = ld_dev(&vin[g]);
        }
        __syncthreads();
        if (sbad) break;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const uint32_t e = e0 + tid + k * T;
            if (e >= e1) continue;
            const uint32_t la = loc[k] & 0xffffu, lb = loc[k] >> 16;
            f3 fv = mk(0, 0, 0);
            float slen = 0;
            uint32_t ev;
            const bool ok = beam_eval(A[k], M[k], hp[la], hp[lb], hv[la], hv[lb], a.bcold, fv, slen, ev);
            if ((ev & EV_BROKE) && !(f2u(M[k].w) & TF_ISB)) atomicAdd(&a.broken[body], 1ull);
            if (ok) stress[k] = slen;
            const uint32_t sl = tid + k * T;
            sfx[sl] = fv.x, sfy[sl] = fv.y, sfz[sl] = fv.z;
            sok[sl] = ok;
        }
        __syncthreads();
        if (has_node) {
            const f3 f = node_force(a, V.w, b0, b1, sfx, sfy, sfz, sok);
            node_finish(a, info, aux, P, V, f, cnt);
            hp[tid] = P, hv[tid] = V;
            st_dev(&a.pm[parity ^ 1][i], P);
            st_dev(&a.vm[parity ^ 1][i], V);
        }
        __builtin_amdgcn_s_waitcnt(0); // the nodes' device-scope stores acknowledged before the progress word
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&flags[c], (unsigned int)(s + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = e0 + tid + k * T;
        if (e < e1) {
            a.iA[e] = A[k], a.iM[e] = M[k];
            if (!(f2u(M[k].w) & TF_ISB)) a.stress[f2u(M[k].z)] = stress[k];
        }
    }
    flush_chunk<T>(a, cnt, c, sred);
}

// ---------------------------------------------------------------------------------------------- frame kernels
__device__ __forceinline__ unsigned int rd_word(const unsigned int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The workgroup's nodes all belong to one body when its first and last do (bodies are contiguous): then a
// workgroup reduction and one atomic per word, else one atomic per node (small bodies: little contention).
template <int T>
__device__ __forceinline__ void body_reduce(const KArgs& a, uint32_t i, bool valid, float v2, float px, float py,
                                            float pz, bool want_v2, bool want_box, uint32_t* sred) {
    const uint32_t first = blockIdx.x * T;
    const uint32_t last = (first + T < a.nnodes ? first + T : a.nnodes) - 1;
    const uint32_t bfirst = a.ninfo[first] >> 16, blast = a.ninfo[last] >> 16;
    uint32_t kv = valid && want_v2 && v2 > 0 ? f2u(v2) : 0u;
    const float p[3] = {px, py, pz};
    uint32_t kmn[3], kmx[3];
    for (int k = 0; k < 3; k++) {
        const bool ok = valid && want_box && p[k] == p[k];
        kmn[k] = ok ? fkey(p[k]) : 0xffffffffu;
        kmx[k] = ok ? fkey(p[k]) : 0u;
    }
    if (bfirst != blast) {
        // several bodies in the workgroup: a wave whose nodes are one body's reduces first and makes one atomic per
        // word, any other wave one atomic per node
        const uint32_t bd = valid ? (a.ninfo[i] >> 16) : 0xffffffffu;
        const uint32_t b0 = (uint32_t)__shfl((int)bd, 0);
        if (b0 != 0xffffffffu && __all(!valid || bd == b0)) {
            for (int off = 32; off > 0; off >>= 1) {
                kv = max(kv, (uint32_t)__shfl_xor((int)kv, off));
                for (int k = 0; k < 3; k++) {
                    kmn[k] = min(kmn[k], (uint32_t)__shfl_xor((int)kmn[k], off));
                    kmx[k] = max(kmx[k], (uint32_t)__shfl_xor((int)kmx[k], off));
                }
            }
            if ((threadIdx.x & 63) == 0) {
                unsigned int* w = a.body_frame + b0 * BF_N;
                if (want_v2 && kv && kv > rd_word(&w[BF_V2])) atomicMax(&w[BF_V2], kv);
                if (want_box)
                    for (int k = 0; k < 3; k++) {
                        if (kmn[k] < rd_word(&w[BF_MN + k])) atomicMin(&w[BF_MN + k], kmn[k]);
                        if (kmx[k] > rd_word(&w[BF_MX + k])) atomicMax(&w[BF_MX + k], kmx[k]);
                    }
            }
            return;
        }
        if (!valid) return;
        unsigned int* w = a.body_frame + bd * BF_N;
        if (kv) atomicMax(&w[BF_V2], kv);
        for (int k = 0; k < 3; k++)
            if (want_box && p[k] == p[k]) atomicMin(&w[BF_MN + k], kmn[k]), atomicMax(&w[BF_MX + k], kmx[k]);
        return;
    }
    for (int off = 32; off > 0; off >>= 1) {
        kv = max(kv, (uint32_t)__shfl_xor((int)kv, off));
        for (int k = 0; k < 3; k++) {
            kmn[k] = min(kmn[k], (uint32_t)__shfl_xor((int)kmn[k], off));
            kmx[k] = max(kmx[k], (uint32_t)__shfl_xor((int)kmx[k], off));
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sred[wv * 7] = kv;
        for (int k = 0; k < 3; k++) sred[wv * 7 + 1 + k] = kmn[k], sred[wv * 7 + 4 + k] = kmx[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w2 = 1; w2 < T / 64; w2++) {
            sred[0] = max(sred[0], sred[w2 * 7]);
            for (int k = 0; k < 3; k++)
                sred[1 + k] = min(sred[1 + k], sred[w2 * 7 + 1 + k]), sred[4 + k] = max(sred[4 + k], sred[w2 * 7 + 4 + k]);
        }
        // (the words only rise (fall) within a frame: a workgroup whose value does not pass the one it reads leaves
        // the word alone - a stale read costs an atomic that changes nothing, never a missed one - so that the many
        // workgroups of a large body do not queue on one address)
        unsigned int* w = a.body_frame + bfirst * BF_N;
        if (want_v2 && sred[0] && sred[0] > rd_word(&w[BF_V2])) atomicMax(&w[BF_V2], sred[0]);
        if (want_box)
            for (int k = 0; k < 3; k++) {
                if (sred[1 + k] < rd_word(&w[BF_MN + k])) atomicMin(&w[BF_MN + k], sred[1 + k]);
                if (sred[4 + k] > rd_word(&w[BF_MX + k])) atomicMax(&w[BF_MX + k], sred[4 + k]);
            }
    }
}

// The frame's start (src/phys/world.cpp:3217-3222): every body's top |v|^2 over all its nodes (NaN left out, as
// std::max leaves it).
__global__ __launch_bounds__(256) void k_frame_begin(KArgs a, int parity) {
    __shared__ uint32_t sred[4 * 7];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < a.nnodes;
    float v2 = 0;
    if (valid) {
        const float4 V = a.vm[parity][i];
        const f3 v = mk(V.x, V.y, V.z);
        v2 = dot(v, v);
    }
    body_reduce<256>(a, i, valid, v2, 0, 0, 0, true, false, sred);
}

// The frame's end: each body's bounds (vmin / vmax over its nodes: NaN coordinates leave them unchanged).
__global__ __launch_bounds__(256) void k_frame_end(KArgs a, int parity) {
    __shared__ uint32_t sred[4 * 7];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < a.nnodes;
    float4 P = make_float4(0, 0, 0, 0);
    if (valid) P = a.pm[parity][i];
    body_reduce<256>(a, i, valid, 0, P.x, P.y, P.z, false, true, sred);
}

// Resets the per-body frame words: top speed 0, bounds keys to the empty box (1e30 / -1e30, as the reference's
// integrate_nodes starts them).
__global__ void k_frame_clear(KArgs a) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.nbodies) return;
    unsigned int* w = a.body_frame + b * BF_N;
    w[BF_V2] = 0;
    for (int k = 0; k < 3; k++) w[BF_MN + k] = fkey(1e30f), w[BF_MX + k] = fkey(-1e30f);
}

// Bounds and top speed of the current state, all bodies (for bl_get_stats).
__global__ void k_bounds(KArgs a, int parity, unsigned int* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nnodes) return;
    const float4 P = a.pm[parity][i], V = a.vm[parity][i];
    const float v2 = V.x * V.x + V.y * V.y + V.z * V.z;
    if (P.w > 0 && v2 == v2) atomicMax(&out[6], f2u(v2)); // non-negative floats order as their bits
    const float p[3] = {P.x, P.y, P.z};
    for (int k = 0; k < 3; k++) {
        if (p[k] != p[k]) continue;
        atomicMin(&out[k], fkey(p[k]));
        atomicMax(&out[3 + k], fkey(p[k]));
    }
}

// The frame's end for the chunk slots: counts into the counters, top speed into the body's frame word; slots cleared.
__global__ void k_chunk_fold(KArgs a) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nchunks) return;
    const uint4 t = a.cstat[c];
    if (t.x) atomicAdd(&a.counters[CNT_STATIC], (unsigned long long)t.x);
    if (t.y) atomicAdd(&a.counters[CNT_CLAMPS], (unsigned long long)t.y);
    if (t.z) atomicAdd(&a.counters[CNT_NANS], (unsigned long long)t.z);
    if (u2f(t.w) > 0) atomicMax(&a.body_frame[a.chunk_body[c] * BF_N + BF_V2], t.w);
    a.cstat[c] = make_uint4(0, 0, 0, 0);
}

__global__ void k_chunk_clear(KArgs a) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.nchunks) a.cstat[c] = make_uint4(0, 0, 0, 0);
}

__global__ void k_clear_words(unsigned int* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------------------------- launches
#define BL_OK_OR_HIP() (hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP)

// the last cooperative launch's status (its name goes into the error message)
thread_local hipError_t g_coop_error = hipSuccess;
const char* coop_error_name() { return hipGetErrorName(g_coop_error); }

template <int T, int PER, int WAVES>
static int launch_resident_t(const KArgs& a, hipStream_t st, int parity, int substeps, unsigned int* bar, int* err,
                             uint32_t spin_limit) {
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, st, bar, 1u);
    KArgs aa = a;
    int p = parity, s = substeps;
    void* args[] = {&aa, &p, &s, &bar, &err, &spin_limit};
    const hipError_t e = hipLaunchCooperativeKernel((const void*)k_resident<T, PER, WAVES>, dim3(a.nchunks), dim3(T), args, 0, st);
    g_coop_error = e;
    if (e != hipSuccess)
        return BL_ERR_HIP;
    return BL_OK_OR_HIP();
}

int resident_fits(int block, int per, uint32_t nchunks) {
    int dev = 0, cus = 0, nbpm = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e = hipErrorInvalidValue;
    if (block == 256 && per == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpm, k_resident<256, 4, 4>, 256, 0);
    else if (block == 256 && per == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpm, k_resident<256, 2, 6>, 256, 0);
    else if (block == 256 && per == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpm, k_resident<256, 8, 2>, 256, 0);
    else if (block == 512 && per == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpm, k_resident<512, 4, 4>, 512, 0);
    if (e != hipSuccess) return 0;
    return (uint64_t)nbpm * (uint64_t)cus >= nchunks ? 1 : 0;
}

int launch_resident(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t, unsigned int* bar,
                    int* err, uint32_t spin_limit) {
    const int per = t.incidences_per_thread;
    if (t.block == 256 && per == 4) return launch_resident_t<256, 4, 4>(a, st, parity, substeps, bar, err, spin_limit);
    if (t.block == 256 && per == 2) return launch_resident_t<256, 2, 6>(a, st, parity, substeps, bar, err, spin_limit);
    if (t.block == 256 && per == 8) return launch_resident_t<256, 8, 2>(a, st, parity, substeps, bar, err, spin_limit);
    if (t.block == 512 && per == 4) return launch_resident_t<512, 4, 4>(a, st, parity, substeps, bar, err, spin_limit);
    return BL_ERR_ARG;
}

// The launch of one form; every (block, incidences per thread) pair the host may chunk for has its own instance
// (the chunks hold at most block x per incidences and block nodes), any other pair is refused.
int launch_substeps(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t) {
    const int per = t.incidences_per_thread, T = t.block;
    if (t.mode == BL_MODE_PERSISTENT) {
        if (T == 1024 && per == 4) hipLaunchKernelGGL((k_body_persistent<1024, 4>), dim3(a.nbodies), dim3(1024), 0, st, a, parity, substeps);
        else if (T == 256 && per == 4) hipLaunchKernelGGL((k_body_persistent<256, 4>), dim3(a.nbodies), dim3(256), 0, st, a, parity, substeps);
        else return BL_ERR_ARG;
        return BL_OK_OR_HIP();
    }
    for (int s = 0; s < substeps; s++) {
        const int p = parity ^ (s & 1);
        if (T == 256 && per == 1) hipLaunchKernelGGL((k_substep<256, 1>), dim3(a.nchunks), dim3(256), 0, st, a, p);
        else if (T == 256 && per == 2) hipLaunchKernelGGL((k_substep<256, 2>), dim3(a.nchunks), dim3(256), 0, st, a, p);
        else if (T == 256 && per == 4) hipLaunchKernelGGL((k_substep<256, 4>), dim3(a.nchunks), dim3(256), 0, st, a, p);
        else if (T == 256 && per == 8) hipLaunchKernelGGL((k_substep<256, 8>), dim3(a.nchunks), dim3(256), 0, st, a, p);
        else if (T == 512 && per == 2) hipLaunchKernelGGL((k_substep<512, 2>), dim3(a.nchunks), dim3(512), 0, st, a, p);
        else if (T == 512 && per == 4) hipLaunchKernelGGL((k_substep<512, 4>), dim3(a.nchunks), dim3(512), 0, st, a, p);
        else return BL_ERR_ARG;
    }
    return BL_OK_OR_HIP();
}

int launch_bounds(const KArgs& a, hipStream_t st, int parity, unsigned int* out) {
    if (a.nnodes == 0) return BL_OK;
    hipLaunchKernelGGL(k_bounds, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity, out);
    return BL_OK_OR_HIP();
}

int launch_frame_begin(const KArgs& a, hipStream_t st, int parity) {
    hipLaunchKernelGGL(k_frame_clear, dim3((a.nbodies + 63) / 64), dim3(64), 0, st, a);
    if (a.nnodes) hipLaunchKernelGGL(k_frame_begin, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity);
    return BL_OK_OR_HIP();
}

int launch_frame_end(const KArgs& a, hipStream_t st, int parity) {
    if (a.nnodes == 0) return BL_OK;
    hipLaunchKernelGGL(k_frame_end, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity);
    if (a.nchunks) hipLaunchKernelGGL(k_chunk_fold, dim3((a.nchunks + 255) / 256), dim3(256), 0, st, a);
    return BL_OK_OR_HIP();
}

int launch_box(const KArgs& a, hipStream_t st, int parity) {
    hipLaunchKernelGGL(k_frame_clear, dim3((a.nbodies + 63) / 64), dim3(64), 0, st, a);
    if (a.nnodes) hipLaunchKernelGGL(k_frame_end, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity);
    return BL_OK_OR_HIP();
}

int launch_chunk_clear(const KArgs& a, hipStream_t st) {
    if (a.nchunks) hipLaunchKernelGGL(k_chunk_clear, dim3((a.nchunks + 255) / 256), dim3(256), 0, st, a);
    return BL_OK_OR_HIP();
}


template <int T, int PER, int HCAP, int WAVES>
static int launch_halo_t(const KArgs& a, hipStream_t st, int parity, int substeps, unsigned int* flags, int* err,
                         uint32_t spin_limit) {
    hipLaunchKernelGGL(k_clear_words, dim3((a.nchunks + 255) / 256), dim3(256), 0, st, flags, a.nchunks);
    KArgs aa = a;
    int p = parity, s = substeps;
    void* args[] = {&aa, &p, &s, &flags, &err, &spin_limit};
    const hipError_t e = hipLaunchCooperativeKernel((const void*)k_halo<T, PER, HCAP, WAVES>, dim3(a.nchunks), dim3(T), args, 0, st);
    g_coop_error = e;
    if (e != hipSuccess)
        return BL_ERR_HIP;
    return BL_OK_OR_HIP();
}

int halo_capacity(int block, int per) { return block == 256 && per == 8 ? 1024 : 512; }

int halo_fits(int block, int per, uint32_t nchunks) {
    int dev = 0, cus = 0, nbpm = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e = hipErrorInvalidValue;
    if (block == 256 && per == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpm, k_halo<256, 8, 1024, 2>, 256, 0);
    else if (block == 256 && per == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpm, k_halo<256, 4, 512, 4>, 256, 0);
    if (e != hipSuccess) return 0;
    return (uint64_t)nbpm * (uint64_t)cus >= nchunks ? 1 : 0;
}

int launch_halo(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t, unsigned int* flags,
                int* err, uint32_t spin_limit) {
    const int per = t.incidences_per_thread;
    if (t.block == 256 && per == 8) return launch_halo_t<256, 8, 1024, 2>(a, st, parity, substeps, flags, err, spin_limit);
    if (t.block == 256 && per == 4) return launch_halo_t<256, 4, 512, 4>(a, st, parity, substeps, flags, err, spin_limit);
    return BL_ERR_ARG;
}


// ---------------------------------------------------------------------------------------------- split substep
// The force phase alone (the inter-body contacts come between it and the integration): F = g m + beams.
template <int T, int PER>
__global__ __launch_bounds__(T) void k_forces(KArgs a, int parity) {
    __shared__ float sfx[T * PER], sfy[T * PER], sfz[T * PER];
    __shared__ uint8_t sok[T * PER];
    NodeCounts cnt = {0, 0, 0, 0.0f};
    chunk_substep<T, PER, true>(a.fwg_rng[blockIdx.x], a, parity, sfx, sfy, sfz, sok, cnt);
}

// The static contacts and the integration from F, one thread per node (a workgroup's nodes may belong to several
// bodies): the counts into the workgroup's slot of `wstat` (folded at the frame's end), each body's top speed
// straight into its frame word, so that it is current at every pair rebuild.
__global__ __launch_bounds__(256) void k_integrate(KArgs a, int parity, uint4* wstat) {
    __shared__ uint32_t sred[4 * 7];
    __shared__ uint4 scnt[4];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < a.nnodes;
    NodeCounts cnt = {0, 0, 0, 0.0f};
    uint4 pre = make_uint4(0, 0, 0, 0); // (the workgroup's slot, read at the start)
    if (threadIdx.x == 0) pre = wstat[blockIdx.x];
    if (valid) {
        float4 P = a.pm[parity][i], V = a.vm[parity][i];
        const float4 F = a.F[i];
        node_finish(a, a.ninfo[i], a.naux[i], P, V, mk(F.x, F.y, F.z), cnt);
        a.pm[parity ^ 1][i] = P, a.vm[parity ^ 1][i] = V;
    }
    uint32_t c0 = cnt.contacts, c1 = cnt.clamps, c2 = cnt.nans;
    for (int off = 32; off > 0; off >>= 1) c0 += __shfl_xor(c0, off), c1 += __shfl_xor(c1, off), c2 += __shfl_xor(c2, off);
    if ((threadIdx.x & 63) == 0) scnt[threadIdx.x >> 6] = make_uint4(c0, c1, c2, 0);
    body_reduce<256>(a, i, valid, cnt.max_v2, 0, 0, 0, true, false, sred);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 t = pre;
        for (int w = 0; w < 4; w++) t.x += scnt[w].x, t.y += scnt[w].y, t.z += scnt[w].z;
        wstat[blockIdx.x] = t;
    }
}

__global__ void k_wstat_fold(KArgs a, uint4* wstat, uint32_t n) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const uint4 t = wstat[w];
    if (t.x) atomicAdd(&a.counters[CNT_STATIC], (unsigned long long)t.x);
    if (t.y) atomicAdd(&a.counters[CNT_CLAMPS], (unsigned long long)t.y);
    if (t.z) atomicAdd(&a.counters[CNT_NANS], (unsigned long long)t.z);
    wstat[w] = make_uint4(0, 0, 0, 0);
}

// The force phase and the contacts' narrow phase in one launch (the narrow phase reads the positions and the
// candidate lists only, and writes its own hit buffers). The narrow phase's blocks come first, so that their chains
// of dependent reads start with the launch, under the force chunks; padded to a multiple of 8, so that the force
// chunks keep the XCDs their table gives them.
template <int PER>
__global__ __launch_bounds__(256) void k_forces_narrow(KArgs a, int parity, PArgs p, uint32_t nblk) {
    __shared__ float sfx[256 * PER], sfy[256 * PER], sfz[256 * PER];
    __shared__ uint8_t sok[256 * PER];
    const uint32_t nb8 = (nblk + 7) & ~7u;
    if (blockIdx.x >= nb8) {
        NodeCounts cnt = {0, 0, 0, 0.0f};
        chunk_substep<256, PER, true>(a.fwg_rng[blockIdx.x - nb8], a, parity, sfx, sfy, sfz, sok, cnt);
    } else if (blockIdx.x < nblk) {
        narrow_block(a, p, parity, blockIdx.x);
    }
}

int launch_forces_narrow(const KArgs& a, hipStream_t st, int parity, const bl_tuning& t, const PArgs& p, uint32_t nblk) {
    const int per = t.incidences_per_thread;
    const dim3 g(a.nfchunks + ((nblk + 7) & ~7u));
    if (t.block != 256) return BL_ERR_ARG;
    if (per == 1) hipLaunchKernelGGL((k_forces_narrow<1>), g, dim3(256), 0, st, a, parity, p, nblk);
    else if (per == 2) hipLaunchKernelGGL((k_forces_narrow<2>), g, dim3(256), 0, st, a, parity, p, nblk);
    else if (per == 4) hipLaunchKernelGGL((k_forces_narrow<4>), g, dim3(256), 0, st, a, parity, p, nblk);
    else if (per == 8) hipLaunchKernelGGL((k_forces_narrow<8>), g, dim3(256), 0, st, a, parity, p, nblk);
    else return BL_ERR_ARG;
    return BL_OK_OR_HIP();
}

int launch_forces(const KArgs& a, hipStream_t st, int parity, const bl_tuning& t) {
    const int per = t.incidences_per_thread, T = t.block;
    if (a.nfchunks == 0) return BL_OK;
    if (T == 256 && per == 1) hipLaunchKernelGGL((k_forces<256, 1>), dim3(a.nfchunks), dim3(256), 0, st, a, parity);
    else if (T == 256 && per == 2) hipLaunchKernelGGL((k_forces<256, 2>), dim3(a.nfchunks), dim3(256), 0, st, a, parity);
    else if (T == 256 && per == 4) hipLaunchKernelGGL((k_forces<256, 4>), dim3(a.nfchunks), dim3(256), 0, st, a, parity);
    else if (T == 256 && per == 8) hipLaunchKernelGGL((k_forces<256, 8>), dim3(a.nfchunks), dim3(256), 0, st, a, parity);
    else if (T == 512 && per == 2) hipLaunchKernelGGL((k_forces<512, 2>), dim3(a.nfchunks), dim3(512), 0, st, a, parity);
    else if (T == 512 && per == 4) hipLaunchKernelGGL((k_forces<512, 4>), dim3(a.nfchunks), dim3(512), 0, st, a, parity);
    else if (T == 1024 && per == 4) hipLaunchKernelGGL((k_forces<1024, 4>), dim3(a.nfchunks), dim3(1024), 0, st, a, parity);
    else return BL_ERR_ARG;
    return BL_OK_OR_HIP();
}

int launch_integrate(const KArgs& a, hipStream_t st, int parity, uint4* wstat) {
    if (a.nnodes) hipLaunchKernelGGL(k_integrate, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity, wstat);
    return BL_OK_OR_HIP();
}

// ---------------------------------------------------------------------------------------------- shocks, slides
// SoftBody::compute_shock_forces (src/phys/softbody.cpp:472-572), one shock. The beam's L, strength and flags are
// its a-end incidence's; a break marks both incidences.
__device__ __forceinline__ void shock_prog(const bl_shock& s, bool out, float diff, float L, float& kk, float& dd) {
    if (out) {
        kk = s.spring_out;
        dd = s.damp_out;
        const float fr = s.long_bound != 0 ? smin((diff / (s.long_bound * L)) * (diff / (s.long_bound * L)), 1.0f) : 1.0f;
        kk += s.prog_spring_out * kk * fr;
        dd += s.prog_damp_out * dd * fr;
    } else {
        kk = s.spring_in;
        dd = s.damp_in;
        const float fr = s.short_bound != 0 ? smin((diff / (s.short_bound * L)) * (diff / (s.short_bound * L)), 1.0f) : 1.0f;
        kk += s.prog_spring_in * kk * fr;
        dd += s.prog_damp_in * dd * fr;
    }
}

__device__ void shock_eval(const KArgs& a, const bl_shock& s, const uint4 x, int parity, uint32_t body) {
    float4 M = a.iM[x.x];
    const uint32_t tf = f2u(M.w);
    if (tf & TF_BROKEN) return;
    const float4 A = a.iA[x.x];
    const float4 Pa = a.pm[parity][x.z], Pb = a.pm[parity][x.w], Va = a.vm[parity][x.z], Vb = a.vm[parity][x.w];
    const f3 dis = sub(mk(Pa.x, Pa.y, Pa.z), mk(Pb.x, Pb.y, Pb.z));
    const float len2 = dot(dis, dis);
    if (len2 < 1e-12f) return;
    const float inv_len = 1.0f / fsqrt(len2);
    const float len = len2 * inv_len;
    const float L = A.x;
    const float diff = len - L;
    const float v = dot(sub(mk(Va.x, Va.y, Va.z), mk(Vb.x, Vb.y, Vb.z)), dis) * inv_len;
    float k, d;
    const float lb = s.long_bound * L, sb = s.short_bound * L;
    if (s.type == 1) {
        k = s.spring;
        d = s.damp;
        float r = 0;
        if (diff > lb) r = diff - lb;
        else if (diff < -sb) r = -diff - sb;
        if (r != 0) {
            r = smin(r, 1.0f);
            k += (s.bound_spring - k) * r;
            d += (s.bound_damp - d) * r;
        }
    } else if (s.type == 2) {
        shock_prog(s, v > 0, diff, L, k, d);
        if (s.soft_bump) {
            const float Lp = 0.8f * L, lpre = s.long_bound * Lp, spre = -s.short_bound * Lp;
            if (diff > lpre) {
                shock_prog(s, true, diff, L, k, d);
                const float q = (diff - lpre) * 5.0f / (s.long_bound * L);
                const float f2 = s.long_bound != 0 ? smin(q * q, 1.0f) : 1.0f;
                k += (k + 100) * s.prog_spring_out * f2;
                d += (d + 100) * s.prog_damp_out * f2;
                if (v < 0) k = s.spring_in, d = s.damp_in;
            } else if (diff < spre) {
                shock_prog(s, false, diff, L, k, d);
                const float q = (diff - spre) * 5.0f / (s.short_bound * L);
                const float f2 = s.short_bound != 0 ? smin(q * q, 1.0f) : 1.0f;
                k += (k + 100) * s.prog_spring_out * f2;
                d += (d + 100) * s.prog_damp_out * f2;
                if (v > 0) k = s.spring_out, d = s.damp_out;
            }
            if (diff > lb || diff < -sb) {
                k = smax(k, s.bound_spring);
                d = smax(d, s.bound_damp);
            }
        } else if (diff > lb || diff < -sb) {
            k = s.bound_spring;
            d = s.bound_damp;
        }
    } else {
        k = s.spring_in;
        d = s.damp_in;
        if (diff > lb) {
            const float r = smin(diff - lb, 1.0f);
            k += (s.bound_spring - k) * r;
            d += (s.bound_damp - d) * r;
        } else if (diff < -sb) {
            const float r = smin(-diff - sb, 1.0f);
            k += (s.bound_spring - k) * r;
            d += (s.bound_damp - d) * r;
        } else if (v > 0) {
            const float w = clampf(ffabs(v), 0.15f, 20.0f);
            k = s.spring_out;
            d = (s.damp_out * s.damp_out_slow * smin(w, s.split_out) + s.damp_out * s.damp_out_fast * smax(0.0f, w - s.split_out)) / w;
        } else if (v < 0) {
            const float w = clampf(ffabs(v), 0.15f, 20.0f);
            k = s.spring_in;
            d = (s.damp_in * s.damp_in_slow * smin(w, s.split_in) + s.damp_in * s.damp_in_fast * smax(0.0f, w - s.split_in)) / w;
        }
    }
    const float slen = -k * diff - d * v;
    if ((tf & TF_BRK) && ffabs(slen) > A.w && !(tf & TF_NOBREAK)) {
        M.w = u2f(tf | TF_BROKEN);
        a.iM[x.x] = M;
        float4 Mb = a.iM[x.y];
        Mb.w = u2f(f2u(Mb.w) | TF_BROKEN);
        a.iM[x.y] = Mb;
        atomicAdd(&a.broken[body], 1ull);
        return;
    }
    a.stress[f2u(M.z)] = slen;
    const f3 fv = mul(dis, slen * inv_len);
    float4 Fa = a.F[x.z], Fb = a.F[x.w];
    Fa = make_float4(Fa.x + fv.x, Fa.y + fv.y, Fa.z + fv.z, Fa.w);
    a.F[x.z] = Fa;
    Fb = a.F[x.w]; // (the same node at both ends reads the force just written, as the reference's f[] does)
    Fb = make_float4(Fb.x - fv.x, Fb.y - fv.y, Fb.z - fv.z, Fb.w);
    a.F[x.w] = Fb;
}

// SoftBody::compute_slide_forces (src/phys/softbody.cpp:574-623), one slide node
__device__ __forceinline__ float slide_seg_d2(const float4* P, const uint32_t* rail, int i, f3 p, float& t) {
    const f3 a = ld3(P[rail[i]]), b = ld3(P[rail[i + 1]]), ab = sub(b, a);
    const float l2 = dot(ab, ab);
    t = l2 > 1e-12f ? clampf(dot(sub(p, a), ab) / l2, 0.0f, 1.0f) : 0.0f;
    const f3 q = sub(add(a, mul(ab, t)), p);
    return dot(q, q);
}

__device__ void slide_eval(const KArgs& a, bl_slide& s, const uint32_t* rail_all, int parity) {
    if (s.broken || s.nrail < 2) return;
    const float4* P = a.pm[parity];
    const uint32_t* rail = rail_all + s.rail0;
    const f3 p = ld3(P[s.node]);
    const int nseg = (int)s.nrail - 1;
    float t, best_t = 0;
    if (s.seg < 0) {
        float best = 1e30f;
        for (int i = 0; i < nseg; i++) {
            const float dd = slide_seg_d2(P, rail, i, p, t);
            if (dd < best) best = dd, s.seg = i, best_t = t;
        }
        if (s.seg < 0) return; // (every distance not a number: no nearest segment - the reference would read rail[-1])
    }
    float best = slide_seg_d2(P, rail, s.seg, p, best_t);
    int bi = s.seg;
    if (s.seg > 0) {
        const float dd = slide_seg_d2(P, rail, s.seg - 1, p, t);
        if (dd < best) best = dd, bi = s.seg - 1, best_t = t;
    }
    if (s.seg + 1 < nseg) {
        const float dd = slide_seg_d2(P, rail, s.seg + 1, p, t);
        if (dd < best) best = dd, bi = s.seg + 1, best_t = t;
    }
    s.seg = bi;
    const uint32_t na = rail[bi], nb = rail[bi + 1];
    const f3 pa = ld3(P[na]);
    const f3 ideal = add(pa, mul(sub(ld3(P[nb]), pa), best_t));
    const f3 d = sub(ideal, p);
    const float dist = length(d);
    const float stretch = dist - s.threshold;
    if (stretch <= 0 || dist < 1e-9f) return;
    const f3 F = mul(d, s.k * stretch / dist);
    if (s.break_force > 0 && length(F) > s.break_force) {
        s.broken = 1;
        return;
    }
    float4 f = a.F[s.node];
    a.F[s.node] = make_float4(f.x + F.x, f.y + F.y, f.z + F.z, f.w);
    const f3 Fa = mul(F, 1.0f - best_t), Fb = mul(F, best_t);
    f = a.F[na];
    a.F[na] = make_float4(f.x - Fa.x, f.y - Fa.y, f.z - Fa.z, f.w);
    f = a.F[nb];
    a.F[nb] = make_float4(f.x - Fb.x, f.y - Fb.y, f.z - Fb.z, f.w);
}

// quaternions (x, y, z, w) as src/core/math.h:131-222 computes them (sin, cos, atan2 as glibc's)
struct qd {
    float x, y, z, w;
};
__device__ __forceinline__ qd qmk(float x, float y, float z, float w) { return qd{x, y, z, w}; }
__device__ __forceinline__ qd qmul(qd a, qd b) {
    return qmk(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
               a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
__device__ __forceinline__ qd qscale(qd a, float s) { return qmk(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ qd qconj(qd a) { return qmk(-a.x, -a.y, -a.z, a.w); }
__device__ __forceinline__ qd qnormalize(qd a) {
    const float l = fsqrt(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
    return l > 1e-12f ? qscale(a, 1.0f / l) : qmk(0, 0, 0, 1);
}
__device__ __forceinline__ f3 qrotate(qd q, f3 p) {
    const f3 u = mk(q.x, q.y, q.z);
    const f3 t = mul(cross(u, p), 2.0f);
    return add(add(p, mul(t, q.w)), cross(u, t));
}
__device__ __forceinline__ qd qaxis_angle(f3 axis, float angle) {
    const float h = angle * 0.5f, sn = bl_sinf(h);
    const f3 a = normalize(axis);
    return qmk(a.x * sn, a.y * sn, a.z * sn, bl_cosf(h));
}
__device__ __forceinline__ f3 qlog(qd qin) {
    const qd q = qin.w < 0 ? qscale(qin, -1.0f) : qin;
    const f3 v = mk(q.x, q.y, q.z);
    const float sn = length(v);
    if (sn < 1e-8f) return mul(v, 2.0f);
    const float angle = 2.0f * bl_atan2f(sn, q.w);
    return mul(v, angle / sn);
}
__device__ __forceinline__ qd ld_q(const float* f) { return qmk(f[0], f[1], f[2], f[3]); }
__device__ __forceinline__ qd ld_q(float4 f) { return qmk(f.x, f.y, f.z, f.w); }

__device__ __forceinline__ void add_force(const KArgs& a, uint32_t n, f3 x) {
    const float4 f = a.F[n];
    a.F[n] = make_float4(f.x + x.x, f.y + x.y, f.z + x.z, f.w);
}
__device__ __forceinline__ void sub_force(const KArgs& a, uint32_t n, f3 x) {
    const float4 f = a.F[n];
    a.F[n] = make_float4(f.x - x.x, f.y - x.y, f.z - x.z, f.w);
}

// SoftBody::compute_joint_forces (src/phys/softbody.cpp:625-679), one joint; its frames' torques in x.ftq
__device__ void joint_eval(const KArgs& a, const XArgs& x, bl_joint& j, int parity, bool deform, bool brk) {
    if (j.broken) return;
    const uint32_t pf = j.parent_frame, pnode = x.fnode[pf];
    const float4 Pp = a.pm[parity][pnode], Pc = a.pm[parity][j.child_node];
    const float4 Vp = a.vm[parity][pnode], Vc = a.vm[parity][j.child_node];
    const qd pq = ld_q(x.fq[pf]);
    const float4 pw4 = x.fw[pf];
    const f3 pw = mk(pw4.x, pw4.y, pw4.z);
    f3 r = qrotate(pq, mk(j.local_offset[0], j.local_offset[1], j.local_offset[2]));
    f3 err = sub(mk(Pc.x, Pc.y, Pc.z), add(mk(Pp.x, Pp.y, Pp.z), r));
    if (j.break_force > 0 && deform) {
        const float ylen = j.break_force / j.k_lin;
        const float el = length(err);
        if (el > ylen) {
            const float excess = el - ylen;
            const f3 dir = divs(err, el);
            const f3 lo = add(mk(j.local_offset[0], j.local_offset[1], j.local_offset[2]), qrotate(qconj(pq), mul(dir, excess)));
            j.local_offset[0] = lo.x, j.local_offset[1] = lo.y, j.local_offset[2] = lo.z;
            r = qrotate(pq, lo);
            err = mul(dir, ylen);
            j.plastic_lin += excess;
        }
    }
    const f3 vrel = sub(mk(Vc.x, Vc.y, Vc.z), add(mk(Vp.x, Vp.y, Vp.z), cross(pw, r)));
    const f3 F = sub(mul(err, -j.k_lin), mul(vrel, j.d_lin));
    add_force(a, j.child_node, F);
    sub_force(a, pnode, F);
    float4 pt = x.ftq[pf];
    const f3 tq = cross(r, mk(-F.x, -F.y, -F.z));
    pt = make_float4(pt.x + tq.x, pt.y + tq.y, pt.z + tq.z, 0.0f);
    float s = j.break_force > 0 ? length(err) * j.k_lin / j.break_force : 0.0f;
    if (j.child_frame >= 0) {
        const uint32_t cf = (uint32_t)j.child_frame;
        qd qe = qmul(ld_q(x.fq[cf]), qconj(qmul(pq, ld_q(j.rel_rot))));
        f3 theta = qlog(qe);
        const float ang = length(theta);
        float yield = j.yield_angle > 0 ? j.yield_angle : 1e9f;
        if (j.break_torque > 0) yield = smin(yield, j.break_torque / j.k_ang);
        if (ang > yield && deform) {
            const float excess = ang - yield;
            const f3 axis = divs(theta, ang);
            const qd bend = qaxis_angle(axis, excess);
            const qd rr = qnormalize(qmul(qmul(qmul(qconj(pq), bend), pq), ld_q(j.rel_rot)));
            j.rel_rot[0] = rr.x, j.rel_rot[1] = rr.y, j.rel_rot[2] = rr.z, j.rel_rot[3] = rr.w;
            theta = mul(axis, yield);
            j.plastic_ang += excess;
        }
        const float4 cw4 = x.fw[cf];
        const f3 T = sub(mul(theta, -j.k_ang), mul(sub(mk(cw4.x, cw4.y, cw4.z), pw), j.d_ang));
        // (the child's torque first, then the parent's: a joint between a frame and itself sums both in this order)
        x.ftq[pf] = pt;
        float4 ct = x.ftq[cf];
        x.ftq[cf] = make_float4(ct.x + T.x, ct.y + T.y, ct.z + T.z, 0.0f);
        pt = x.ftq[pf];
        pt = make_float4(pt.x - T.x, pt.y - T.y, pt.z - T.z, 0.0f);
        if (j.break_torque > 0) s = smax(s, length(theta) * j.k_ang / j.break_torque);
    }
    x.ftq[pf] = pt;
    j.stress = s;
    const bool snap = (j.break_force > 0 && j.plastic_lin > j.max_stretch) || (j.break_torque > 0 && j.plastic_ang > j.max_bend);
    if (brk && snap) j.broken = 1;
}

// SoftBody::compute_wheel_forces (src/phys/softbody.cpp:689-767), one node wheel, first = true (the controllers'
// substep): the drive and brake torques as tangential tread forces, the tread speed, the arm's reaction torque
__device__ void wheel_eval(const KArgs& a, const XArgs& x, bl_wheel& w, int parity, float& speed_sum, float& spin_sum,
                           int& nprop) {
    if (w.detached || w.n_nodes == 0) return;
    const float dt = a.dt;
    w.drive_torque = w.torque;
    if (w.brake > 0) {
        const float stop = -w.avg_speed * w.radius * w.mass / dt - w.last_retorque;
        w.torque += w.speed > 0 ? clampf(stop, -w.brake, 0.0f) : clampf(stop, 0.0f, w.brake);
    }
    const float4 H0 = a.pm[parity][w.axle0], H1 = a.pm[parity][w.axle1];
    const float4 U0 = a.vm[parity][w.axle0], U1 = a.vm[parity][w.axle1];
    const f3 axis = normalize(sub(mk(H1.x, H1.y, H1.z), mk(H0.x, H0.y, H0.z)));
    const int nn = (int)w.n_nodes;
    const float per = w.torque / (float)nn;
    float expected = w.speed;
    float speed = 0;
    for (int j = 0; j < nn; j++) {
        const float4 hp = (j & 1) ? H1 : H0, hv = (j & 1) ? U1 : U0;
        const uint32_t ni = x.wn[w.node0 + j];
        const float4 P = a.pm[parity][ni], V = a.vm[parity][ni];
        f3 rad = sub(mk(P.x, P.y, P.z), mk(hp.x, hp.y, hp.z));
        const float rl2 = dot(rad, rad);
        if (rl2 < 1e-8f) continue;
        const float inv = 1.0f / fsqrt(rl2);
        if (w.propulsed == 2) rad = mk(-rad.x, -rad.y, -rad.z);
        const f3 dir = mul(cross(axis, rad), inv);
        add_force(a, ni, mul(dir, per * inv));
        speed += dot(sub(mk(V.x, V.y, V.z), mk(hv.x, hv.y, hv.z)), dir);
    }
    w.speed = speed / (float)nn;
    w.avg_speed = 0.99f * w.avg_speed + 0.1f * w.speed;
    if (w.propulsed == 1) {
        speed_sum += w.speed;
        spin_sum += w.speed / smax(0.05f, w.radius);
        nprop++;
    }
    if (w.mass > 0 && w.radius > 0) {
        expected += (w.last_torque / w.radius) / w.mass * dt;
        w.last_retorque = w.mass * (w.speed - expected) / dt;
    }
    if (w.arm >= 0 && w.near_attach >= 0 && ffabs(w.torque) > 0.01f) {
        const float4 PA = a.pm[parity][w.arm], PN = a.pm[parity][w.near_attach];
        const f3 rr = sub(mk(PA.x, PA.y, PA.z), mk(PN.x, PN.y, PN.z));
        f3 r = sub(rr, mul(axis, dot(rr, axis)));
        const float off = length(sub(rr, r)), rl = length(r);
        if (rl > 0.01f && 2 * off < rl) {
            r = divs(r, rl);
            const f3 cf = mul(mul(cross(axis, r), 0.5f * w.torque / rl), 1.0f - 2.0f * off / rl);
            sub_force(a, (uint32_t)w.arm, cf);
            add_force(a, (uint32_t)w.near_attach, cf);
        }
    }
    w.last_torque = w.torque;
    w.torque = 0;
}

// a thread per body with shocks, joints, wheels or slides: its frames' torques cleared (clear_forces), its shocks,
// joints, wheels (after its controller's inputs) and slides, in order (World::simulate_island's base_forces,
// src/phys/world.cpp:3255-3273)
__global__ void k_extras(KArgs a, XArgs x, int parity) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= x.nbody) return;
    const uint4 r = x.body[3 * i], r2 = x.body[3 * i + 1], r3 = x.body[3 * i + 2];
    for (uint32_t f = r2.z; f < r2.w; f++) x.ftq[f] = make_float4(0, 0, 0, 0);
    for (uint32_t q = r.x; q < r.y; q++) shock_eval(a, x.sh[q], x.shx[q], parity, a.ninfo[x.shx[q].z] >> 16);
    if (r2.x < r2.y) {
        // (the body's allow_deform and allow_break: its beams' flags carry them; a body with joints has a frame node)
        const uint32_t body = a.ninfo[x.fnode[x.jt[r2.x].parent_frame]] >> 16;
        const BodyParam& bp = a.bodies[body];
        for (uint32_t q = r2.x; q < r2.y; q++) {
            bl_joint j = x.jt[q];
            joint_eval(a, x, j, parity, bp.deform != 0, bp.brk != 0);
            x.jt[q] = j;
        }
    }
    if (r3.x < r3.y) {
        float speed_sum = 0, spin_sum = 0;
        int nprop = 0;
        for (uint32_t q = r3.x; q < r3.y; q++) {
            bl_wheel w = x.wh[q];
            const float4 c = x.ctl[q];
            if (c.z != 0) w.torque = c.x, w.brake = c.y;
            wheel_eval(a, x, w, parity, speed_sum, spin_sum, nprop);
            x.wh[q] = w;
        }
        if (nprop > 0) x.wsp[i] = make_float2(speed_sum / (float)nprop, spin_sum / (float)nprop);
    }
    for (uint32_t q = r.z; q < r.w; q++) {
        bl_slide s = x.sl[q];
        slide_eval(a, s, x.rail, parity);
        x.sl[q] = s;
    }
}

int launch_extras(const KArgs& a, const XArgs& x, hipStream_t st, int parity) {
    if (x.nbody) hipLaunchKernelGGL(k_extras, dim3((x.nbody + 63) / 64), dim3(64), 0, st, a, x, parity);
    return BL_OK_OR_HIP();
}

// World::collide_static's tyre nodes (src/phys/world.cpp:799-892): a workgroup per body with tyre nodes. Each tyre
// node's static contact by tyre_contact (a yielding obstacle's contact is the integration's), its normal force into
// F, its need, grip and slide force kept (the nodes are independent: a thread each); then each wheel sums its
// touching tyre nodes' need and grip in node order (a thread per wheel: the order the float sums need); then each
// node gets the force that would hold it when its wheel's patch holds (summed need within summed grip), else its
// slide force. The integration counts these contacts and leaves their forces as they are. (Global scratch across
// the workgroup's barriers: its waves share one CU's L1.)
__global__ __launch_bounds__(256) void k_tyres(KArgs a, XArgs x, int parity) {
    const uint32_t i = blockIdx.x;
    const uint4 r3 = x.body[3 * i + 2];
    if (r3.z >= r3.w) return;
    for (uint32_t q = r3.z + threadIdx.x; q < r3.w; q += blockDim.x) {
        const uint2 e = x.ty[q];
        x.tslide[q] = make_float4(0, 0, 0, 0);
        const uint32_t info = a.ninfo[e.x];
        const float4 P = a.pm[parity][e.x], V = a.vm[parity][e.x];
        if (!(info & BL_NF_GROUND) || !(P.w > 0) || P.y < a.kill_y) continue;
        const float2 aux = a.naux[e.x];
        const bool nt = near_terrain(a, P.y, aux.x, info >> 16);
        if (!nt && a.nbox <= 0 && a.ncyl <= 0) continue;
        f3 n;
        float depth, max_force;
        uint32_t surf;
        if (!collide_point(a, mk(P.x, P.y, P.z), aux.x, nt, n, depth, surf, max_force) || max_force > 0) continue;
        const ground_model& gm = k_gm[surf < BL_SURF_N ? surf : 0];
        const float4 F = a.F[e.x];
        const tyre_hit t = tyre_contact(mk(F.x, F.y, F.z), mk(V.x, V.y, V.z), V.w, n, a.dt, gm, depth, aux.y * x.tyre_grip);
        a.F[e.x] = make_float4(F.x + t.normal_force.x, F.y + t.normal_force.y, F.z + t.normal_force.z, F.w);
        if (!t.touching) continue;
        x.tneed[q] = make_float4(t.need.x, t.need.y, t.need.z, t.cap);
        x.tslide[q] = make_float4(t.slide.x, t.slide.y, t.slide.z, 1.0f);
    }
    __syncthreads();
    for (uint32_t w = r3.x + threadIdx.x; w < r3.y; w += blockDim.x) {
        float4 s = make_float4(0, 0, 0, 0);
        for (uint32_t q = r3.z; q < r3.w; q++) {
            if (x.ty[q].y != w || x.tslide[q].w == 0) continue;
            const float4 nd = x.tneed[q];
            s = make_float4(s.x + nd.x, s.y + nd.y, s.z + nd.z, s.w + nd.w);
        }
        x.tsum[w] = s;
    }
    __syncthreads();
    for (uint32_t q = r3.z + threadIdx.x; q < r3.w; q += blockDim.x) {
        const float4 sl = x.tslide[q];
        if (sl.w == 0) continue;
        const uint2 e = x.ty[q];
        const float4 nd = x.tneed[q];
        bool stick;
        if (e.y != 0xffffffffu) {
            const float4 s = x.tsum[e.y];
            stick = length(mk(s.x, s.y, s.z)) <= s.w;
        } else {
            stick = length(mk(nd.x, nd.y, nd.z)) <= nd.w;
        }
        add_force(a, e.x, stick ? mk(nd.x, nd.y, nd.z) : mk(sl.x, sl.y, sl.z));
    }
}

int launch_tyres(const KArgs& a, const XArgs& x, hipStream_t st, int parity) {
    if (x.ty && x.nbody) hipLaunchKernelGGL(k_tyres, dim3(x.nbody), dim3(256), 0, st, a, x, parity);
    return BL_OK_OR_HIP();
}

__device__ __forceinline__ f3 normalize_or(f3 a, f3 fb) {
    const float l = length(a);
    return l > 1e-12f ? mul(a, 1.0f / l) : fb;
}

// from_mat3 (src/core/math.h:171-201): the rotation of the matrix with columns c0, c1, c2
__device__ qd quat_from_cols(f3 c0, f3 c1, f3 c2) {
    const float m00 = c0.x, m11 = c1.y, m22 = c2.z;
    const float tr = m00 + m11 + m22;
    qd q;
    if (tr > 0) {
        const float s = fsqrt(tr + 1.0f) * 2;
        q = qmk((c1.z - c2.y) / s, (c2.x - c0.z) / s, (c0.y - c1.x) / s, 0.25f * s);
    } else if (m00 > m11 && m00 > m22) {
        const float s = fsqrt(1.0f + m00 - m11 - m22) * 2;
        q = qmk(0.25f * s, (c1.x + c0.y) / s, (c2.x + c0.z) / s, (c1.z - c2.y) / s);
    } else if (m11 > m22) {
        const float s = fsqrt(1.0f + m11 - m00 - m22) * 2;
        q = qmk((c1.x + c0.y) / s, 0.25f * s, (c2.y + c1.z) / s, (c2.x - c0.z) / s);
    } else {
        const float s = fsqrt(1.0f + m22 - m00 - m11) * 2;
        q = qmk((c2.x + c0.z) / s, (c2.y + c1.z) / s, 0.25f * s, (c0.y - c1.x) / s);
    }
    return qnormalize(q);
}

// SoftBody::integrate_frames (src/phys/softbody.cpp:997-1021): a thread per frame, after the nodes' integration
// (P: the integrated positions); a kinematic frame takes its triad's orientation (x towards ref_x, z normal to the
// plane of ref_y), a dynamic one integrates its torque
__global__ void k_frames(const float4* __restrict__ P, XArgs x, float dt) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= x.nframes) return;
    const float4 W = x.fw[f], T = x.ftq[f];
    const int2 rf = x.fref[f];
    if (rf.x >= 0) {
        const float4 pn = P[x.fnode[f]], px = P[rf.x], py = P[rf.y];
        const f3 p = mk(pn.x, pn.y, pn.z);
        const f3 e1 = normalize_or(sub(mk(px.x, px.y, px.z), p), mk(1, 0, 0));
        f3 e3 = normalize_or(cross(e1, sub(mk(py.x, py.y, py.z), p)), mk(0, 0, 1));
        const f3 t = cross(e1, e3);
        if (dot(t, t) < 1e-12f) e3 = normalize_or(cross(e1, mk(0, 1, 0)), mk(0, 0, 1));
        const f3 e2 = cross(e3, e1);
        const qd q = qnormalize(quat_from_cols(e1, e2, e3));
        x.fq[f] = make_float4(q.x, q.y, q.z, q.w);
        x.fw[f] = make_float4(0, 0, 0, W.w);
        x.ftq[f] = make_float4(0, 0, 0, 0);
        return;
    }
    if (W.w <= 0) return;
    f3 w = add(mk(W.x, W.y, W.z), mul(mk(T.x, T.y, T.z), W.w * dt));
    w = mul(w, 1.0f - 0.5f * dt);
    x.fw[f] = make_float4(w.x, w.y, w.z, W.w);
    const qd dq = qmk(w.x, w.y, w.z, 0);
    const qd q = ld_q(x.fq[f]);
    const qd s = qscale(qmul(dq, q), 0.5f * dt);
    const qd n = qnormalize(qmk(q.x + s.x, q.y + s.y, q.z + s.z, q.w + s.w));
    x.fq[f] = make_float4(n.x, n.y, n.z, n.w);
}

int launch_frames(const KArgs& a, const XArgs& x, hipStream_t st, int parity_new) {
    if (x.nframes)
        hipLaunchKernelGGL(k_frames, dim3((x.nframes + 63) / 64), dim3(64), 0, st, (const float4*)a.pm[parity_new], x, a.dt);
    return BL_OK_OR_HIP();
}

int launch_wstat_fold(const KArgs& a, hipStream_t st, uint4* wstat) {
    const uint32_t n = (a.nnodes + 255) / 256;
    if (n) hipLaunchKernelGGL(k_wstat_fold, dim3((n + 255) / 256), dim3(256), 0, st, a, wstat, n);
    return BL_OK_OR_HIP();
}

// Incidences of a chunk: block x incidences per thread (the node cap of a chunk is the block).
int chunk_capacity(int block, int per) { return block * per; }

} // namespace blg
