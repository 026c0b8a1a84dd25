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
//              iA float4 {L, k, d, strength}, iM float4 {max_pos, max_neg, beam index, flags}, iAB uint2 {a, b}.
//              Each end keeps its own copy of the beam's mutable state; both copies see the same inputs and run
//              the same arithmetic, so they stay identical and no end ever reads what the other writes.
//   chunks     contiguous node ranges whose incidences fit in LDS (ch_n0[]): one workgroup each.
//
// A chunk's workgroup evaluates its incidences in parallel (one thread per incidence, forces into LDS), then one
// thread per node adds them up in beam-index order - the order in which the reference's serial loop scatters
// f[a] += fv, f[b] -= fv - so that every node force is bit-identical without atomics.
#include "bl_world.h"

#pragma clang fp contract(off)

namespace blg {

__constant__ ground_model k_gm[BL_SURF_N] = BL_GROUND_MODELS;

// ---------------------------------------------------------------------------------------------- incidence phase
// One incidence of compute_beam_forces (src/phys/softbody.cpp:416-468): the beam's force, signed for this end,
// and the beam's yield / break update on this end's copy of its state.
__device__ __forceinline__ bool incidence_force(uint32_t e, const KArgs& a, const float4* __restrict__ pin,
                                                const float4* __restrict__ vin, f3& out) {
    const float4 M = a.iM[e];
    uint32_t tf = f2u(M.w);
    if (tf & (TF_BROKEN | TF_SKIP)) return false;
    const float4 A = a.iA[e];
    const uint2 ab = a.iAB[e];
    const float4 Pa = pin[ab.x], Pb = pin[ab.y];
    const f3 dis = sub(mk(Pa.x, Pa.y, Pa.z), mk(Pb.x, Pb.y, Pb.z));
    const float len2 = dot(dis, dis);
    if (len2 < 1e-12f) return false;
    const float4 Va = vin[ab.x], Vb = vin[ab.y];
    const float inv_len = 1.0f / fsqrt(len2);
    const float len = len2 * inv_len;
    float L = A.x, strength = A.w, max_pos = M.x, max_neg = M.y;
    const float diff = len - L;
    const f3 dv = sub(mk(Va.x, Va.y, Va.z), mk(Vb.x, Vb.y, Vb.z));
    float k = A.y, d = A.z;
    const uint32_t type = (tf >> TF_TYPE_SHIFT) & 3u;
    const bool brk = (tf & TF_BRK) != 0;
    const uint32_t beam = f2u(M.z);
    const bool owner = (tf & TF_ISB) == 0;
    if (type == BL_BT_ROPE) {
        if (diff < 0) { k = 0; d *= 0.1f; }
    } else if (type == BL_BT_SUPPORT) {
        if (diff > 0) {
            k = 0;
            d *= 0.1f;
            if (brk && diff > L * a.bcold[beam].z) {
                a.iM[e] = make_float4(max_pos, max_neg, M.z, u2f(tf | TF_BROKEN));
                if (owner) atomicAdd(&a.broken[a.ninfo[ab.x] >> 16], 1ull);
                return false;
            }
        }
    }
    float slen = -k * diff - d * dot(dv, dis) * inv_len;
    bool changed = false;
    if (!(tf & TF_NODEFORM) && k != 0) {
        if (slen > max_pos && diff < 0) {
            const float4 cold = a.bcold[beam]; // plastic, L0, support_limit
            const float yield_len = max_pos / k;
            const float def = diff + yield_len * (1.0f - cold.x);
            const float Lold = L;
            L = smax(smin(0.1f, 0.3f * cold.y), L + def);
            slen = slen - (slen - max_pos) * 0.5f;
            if (L < Lold) max_pos *= Lold / L;
            changed = true;
        } else if (slen < max_neg && diff > 0) {
            const float4 cold = a.bcold[beam];
            const float yield_len = max_neg / k;
            const float def = diff + yield_len * (1.0f - cold.x);
            const float Lold = L;
            L = L + def;
            slen = slen - (slen - max_neg) * 0.5f;
            if (Lold > 0 && L > Lold) max_neg *= L / Lold;
            if (!(tf & TF_DUCTILE)) strength -= def * k;
            changed = true;
        }
    }
    if (brk && ffabs(slen) > strength && !(tf & TF_NOBREAK)) {
        a.iA[e] = make_float4(L, A.y, A.z, strength);
        a.iM[e] = make_float4(max_pos, max_neg, M.z, u2f(tf | TF_BROKEN));
        if (owner) atomicAdd(&a.broken[a.ninfo[ab.x] >> 16], 1ull);
        return false;
    }
    if (changed) {
        a.iA[e] = make_float4(L, A.y, A.z, strength);
        a.iM[e] = make_float4(max_pos, max_neg, M.z, M.w);
    }
    if (owner) a.stress[beam] = slen;
    const f3 fv = mul(dis, slen * inv_len);
    // f[b] -= fv is f[b] + (-fv) to the bit: the b end stores the negated force and the gather only adds
    out = owner ? fv : mk(-fv.x, -fv.y, -fv.z);
    return true;
}

// ---------------------------------------------------------------------------------------------- node phase
// StaticWorld::collide_point, the box part (src/phys/static_world.cpp:364-374,407-454), boxes in index order.
__device__ __forceinline__ bool collide_point(const KArgs& a, f3 p, float r, f3& normal, float& depth_out,
                                              uint32_t& surface, float& max_force) {
    bool hit = false;
    float out_depth = 0;
    max_force = 0;
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
    depth_out = out_depth;
    return hit;
}

struct NodeCounts {
    uint32_t contacts, clamps, nans;
    float max_v2; // top |v|^2 of the integrated nodes (integrate_nodes' max_v2)
};

// One node after its force sum: World::collide_static (src/phys/world.cpp:799-869, boxes only) then
// SoftBody::integrate_nodes (src/phys/softbody.cpp:773-791).
__device__ __forceinline__ void node_finish(uint32_t i, const KArgs& a, float4 P, float4 Vv, f3 f, float4* __restrict__ pout,
                                            float4* __restrict__ vout, NodeCounts& cnt) {
    f3 p = mk(P.x, P.y, P.z), v = mk(Vv.x, Vv.y, Vv.z);
    float inv_mass = P.w;
    const float mass = Vv.w;
    const uint32_t info = a.ninfo[i];
    if ((info & BL_NF_GROUND) && inv_mass > 0) {
        if (p.y < a.kill_y) {
            v = mk(0, 0, 0);
            inv_mass = 0;
        } else if (a.nbox > 0) {
            const float2 aux = a.naux[i]; // radius, friction x the body's ground friction
            f3 n;
            float depth, max_force;
            uint32_t surf;
            if (collide_point(a, p, aux.x, n, depth, surf, max_force)) {
                cnt.contacts++;
                const ground_model& gm = k_gm[surf < BL_SURF_N ? surf : 0];
                const BodyParam& bp = a.bodies[info >> 16];
                if (max_force > 0) {
                    f3 fc = primitive_collision(f, v, mass, n, a.dt, gm, depth, aux.y, bp.push_max, bp.slop, -1.0f);
                    const float fl = length(fc);
                    if (fl > max_force) fc = mul(fc, max_force / fl);
                    f = add(f, fc);
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
        if (v2 > kmax * kmax) {
            v = mul(v, kmax / fsqrt(v2));
            v2 = kmax * kmax;
            cnt.clamps++;
        } else if (!(v2 >= 0.0f)) {
            v = mk(0, 0, 0);
            v2 = 0;
            cnt.nans++;
        }
        p = add(p, mul(v, a.dt));
        cnt.max_v2 = v2 > cnt.max_v2 ? v2 : cnt.max_v2;
    }
    pout[i] = make_float4(p.x, p.y, p.z, inv_mass);
    vout[i] = make_float4(v.x, v.y, v.z, mass);
}

// ---------------------------------------------------------------------------------------------- one chunk
template <int T, int CAP>
__device__ __forceinline__ void chunk_substep(uint32_t c, const KArgs& a, int parity, float* sfx, float* sfy,
                                              float* sfz, uint8_t* sok, NodeCounts& cnt) {
    const float4* __restrict__ pin = a.pm[parity];
    const float4* __restrict__ vin = a.vm[parity];
    float4* __restrict__ pout = a.pm[parity ^ 1];
    float4* __restrict__ vout = a.vm[parity ^ 1];
    const uint32_t n0 = a.ch_n0[c], n1 = a.ch_n0[c + 1];
    const uint32_t e0 = a.noff[n0], e1 = a.noff[n1];
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += T) {
        f3 fv;
        const bool ok = incidence_force(e, a, pin, vin, fv);
        const uint32_t s = e - e0;
        sfx[s] = fv.x, sfy[s] = fv.y, sfz[s] = fv.z;
        sok[s] = ok;
    }
    __syncthreads();
    for (uint32_t i = n0 + threadIdx.x; i < n1; i += T) {
        const float4 P = pin[i], Vv = vin[i];
        // clear_forces (src/phys/softbody.cpp:381): f = g m
        f3 f = mul(a.g, Vv.w);
        const uint32_t b0 = a.noff[i] - e0, b1 = a.noff[i + 1] - e0;
        for (uint32_t s = b0; s < b1; s++)
            if (sok[s]) f = add(f, mk(sfx[s], sfy[s], sfz[s]));
        node_finish(i, a, P, Vv, f, pout, vout, cnt);
    }
}

__device__ __forceinline__ void flush_counts(const KArgs& a, NodeCounts cnt) {
    uint32_t c0 = cnt.contacts, c1 = cnt.clamps, c2 = cnt.nans;
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off);
        c1 += __shfl_xor(c1, off);
        c2 += __shfl_xor(c2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c0) atomicAdd(&a.counters[CNT_STATIC], (unsigned long long)c0);
        if (c1) atomicAdd(&a.counters[CNT_CLAMPS], (unsigned long long)c1);
        if (c2) atomicAdd(&a.counters[CNT_NANS], (unsigned long long)c2);
    }
}

// The workgroup's top |v|^2 onto its body's frame word (all nodes of a workgroup belong to one body).
template <int T>
__device__ __forceinline__ void flush_speed(const KArgs& a, float v2, uint32_t body, float* sred) {
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(v2, off);
        v2 = o > v2 ? o : v2;
    }
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v2;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = sred[0];
        for (int w = 1; w < T / 64; w++) m = sred[w] > m ? sred[w] : m;
        if (m > 0) atomicMax(&a.body_frame[body * BF_N + BF_V2], f2u(m));
    }
}

// One substep over all chunks: one workgroup per chunk, chunks laid out contiguously per XCD (the hardware deals
// workgroup b to XCD b % 8; logical chunk ranges per XCD keep a node's neighbours in that XCD's L2).
template <int T, int CAP>
__global__ __launch_bounds__(T) void k_substep(KArgs a, int parity) {
    __shared__ float sfx[CAP], sfy[CAP], sfz[CAP];
    __shared__ uint8_t sok[CAP];
    const uint32_t C = a.nchunks;
    uint32_t c = blockIdx.x;
    if (a.xcd_remap) {
        const uint32_t per = C / 8, rem = C % 8, x = blockIdx.x % 8, idx = blockIdx.x / 8;
        c = x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
    }
    NodeCounts cnt = {0, 0, 0, 0.0f};
    chunk_substep<T, CAP>(c, a, parity, sfx, sfy, sfz, sok, cnt);
    flush_counts(a, cnt);
    __shared__ float sred[T / 64];
    flush_speed<T>(a, cnt.max_v2, a.ninfo[a.ch_n0[c]] >> 16, sred);
}

// Many substeps per launch, one workgroup per body: the body's chunks one after another, a workgroup barrier
// between phases (bodies are independent without inter-body contacts).
template <int T, int CAP>
__global__ __launch_bounds__(T) void k_body_persistent(KArgs a, int parity0, int substeps) {
    __shared__ float sfx[CAP], sfy[CAP], sfz[CAP];
    __shared__ uint8_t sok[CAP];
    const uint32_t body = blockIdx.x;
    const uint32_t c0 = a.body_c0[body], c1 = a.body_c0[body + 1];
    NodeCounts cnt = {0, 0, 0, 0.0f};
    for (int s = 0; s < substeps; s++) {
        const int parity = parity0 ^ (s & 1);
        for (uint32_t c = c0; c < c1; c++) {
            chunk_substep<T, CAP>(c, a, parity, sfx, sfy, sfz, sok, cnt);
            __syncthreads();
        }
    }
    flush_counts(a, cnt);
    __shared__ float sred[T / 64];
    flush_speed<T>(a, cnt.max_v2, body, sred);
}

// Bounds and top speed of the current state (SoftBody::integrate_nodes' mn/mx/max_v2 over all nodes).
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

// The frame's start (src/phys/world.cpp:3217-3222): every body's top |v|^2 over all its nodes (NaN left out, as
// std::max leaves it), and its bounds keys reset.
__global__ void k_frame_begin(KArgs a, int parity) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.nbodies) {
        unsigned int* w = a.body_frame + i * BF_N;
        w[BF_MN] = w[BF_MN + 1] = w[BF_MN + 2] = fkey(1e30f);
        w[BF_MX] = w[BF_MX + 1] = w[BF_MX + 2] = fkey(-1e30f);
    }
    if (i >= a.nnodes) return;
    const float4 V = a.vm[parity][i];
    const f3 v = mk(V.x, V.y, V.z);
    const float v2 = dot(v, v);
    if (v2 > 0) atomicMax(&a.body_frame[(a.ninfo[i] >> 16) * BF_N + BF_V2], f2u(v2));
}

// The frame's end: each body's bounds (vmin / vmax over its nodes: NaN coordinates leave them unchanged).
__global__ void k_frame_end(KArgs a, int parity) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nnodes) return;
    const float4 P = a.pm[parity][i];
    unsigned int* w = a.body_frame + (a.ninfo[i] >> 16) * BF_N;
    const float p[3] = {P.x, P.y, P.z};
    for (int k = 0; k < 3; k++) {
        if (p[k] != p[k]) continue;
        atomicMin(&w[BF_MN + k], fkey(p[k]));
        atomicMax(&w[BF_MX + k], fkey(p[k]));
    }
}

__global__ void k_clear_u32(unsigned int* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------------------------- launches
int launch_substeps(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t) {
    if (t.mode == BL_MODE_PERSISTENT) {
        if (t.block == 512) hipLaunchKernelGGL((k_body_persistent<512, 2048>), dim3(a.nbodies), dim3(512), 0, st, a, parity, substeps);
        else if (t.block == 1024) hipLaunchKernelGGL((k_body_persistent<1024, 4096>), dim3(a.nbodies), dim3(1024), 0, st, a, parity, substeps);
        else hipLaunchKernelGGL((k_body_persistent<256, 1024>), dim3(a.nbodies), dim3(256), 0, st, a, parity, substeps);
        return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
    }
    for (int s = 0; s < substeps; s++) {
        const int p = parity ^ (s & 1);
        if (t.block == 512) hipLaunchKernelGGL((k_substep<512, 2048>), dim3(a.nchunks), dim3(512), 0, st, a, p);
        else if (t.block == 1024) hipLaunchKernelGGL((k_substep<1024, 4096>), dim3(a.nchunks), dim3(1024), 0, st, a, p);
        else hipLaunchKernelGGL((k_substep<256, 1024>), dim3(a.nchunks), dim3(256), 0, st, a, p);
    }
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

int launch_bounds(const KArgs& a, hipStream_t st, int parity, unsigned int* out) {
    if (a.nnodes == 0) return BL_OK;
    hipLaunchKernelGGL(k_bounds, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity, out);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

int launch_frame_begin(const KArgs& a, hipStream_t st, int parity) {
    const uint32_t nw = a.nbodies * BF_N;
    hipLaunchKernelGGL(k_clear_u32, dim3((nw + 255) / 256), dim3(256), 0, st, a.body_frame, nw);
    const uint32_t n = a.nnodes > a.nbodies ? a.nnodes : a.nbodies;
    hipLaunchKernelGGL(k_frame_begin, dim3((n + 255) / 256), dim3(256), 0, st, a, parity);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

int launch_frame_end(const KArgs& a, hipStream_t st, int parity) {
    if (a.nnodes == 0) return BL_OK;
    hipLaunchKernelGGL(k_frame_end, dim3((a.nnodes + 255) / 256), dim3(256), 0, st, a, parity);
    return hipGetLastError() == hipSuccess ? BL_OK : BL_ERR_HIP;
}

int chunk_capacity(int block) { return block == 512 ? 2048 : block == 1024 ? 4096 : 1024; }

} // namespace blg
