// The C ABI (include/beamlab_gpu.h): context, device memory, uploads, launches and read-back.
#include "bl_contacts.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#pragma clang fp contract(off)

using namespace blg;

namespace {

thread_local std::string g_last_error;

#define HIPCHK(x)                                                                                                   \
    do {                                                                                                            \
        hipError_t e_ = (x);                                                                                        \
        if (e_ != hipSuccess) {                                                                                     \
            g_last_error = std::string(#x) + ": " + hipGetErrorString(e_);                                          \
            return BL_ERR_HIP;                                                                                      \
        }                                                                                                           \
    } while (0)

struct HostBody {
    std::vector<bl_node> nodes;
    std::vector<bl_node_info> info;
    std::vector<bl_beam> beams;
    std::vector<bl_triangle> tris;
    std::vector<bl_capsule> caps;
    std::vector<bl_shock> shocks;
    std::vector<bl_slide> slides; // (seg and broken: the step's state)
    std::vector<uint32_t> rail;
    std::vector<bl_frame> frames; // (q, w and the joints' state: the step's)
    std::vector<bl_joint> joints;
    std::vector<bl_wheel> wheels; // (torque, brake and the speeds: the step's)
    std::vector<uint32_t> wnodes;
    std::vector<float4> ctl;      // a controller's torque, brake, on (pre_substep) per wheel
    float wsp[2] = {0, 0};        // SoftBody::wheel_speed, wheel_spin
    bool tyres = false;           // it has tyre nodes
    double guard_ke = 0, guard_pe = 0; // the energy guard's motion energy at the frame's start
    int guard_quiet = 0, guard_cuts = 0;
    bl_body_opts opts;
    uint32_t node0 = 0, beam0 = 0, tri0 = 0, cap0 = 0;
    uint32_t slide0 = 0; // its first slide in the device array
    uint32_t frame0 = 0, joint0 = 0; // its first frame and joint in the device arrays
    uint32_t wheel0 = 0;             // its first wheel in the device array
    int32_t xi = -1;                 // its index among the bodies with extras (the device's per-body records)
    // as of the next frame's start: top speed of the last frame and bounds (World::build_islands reads them)
    float max_speed = 0;
    f3 mn = mk(1e30f, 1e30f, 1e30f), mx = mk(-1e30f, -1e30f, -1e30f);
    void compute_bounds() {
        mn = mk(1e30f, 1e30f, 1e30f), mx = mk(-1e30f, -1e30f, -1e30f);
        for (const bl_node& x : nodes) mn = vmin(mn, mk(x.p[0], x.p[1], x.p[2])), mx = vmax(mx, mk(x.p[0], x.p[1], x.p[2]));
    }
};

// SoftBody::hull_reach (src/phys/softbody.h:751-758): how far past its bounds the body reaches for other bodies'
// nodes - its hull depth, when it has a hull (one-sided) triangle
inline float hull_reach(const HostBody& hb) {
    for (const bl_triangle& t : hb.tris)
        if (!t.two_sided) return hb.opts.hull_depth;
    return 0.0f;
}

template <class T>
struct DevArr {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return BL_OK;
        HIPCHK(hipMalloc((void**)&p, sizeof(T) * count));
        return BL_OK;
    }
    int upload(const std::vector<T>& v) {
        int rc = alloc(v.size());
        if (rc) return rc;
        if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        return BL_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

} // namespace

// Page-locked host memory (hipHostMalloc): copies to and from it run on the copy engines without staging.
struct HostPinned {
    void* p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t n) {
        if (p && n <= bytes) return true;
        release();
        if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = n;
        return true;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct bl_ctx {
    bl_settings s{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<HostBody> bodies;
    std::vector<bl_box> boxes;
    bool built = false;
    int built_block = 0, built_per = 0; // the chunking of the device arrays
    int parity = 0;
    bl_tuning tune{};
    int64_t substeps = 0;
    // device
    DevArr<float4> pm[2], vm[2], iA, iM, bcold;
    DevArr<uint32_t> ninfo, noff, ch_n0, body_c0, chunk_body;
    DevArr<uint4> ch_rng, fch_rng, wg_rng, fwg_rng;
    DevArr<uint32_t> fwg;
    DevArr<uint4> cstat;
    DevArr<uint32_t> halo_off, halo, nbr_off, nbr, iloc, flags, wg_chunk;
    DevArr<float2> naux;
    DevArr<uint32_t> iOth;
    DevArr<float> stress;
    DevArr<BodyParam> bparam;
    DevArr<bl_box> dboxes;
    // the static world's cylinders and terrain (host copies; the terrain's 16 x 16 vertex blocks' tops), and each
    // body's terrain top over its reach for the frame (World::build_islands)
    std::vector<bl_cylinder> cyls;
    int tnx = 0, tnz = 0, tbx = 0, tbz = 0;
    float tcell = 1, tox = 0, toz = 0;
    std::vector<float> th, tblock, htmax;
    std::vector<uint8_t> ts;
    DevArr<bl_cylinder> dcyl;
    DevArr<float> dth, dtmax;
    DevArr<uint8_t> dts;
    DevArr<unsigned long long> counters, broken;
    DevArr<unsigned int> bounds, body_frame, bar;
    DevArr<int> err;
    // inter-body contacts
    DevArr<float4> F;
    DevArr<uint4> tri_abc, cap_data;
    DevArr<float> tri_rest2;
    DevArr<uint32_t> tri_torn;
    DevArr<uint8_t> tri_hull;
    DevArr<unsigned int> cbox;
    DevArr<uint4> wstat; // per workgroup of the node-parallel integration: its counts of the frame
    // shocks and slide nodes (XArgs): bodies that have them step in the split form (forces, these, integration)
    DevArr<uint4> xbody, shx;
    DevArr<bl_shock> xsh;
    DevArr<bl_slide> xsl;
    DevArr<uint32_t> xrail;
    DevArr<bl_joint> xjt;
    DevArr<float4> fq, fw, ftq;
    DevArr<uint32_t> fnode;
    DevArr<int2> fref;
    DevArr<bl_wheel> xwh;
    DevArr<uint32_t> xwn;
    DevArr<float4> xctl;
    DevArr<float2> xwsp;
    DevArr<uint2> xty;
    DevArr<float4> xtsum, xtneed, xtslide;
    XArgs xargs{};
    std::map<std::pair<int, int>, hipGraphExec_t> sgraphs; // the split form's captured substeps
    ContactEngine contacts;
    HostView hv;
    bool contact_geometry = false; // some body has triangles, capsules or self collision
    bool frame_pending = false;   // a frame enqueued whose post-frame check has not run
    bool no_coop = false;         // the runtime refused a cooperative launch: the automatic choice steps
    // pinned host staging: the frame-end readback (error word, then the body frame words), queued behind the
    // frame's kernels so that the post-frame check only waits; the first rebuild's bounds words on their way up
    HostPinned fin, w0;
    bool fin_queued = false;
    hipEvent_t w0_ev = nullptr;
    int64_t repairs = 0;
    std::vector<uint32_t> a_slot;   // global beam -> incidence of its a end
    std::vector<uint32_t> h_noff;
    KArgs args{};
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    // captured contact segments (the substeps from one pair rebuild to the next), for the lists' current buffers
    struct Segment {
        int len, parity, tail;
        hipGraphExec_t ex;
    };
    std::vector<Segment> segs;
    PArgs seg_p{};
    uint32_t seg_nblk = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> kev;  // per-launch events of bl_time_kernels

    void drop_segments() {
        for (auto& g : segs) (void)hipGraphExecDestroy(g.ex);
        segs.clear();
    }
    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        for (auto& kv : sgraphs) (void)hipGraphExecDestroy(kv.second);
        sgraphs.clear();
        drop_segments();
    }
    void release() {
        drop_graphs();
        for (int k = 0; k < 2; k++) pm[k].release(), vm[k].release();
        iA.release(), iM.release(), bcold.release(), ninfo.release(), noff.release(), ch_n0.release(), ch_rng.release(), fch_rng.release(), fwg.release(), wg_rng.release(), fwg_rng.release();
        body_c0.release(), naux.release(), iOth.release(), stress.release(), bparam.release(), dboxes.release();
        dcyl.release(), dth.release(), dtmax.release(), dts.release();
        counters.release(), broken.release(), bounds.release(), body_frame.release(), bar.release(), err.release();
        chunk_body.release(), cstat.release();
        halo_off.release(), halo.release(), nbr_off.release(), nbr.release(), iloc.release(), flags.release();
        wg_chunk.release();
        F.release(), tri_abc.release(), cap_data.release(), tri_rest2.release(), tri_torn.release(), tri_hull.release(), cbox.release();
        wstat.release();
        xbody.release(), shx.release(), xsh.release(), xsl.release(), xrail.release();
        xjt.release(), fq.release(), fw.release(), ftq.release(), fnode.release(), fref.release();
        xwh.release(), xwn.release(), xctl.release(), xwsp.release(), xty.release(), xtsum.release(), xtneed.release();
        xtslide.release();
        xargs = XArgs{};
        contacts.release();
        frame_pending = false;
        fin_queued = false;
        built = false;
    }
};

namespace {

int effective_block(const bl_ctx* c) { return c->tune.block == 512 || c->tune.block == 1024 ? c->tune.block : 256; }
int effective_per(const bl_ctx* c) {
    const int p = c->tune.incidences_per_thread;
    return p == 1 || p == 2 || p == 8 ? p : 4;
}

// The (block, incidences per thread) pairs each form is instantiated for.
bool tuning_valid(const bl_tuning& t) {
    const int b = t.block ? t.block : 256, p = t.incidences_per_thread ? t.incidences_per_thread : 4;
    switch (t.mode) {
    case BL_MODE_AUTO: return (b == 256 && (p == 1 || p == 2 || p == 4 || p == 8)) || (b == 512 && (p == 2 || p == 4));
    case BL_MODE_STEPPED: return (b == 256 && (p == 1 || p == 2 || p == 4 || p == 8)) || (b == 512 && (p == 2 || p == 4));
    case BL_MODE_PERSISTENT: return (b == 256 || b == 1024) && p == 4;
    case BL_MODE_RESIDENT: return (b == 256 && (p == 2 || p == 4 || p == 8)) || (b == 512 && p == 4);
    case BL_MODE_HALO: return b == 256 && (p == 4 || p == 8);
    default: return false;
    }
}

constexpr uint32_t kSpinLimit = 1u << 20; // grid-barrier polls (about a second) before a resident launch gives up

bool unsupported_opts(const bl_body_opts& o) {
    return o.air_drag != 0 || o.aero_cda != 0 || o.rest_damp != 0;
}

// Device state -> host copies of every body (nodes, beam state), so that a rebuild starts from the current state.
// some body has shocks or slide nodes: the world steps in the split form (F holds the node forces between phases)
bool has_extras(const bl_ctx* c) {
    for (const HostBody& hb : c->bodies)
        if (!hb.shocks.empty() || !hb.slides.empty() || !hb.joints.empty() || !hb.frames.empty() || !hb.wheels.empty() ||
            hb.tyres)
            return true;
    return false;
}

int pull_state(bl_ctx* c) {
    if (!c->built) return BL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t N = c->pm[0].n;
    std::vector<float4> P(N), V(N);
    if (N) {
        HIPCHK(hipMemcpy(P.data(), c->pm[c->parity].p, sizeof(float4) * N, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(V.data(), c->vm[c->parity].p, sizeof(float4) * N, hipMemcpyDeviceToHost));
    }
    const size_t E = c->iA.n;
    std::vector<float4> A(E), M(E);
    if (E) {
        HIPCHK(hipMemcpy(A.data(), c->iA.p, sizeof(float4) * E, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(M.data(), c->iM.p, sizeof(float4) * E, hipMemcpyDeviceToHost));
    }
    const size_t B = c->stress.n;
    std::vector<float> S(B);
    if (B) HIPCHK(hipMemcpy(S.data(), c->stress.p, sizeof(float) * B, hipMemcpyDeviceToHost));
    std::vector<uint32_t> torn(c->tri_torn.n);
    if (!torn.empty()) HIPCHK(hipMemcpy(torn.data(), c->tri_torn.p, 4 * torn.size(), hipMemcpyDeviceToHost));
    std::vector<bl_slide> sl(c->xsl.n);
    if (!sl.empty()) HIPCHK(hipMemcpy(sl.data(), c->xsl.p, sizeof(bl_slide) * sl.size(), hipMemcpyDeviceToHost));
    std::vector<bl_wheel> wh(c->xwh.n);
    if (!wh.empty()) HIPCHK(hipMemcpy(wh.data(), c->xwh.p, sizeof(bl_wheel) * wh.size(), hipMemcpyDeviceToHost));
    std::vector<float2> wsp(c->xwsp.n);
    if (!wsp.empty()) HIPCHK(hipMemcpy(wsp.data(), c->xwsp.p, sizeof(float2) * wsp.size(), hipMemcpyDeviceToHost));
    std::vector<bl_joint> jt(c->xjt.n);
    if (!jt.empty()) HIPCHK(hipMemcpy(jt.data(), c->xjt.p, sizeof(bl_joint) * jt.size(), hipMemcpyDeviceToHost));
    std::vector<float4> fq(c->fq.n), fw(c->fw.n);
    if (!fq.empty()) {
        HIPCHK(hipMemcpy(fq.data(), c->fq.p, sizeof(float4) * fq.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(fw.data(), c->fw.p, sizeof(float4) * fw.size(), hipMemcpyDeviceToHost));
    }
    for (HostBody& hb : c->bodies) {
        for (size_t t = 0; t < hb.tris.size() && hb.tri0 + t < torn.size(); t++) hb.tris[t].torn = torn[hb.tri0 + t] ? 1 : 0;
        for (size_t q = 0; q < hb.slides.size() && hb.slide0 + q < sl.size(); q++)
            hb.slides[q].seg = sl[hb.slide0 + q].seg, hb.slides[q].broken = sl[hb.slide0 + q].broken;
        for (size_t q = 0; q < hb.frames.size() && hb.frame0 + q < fq.size(); q++) {
            const float4 a = fq[hb.frame0 + q], b = fw[hb.frame0 + q];
            bl_frame& f = hb.frames[q];
            f.q[0] = a.x, f.q[1] = a.y, f.q[2] = a.z, f.q[3] = a.w, f.w[0] = b.x, f.w[1] = b.y, f.w[2] = b.z;
        }
        for (size_t q = 0; q < hb.wheels.size() && hb.wheel0 + q < wh.size(); q++) {
            const bl_wheel& d = wh[hb.wheel0 + q];
            bl_wheel& w = hb.wheels[q];
            w.torque = d.torque, w.drive_torque = d.drive_torque, w.brake = d.brake, w.speed = d.speed;
            w.avg_speed = d.avg_speed, w.last_torque = d.last_torque, w.last_retorque = d.last_retorque;
        }
        if (hb.xi >= 0 && (size_t)hb.xi < wsp.size()) hb.wsp[0] = wsp[hb.xi].x, hb.wsp[1] = wsp[hb.xi].y;
        for (size_t q = 0; q < hb.joints.size() && hb.joint0 + q < jt.size(); q++) {
            const bl_joint& d = jt[hb.joint0 + q];
            bl_joint& j = hb.joints[q];
            for (int k = 0; k < 3; k++) j.local_offset[k] = d.local_offset[k];
            for (int k = 0; k < 4; k++) j.rel_rot[k] = d.rel_rot[k];
            j.plastic_lin = d.plastic_lin, j.plastic_ang = d.plastic_ang, j.stress = d.stress, j.broken = d.broken;
        }
        for (size_t i = 0; i < hb.nodes.size(); i++) {
            const float4 p = P[hb.node0 + i], v = V[hb.node0 + i];
            bl_node& n = hb.nodes[i];
            n.p[0] = p.x, n.p[1] = p.y, n.p[2] = p.z, n.inv_mass = p.w;
            n.v[0] = v.x, n.v[1] = v.y, n.v[2] = v.z, n.mass = v.w;
        }
        for (size_t j = 0; j < hb.beams.size(); j++) {
            const uint32_t e = c->a_slot[hb.beam0 + j];
            bl_beam& b = hb.beams[j];
            b.stress = S[hb.beam0 + j];
            if (e == 0xffffffffu) continue;
            b.L = A[e].x, b.strength = A[e].w, b.max_pos = M[e].x, b.max_neg = M[e].y;
            if (f2u(M[e].w) & TF_BROKEN) b.flags |= BL_BF_BROKEN;
        }
    }
    return BL_OK;
}

int build(bl_ctx* c) {
    c->release();
    const int T = effective_block(c);
    const uint32_t CAP = (uint32_t)chunk_capacity(T, effective_per(c));
    const uint32_t NCAP = (uint32_t)T; // nodes per chunk: one per thread
    uint32_t N = 0, B = 0, NT = 0, NC = 0;
    for (HostBody& hb : c->bodies) {
        hb.node0 = N, hb.beam0 = B, hb.tri0 = NT, hb.cap0 = NC;
        N += (uint32_t)hb.nodes.size();
        B += (uint32_t)hb.beams.size();
        NT += (uint32_t)hb.tris.size();
        NC += (uint32_t)hb.caps.size();
    }
    // collision triangles and capsules with global node indices; the island builder's view of the bodies
    std::vector<uint4> tri_abc(NT), cap_data(NC);
    std::vector<float> tri_rest2(NT);
    std::vector<uint32_t> tri_torn(NT);
    std::vector<uint8_t> tri_hull(NT);
    c->hv = HostView{};
    c->hv.dt = c->s.dt, c->hv.inter_body = c->s.inter_body_collisions;
    c->contact_geometry = false;
    for (size_t k = 0; k < c->bodies.size(); k++) {
        const HostBody& hb = c->bodies[k];
        for (size_t t = 0; t < hb.tris.size(); t++) {
            const bl_triangle& tr = hb.tris[t];
            tri_abc[hb.tri0 + t] = make_uint4(hb.node0 + tr.a, hb.node0 + tr.b, hb.node0 + tr.c, (uint32_t)k);
            tri_rest2[hb.tri0 + t] = tr.rest_edge2;
            tri_torn[hb.tri0 + t] = tr.torn;
            tri_hull[hb.tri0 + t] = tr.two_sided ? 0 : 1;
        }
        for (size_t q = 0; q < hb.caps.size(); q++) {
            const bl_capsule& cp = hb.caps[q];
            cap_data[hb.cap0 + q] = make_uint4(hb.node0 + cp.a, hb.node0 + cp.b, (uint32_t)k, f2u(cp.radius));
        }
        BodyView v{};
        v.group = hb.opts.collision_group, v.self = hb.opts.self_collision != 0;
        v.ntri = (uint32_t)hb.tris.size(), v.ncap = (uint32_t)hb.caps.size(), v.tri0 = hb.tri0;
        v.nnode = (uint32_t)hb.nodes.size();
        v.collision_radius = hb.opts.collision_radius;
        v.hull_reach = hull_reach(hb);
        for (const bl_node_info& inf : hb.info) v.contacters |= (inf.flags & BL_NF_CONTACTER) ? 1u : 0u;
        c->hv.bodies.push_back(v);
        c->contact_geometry |= v.ntri > 0 || v.ncap > 0 || v.self;
    }
    c->hv.max_speed.assign(c->bodies.size(), 0.0f);
    c->hv.mn.assign(c->bodies.size(), mk(0, 0, 0));
    c->hv.mx.assign(c->bodies.size(), mk(0, 0, 0));
    std::vector<float4> P(N), V(N);
    std::vector<uint32_t> ninfo(N), deg(N + 1, 0);
    std::vector<float2> naux(N);
    std::vector<BodyParam> bp(c->bodies.size());
    for (size_t k = 0; k < c->bodies.size(); k++) {
        const HostBody& hb = c->bodies[k];
        bp[k] = BodyParam{hb.opts.bounce, hb.opts.contact_push_max, hb.opts.contact_slop, hb.opts.collision_radius,
                          hb.opts.contact_friction, hb.opts.hull_depth, hb.opts.face_skin, hull_reach(hb),
                          hb.opts.collision_group, hb.opts.self_collision, hb.opts.faces_only, 0,
                          hb.opts.allow_deform, hb.opts.allow_break, 0, 0};
        const float fric_body = hb.opts.ground_friction * 1.0f; // (resting is off: src/phys/world.cpp:792)
        for (size_t i = 0; i < hb.nodes.size(); i++) {
            const bl_node& n = hb.nodes[i];
            const uint32_t g = hb.node0 + (uint32_t)i;
            P[g] = make_float4(n.p[0], n.p[1], n.p[2], n.inv_mass);
            V[g] = make_float4(n.v[0], n.v[1], n.v[2], n.mass);
            ninfo[g] = (uint32_t)hb.info[i].flags | ((uint32_t)k << 16);
            naux[g] = make_float2(hb.info[i].radius, hb.info[i].friction * fric_body);
        }
        for (const bl_beam& b : hb.beams) deg[hb.node0 + b.a]++, deg[hb.node0 + b.b]++;
    }
    std::vector<uint32_t> noff(N + 1, 0);
    for (uint32_t i = 0; i < N; i++) noff[i + 1] = noff[i] + deg[i];
    const uint32_t E = noff[N];
    std::vector<uint32_t> fill(noff.begin(), noff.end() - 1);
    std::vector<float4> iA(E), iM(E), bcold(B);
    std::vector<uint2> iAB(E);
    std::vector<float> stress(B);
    c->a_slot.assign(B, 0xffffffffu);
    std::vector<uint32_t> b_slot(B, 0xffffffffu);
    for (size_t k = 0; k < c->bodies.size(); k++) {
        const HostBody& hb = c->bodies[k];
        const bl_body_opts& o = hb.opts;
        for (size_t j = 0; j < hb.beams.size(); j++) {
            const bl_beam& b = hb.beams[j];
            const uint32_t gb = hb.beam0 + (uint32_t)j, ga = hb.node0 + b.a, gbn = hb.node0 + b.b;
            uint32_t tf = 0;
            if (b.flags & BL_BF_BROKEN) tf |= TF_BROKEN;
            if (b.flags & BL_BF_SHOCK) tf |= TF_SKIP;
            if ((b.flags & BL_BF_NO_DEFORM) || !o.allow_deform) tf |= TF_NODEFORM;
            if (b.flags & BL_BF_NO_BREAK) tf |= TF_NOBREAK;
            if (o.allow_break) tf |= TF_BRK;
            if (o.ductile) tf |= TF_DUCTILE;
            tf |= (uint32_t)b.type << TF_TYPE_SHIFT;
            bcold[gb] = make_float4(b.plastic, b.L0, b.support_limit, 0);
            stress[gb] = b.stress;
            for (int side = 0; side < 2; side++) {
                const uint32_t node = side ? gbn : ga;
                const uint32_t e = fill[node]++;
                iA[e] = make_float4(b.L, b.k, b.d, b.strength);
                iM[e] = make_float4(b.max_pos, b.max_neg, u2f(gb), u2f(tf | (side ? TF_ISB : 0u)));
                iAB[e] = make_uint2(ga, gbn);
                if (!side) c->a_slot[gb] = e;
                else b_slot[gb] = e;
            }
        }
    }
    // chunks: contiguous node ranges of one body with at most CAP incidences, NCAP nodes and HCAP nodes read (its
    // own and its halo together)
    const uint32_t HCAP = (uint32_t)halo_capacity(T, effective_per(c));
    std::vector<uint32_t> ch_n0, body_c0, chunk_body;
    std::vector<uint32_t> mark(N, 0xffffffffu);
    uint32_t stamp = 0, reach = 0;
    auto count_new = [&](uint32_t node, bool commit) {
        uint32_t add = 0;
        auto touch = [&](uint32_t g) {
            if (mark[g] != stamp) {
                add++;
                if (commit) mark[g] = stamp;
            }
        };
        if (!commit) {
            // (a dry count may see a node twice: an upper bound, enough for the cap)
            add += mark[node] != stamp;
            for (uint32_t e = noff[node]; e < noff[node + 1]; e++) add += (mark[iAB[e].x] != stamp) + (mark[iAB[e].y] != stamp);
            return add;
        }
        touch(node);
        for (uint32_t e = noff[node]; e < noff[node + 1]; e++) touch(iAB[e].x), touch(iAB[e].y);
        return add;
    };
    for (const HostBody& hb : c->bodies) {
        const uint32_t bidx = (uint32_t)body_c0.size();
        body_c0.push_back((uint32_t)ch_n0.size());
        const uint32_t n0 = hb.node0, n1 = hb.node0 + (uint32_t)hb.nodes.size();
        uint32_t start = n0;
        stamp++, reach = 0;
        for (uint32_t i = n0; i < n1; i++) {
            if (deg[i] > CAP || 1 + 2 * deg[i] > HCAP) {
                g_last_error = "a node has more beams than a workgroup holds";
                return BL_ERR_UNSUPPORTED;
            }
            if (i > start && (noff[i + 1] - noff[start] > CAP || i + 1 - start > NCAP || reach + count_new(i, false) > HCAP)) {
                ch_n0.push_back(start), chunk_body.push_back(bidx);
                start = i;
                stamp++, reach = 0;
            }
            reach += count_new(i, true);
        }
        if (n1 > start) ch_n0.push_back(start), chunk_body.push_back(bidx);
    }
    body_c0.push_back((uint32_t)ch_n0.size());
    const uint32_t nchunks = (uint32_t)ch_n0.size();
    // workgroup -> chunk: XCD x runs workgroups x, x + 8, ...; each XCD has its own L2, and a chunk's beams read
    // the nodes of its body. A body of more than 1/32 of all chunks (and at least 8) is split into 8 contiguous
    // ranges, one per XCD (the longer ranges rotating over the XCDs from body to body); every other body goes whole
    // to one XCD, the largest first onto the least loaded, so that its nodes are fetched into one L2 only. Lists
    // longer than their XCD's share hand their tails to the short ones.
    std::vector<std::vector<uint32_t>> xl(8);
    {
        uint32_t rot = 0;
        std::vector<uint32_t> whole;
        for (size_t k = 0; k + 1 < body_c0.size(); k++) {
            const uint32_t c0 = body_c0[k], n = body_c0[k + 1] - c0;
            if (n >= 8 && (uint64_t)n * 32 > nchunks) {
                uint32_t at = c0;
                for (uint32_t q = 0; q < 8; q++) {
                    const uint32_t x = (rot + q) % 8, len = n / 8 + (q < n % 8 ? 1u : 0u);
                    for (uint32_t j = 0; j < len; j++) xl[x].push_back(at++);
                }
                rot = (rot + n % 8) % 8;
            } else if (n) {
                whole.push_back((uint32_t)k);
            }
        }
        std::stable_sort(whole.begin(), whole.end(), [&](uint32_t a, uint32_t b) {
            return body_c0[a + 1] - body_c0[a] > body_c0[b + 1] - body_c0[b];
        });
        for (uint32_t k : whole) {
            uint32_t x = 0;
            for (uint32_t y = 1; y < 8; y++)
                if (xl[y].size() < xl[x].size()) x = y;
            for (uint32_t c = body_c0[k]; c < body_c0[k + 1]; c++) xl[x].push_back(c);
        }
        std::vector<uint32_t> spill;
        for (uint32_t x = 0; x < 8; x++) {
            const size_t share = nchunks > x ? (nchunks - x + 7) / 8 : 0;
            while (xl[x].size() > share) spill.push_back(xl[x].back()), xl[x].pop_back();
        }
        for (uint32_t x = 0; x < 8; x++) {
            const size_t share = nchunks > x ? (nchunks - x + 7) / 8 : 0;
            while (xl[x].size() < share) xl[x].push_back(spill.back()), spill.pop_back();
        }
    }
    std::vector<uint32_t> wg_chunk(nchunks);
    for (uint32_t x = 0; x < 8; x++)
        for (size_t j = 0; j < xl[x].size(); j++) wg_chunk[j * 8 + x] = xl[x][j];
    ch_n0.push_back(N);
    c->h_noff = noff;
    // each incidence's own node as an offset from its chunk's first node (in the flags' high bits), and the other end
    std::vector<uint32_t> iOth(E);
    for (uint32_t k = 0; k + 1 < ch_n0.size(); k++)
        for (uint32_t i = ch_n0[k]; i < ch_n0[k + 1]; i++)
            for (uint32_t e = noff[i]; e < noff[i + 1]; e++) {
                const uint32_t tf = f2u(iM[e].w), off = i - ch_n0[k];
                if (off >= (1u << (32 - TF_OWN_SHIFT))) return BL_ERR_ARG;
                iM[e].w = u2f(tf | (off << TF_OWN_SHIFT));
                iOth[e] = (tf & TF_ISB) ? iAB[e].x : iAB[e].y;
            }
    std::vector<uint4> ch_rng(nchunks);
    for (uint32_t k = 0; k < nchunks; k++) ch_rng[k] = make_uint4(ch_n0[k], ch_n0[k + 1], noff[ch_n0[k]], noff[ch_n0[k + 1]]);
    // the force phase's chunks: those with beams as they are, the beamless ones' nodes in contiguous runs of <= T
    std::vector<uint4> fch_rng;
    {
        uint32_t r0 = 0, r1 = 0;
        std::vector<uint4> packed;
        for (uint32_t k = 0; k < nchunks; k++) {
            const uint4 cr = ch_rng[k];
            if (cr.z != cr.w) continue;
            if (r1 > r0 && (cr.x != r1 || cr.y - r0 > T)) {
                packed.push_back(make_uint4(r0, r1, noff[r0], noff[r0]));
                r0 = r1 = 0;
            }
            if (r1 == r0) r0 = cr.x;
            r1 = cr.y;
        }
        if (r1 > r0) packed.push_back(make_uint4(r0, r1, noff[r0], noff[r0]));
        for (uint32_t k = 0; k < nchunks; k++)
            if (ch_rng[k].z != ch_rng[k].w) fch_rng.push_back(ch_rng[k]);
        fch_rng.insert(fch_rng.end(), packed.begin(), packed.end());
    }
    const uint32_t nf = (uint32_t)fch_rng.size();
    std::vector<uint32_t> fwg;
    {
        uint32_t st[8], len[8];
        for (uint32_t x = 0; x < 8; x++) st[x] = (uint32_t)((uint64_t)nf * x / 8), len[x] = (uint32_t)((uint64_t)nf * (x + 1) / 8) - st[x];
        for (uint32_t j = 0; fwg.size() < nf; j++)
            for (uint32_t x = 0; x < 8; x++)
                if (j < len[x]) fwg.push_back(st[x] + j);
    }
    // each workgroup's chunk range itself (one load at a workgroup's start instead of two dependent ones)
    std::vector<uint4> wg_rng(nchunks), fwg_rng(nf);
    for (uint32_t b = 0; b < nchunks; b++) wg_rng[b] = ch_rng[wg_chunk[b]];
    for (uint32_t b = 0; b < nf; b++) fwg_rng[b] = fch_rng[fwg[b]];
    // halos: per chunk the outside nodes its incidences read (sorted), the chunks owning them, each incidence's
    // two LDS slots
    std::vector<uint32_t> node_chunk(N);
    for (uint32_t k = 0; k < nchunks; k++)
        for (uint32_t i = ch_n0[k]; i < ch_n0[k + 1]; i++) node_chunk[i] = k;
    std::vector<uint32_t> halo_off(nchunks + 1, 0), halo, nbr_off(nchunks + 1, 0), nbr, iloc(E);
    std::vector<uint32_t> ext;
    std::vector<uint32_t> slot(N, 0);
    for (uint32_t k = 0; k < nchunks; k++) {
        const uint32_t a0 = ch_n0[k], a1 = ch_n0[k + 1];
        ext.clear();
        for (uint32_t e = noff[a0]; e < noff[a1]; e++)
            for (uint32_t g : {iAB[e].x, iAB[e].y})
                if (g < a0 || g >= a1) ext.push_back(g);
        std::sort(ext.begin(), ext.end());
        ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
        if ((a1 - a0) + ext.size() > HCAP) {
            g_last_error = "chunk halo over capacity";
            return BL_ERR_UNSUPPORTED;
        }
        for (uint32_t i = a0; i < a1; i++) slot[i] = i - a0;
        for (size_t j = 0; j < ext.size(); j++) slot[ext[j]] = (a1 - a0) + (uint32_t)j;
        for (uint32_t e = noff[a0]; e < noff[a1]; e++) iloc[e] = slot[iAB[e].x] | (slot[iAB[e].y] << 16);
        halo.insert(halo.end(), ext.begin(), ext.end());
        halo_off[k + 1] = (uint32_t)halo.size();
        std::vector<uint32_t> nb;
        for (uint32_t g : ext) nb.push_back(node_chunk[g]);
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        nbr.insert(nbr.end(), nb.begin(), nb.end());
        nbr_off[k + 1] = (uint32_t)nbr.size();
    }

    int rc;
    if ((rc = c->pm[0].upload(P)) || (rc = c->vm[0].upload(V)) || (rc = c->pm[1].alloc(N)) || (rc = c->vm[1].alloc(N)) ||
        (rc = c->ninfo.upload(ninfo)) || (rc = c->naux.upload(naux)) || (rc = c->noff.upload(noff)) ||
        (rc = c->iA.upload(iA)) || (rc = c->iM.upload(iM)) || (rc = c->iOth.upload(iOth)) || (rc = c->bcold.upload(bcold)) ||
        (rc = c->stress.upload(stress)) || (rc = c->ch_n0.upload(ch_n0)) || (rc = c->ch_rng.upload(ch_rng)) || (rc = c->fch_rng.upload(fch_rng)) || (rc = c->fwg.upload(fwg)) || (rc = c->wg_rng.upload(wg_rng)) || (rc = c->fwg_rng.upload(fwg_rng)) || (rc = c->body_c0.upload(body_c0)) ||
        (rc = c->bparam.upload(bp)) || (rc = c->dboxes.upload(c->boxes)) || (rc = c->dcyl.upload(c->cyls)) ||
        (c->tnx && ((rc = c->dth.upload(c->th)) || (rc = c->dts.upload(c->ts)) || (rc = c->dtmax.alloc(c->bodies.size())))) ||
        (rc = c->counters.alloc(CNT_N)) ||
        (rc = c->broken.alloc(std::max<size_t>(1, c->bodies.size()))) || (rc = c->bounds.alloc(8)) ||
        (rc = c->body_frame.alloc(std::max<size_t>(1, c->bodies.size()) * BF_N)) || (rc = c->bar.alloc(1)) ||
        (rc = c->err.alloc(1)) || (rc = c->chunk_body.upload(chunk_body)) ||
        (rc = c->cstat.alloc(std::max<size_t>(1, nchunks))) || (rc = c->halo_off.upload(halo_off)) ||
        (rc = c->halo.upload(halo)) || (rc = c->nbr_off.upload(nbr_off)) || (rc = c->nbr.upload(nbr)) ||
        (rc = c->iloc.upload(iloc)) || (rc = c->flags.alloc(std::max<size_t>(1, nchunks))) ||
        (rc = c->tri_abc.upload(tri_abc)) || (rc = c->tri_rest2.upload(tri_rest2)) || (rc = c->tri_torn.upload(tri_torn)) || (rc = c->tri_hull.upload(tri_hull)) ||
        (rc = c->cap_data.upload(cap_data)) || (rc = c->wg_chunk.upload(wg_chunk)) ||
        (rc = c->F.alloc(c->contact_geometry || has_extras(c) ? N : 0)) || (rc = c->wstat.alloc((N + 255) / 256 + 1)) ||
        (rc = c->cbox.alloc(std::max<size_t>(1, c->bodies.size()) * BF_N)))
        return rc;
    if (c->contact_geometry && (rc = c->contacts.setup(c->hv, N, NT, NC))) {
        g_last_error = "contact buffers: device allocation failed";
        return rc;
    }
    HIPCHK(hipMemset(c->counters.p, 0, sizeof(unsigned long long) * CNT_N));
    HIPCHK(hipMemset(c->broken.p, 0, sizeof(unsigned long long) * c->broken.n));
    HIPCHK(hipMemset(c->err.p, 0, sizeof(int)));
    c->contacts.err_word = c->err.p;
    HIPCHK(hipMemset(c->cstat.p, 0, sizeof(uint4) * c->cstat.n));
    HIPCHK(hipMemset(c->wstat.p, 0, sizeof(uint4) * c->wstat.n));
    KArgs& a = c->args;
    a = KArgs{};
    a.pm[0] = c->pm[0].p, a.pm[1] = c->pm[1].p, a.vm[0] = c->vm[0].p, a.vm[1] = c->vm[1].p;
    a.ninfo = c->ninfo.p, a.naux = c->naux.p, a.noff = c->noff.p;
    a.iA = c->iA.p, a.iM = c->iM.p, a.iOth = c->iOth.p, a.bcold = c->bcold.p, a.stress = c->stress.p;
    a.ch_n0 = c->ch_n0.p, a.ch_rng = c->ch_rng.p, a.body_c0 = c->body_c0.p, a.bodies = c->bparam.p, a.boxes = c->dboxes.p;
    a.counters = c->counters.p, a.broken = c->broken.p, a.body_frame = c->body_frame.p;
    a.cstat = c->cstat.p, a.chunk_body = c->chunk_body.p;
    a.halo_off = c->halo_off.p, a.halo = c->halo.p, a.nbr_off = c->nbr_off.p, a.nbr = c->nbr.p, a.iloc = c->iloc.p;
    a.F = c->F.p, a.tri_abc = c->tri_abc.p, a.tri_rest2 = c->tri_rest2.p, a.tri_torn = c->tri_torn.p, a.tri_hull = c->tri_hull.p;
    a.cap_data = c->cap_data.p, a.ntris = NT, a.ncaps = NC;
    a.g = mk(c->s.gravity[0], c->s.gravity[1], c->s.gravity[2]);
    a.dt = c->s.dt, a.kill_y = c->s.kill_y, a.nbox = (int)c->boxes.size();
    a.cyls = c->dcyl.p, a.ncyl = (int)c->cyls.size();
    if (c->tnx) {
        a.th = c->dth.p, a.ts = c->dts.p, a.tmax = c->dtmax.p;
        a.tnx = c->tnx, a.tnz = c->tnz, a.tcell = c->tcell, a.tox = c->tox, a.toz = c->toz;
    }
    a.nnodes = N, a.nchunks = nchunks, a.nbodies = (uint32_t)c->bodies.size();
    a.wg_chunk = c->wg_chunk.p;
    a.fch_rng = c->fch_rng.p, a.fwg = c->fwg.p, a.nfchunks = nf;
    a.wg_rng = c->wg_rng.p, a.fwg_rng = c->fwg_rng.p;
    c->parity = 0;
    // shocks and slide nodes: per body with any its ranges; per shock its beam's two incidences and nodes; slides
    // with global nodes and rails
    {
        std::vector<uint4> xb, sx;
        std::vector<bl_shock> sh;
        std::vector<bl_slide> sl;
        std::vector<uint32_t> rail, fn;
        std::vector<bl_joint> jt;
        std::vector<float4> fq4, fw4, ctl;
        std::vector<int2> fr2;
        std::vector<bl_wheel> wh;
        std::vector<uint32_t> wn;
        std::vector<uint2> ty;
        for (HostBody& hb : c->bodies) {
            hb.slide0 = (uint32_t)sl.size(), hb.frame0 = (uint32_t)fn.size(), hb.joint0 = (uint32_t)jt.size();
            hb.wheel0 = (uint32_t)wh.size(), hb.xi = -1;
            if (hb.shocks.empty() && hb.slides.empty() && hb.joints.empty() && hb.frames.empty() && hb.wheels.empty() &&
                !hb.tyres)
                continue;
            hb.xi = (int32_t)(xb.size() / 3);
            const uint32_t w0 = (uint32_t)wh.size(), t0 = (uint32_t)ty.size();
            for (size_t k = 0; k < hb.wheels.size(); k++) {
                bl_wheel y = hb.wheels[k];
                y.axle0 += hb.node0, y.axle1 += hb.node0;
                y.arm = y.arm >= 0 ? y.arm + (int32_t)hb.node0 : -1;
                y.near_attach = y.near_attach >= 0 ? y.near_attach + (int32_t)hb.node0 : -1;
                y.node0 = (uint32_t)wn.size();
                for (uint32_t j = 0; j < hb.wheels[k].n_nodes; j++) wn.push_back(hb.node0 + hb.wnodes[hb.wheels[k].node0 + j]);
                wh.push_back(y);
                ctl.push_back(hb.ctl[k]);
            }
            if (hb.tyres) {
                // (World::collide_static's node_wheel, src/phys/world.cpp:873-877: a later wheel listing a node takes it)
                std::vector<uint32_t> nw(hb.nodes.size(), 0xffffffffu);
                for (size_t k = 0; k < hb.wheels.size(); k++)
                    for (uint32_t j = 0; j < hb.wheels[k].n_nodes; j++) nw[hb.wnodes[hb.wheels[k].node0 + j]] = w0 + (uint32_t)k;
                for (size_t i = 0; i < hb.nodes.size(); i++)
                    if (hb.info[i].flags & BL_NF_TYRE) ty.push_back(make_uint2(hb.node0 + (uint32_t)i, nw[i]));
            }
            const uint32_t s0 = (uint32_t)sh.size(), l0 = (uint32_t)sl.size();
            const uint32_t j0 = (uint32_t)jt.size(), f0 = (uint32_t)fn.size();
            for (const bl_frame& f : hb.frames) {
                fn.push_back(hb.node0 + f.node);
                fq4.push_back(make_float4(f.q[0], f.q[1], f.q[2], f.q[3]));
                fw4.push_back(make_float4(f.w[0], f.w[1], f.w[2], f.inv_inertia));
                const int32_t nn = (int32_t)hb.nodes.size();
                const bool kin = f.ref_x >= 0 && f.ref_y >= 0 && f.ref_x < nn && f.ref_y < nn;
                fr2.push_back(kin ? make_int2((int32_t)hb.node0 + f.ref_x, (int32_t)hb.node0 + f.ref_y) : make_int2(-1, -1));
            }
            for (const bl_joint& j : hb.joints) {
                bl_joint y = j;
                y.parent_frame = f0 + j.parent_frame, y.child_node = hb.node0 + j.child_node;
                y.child_frame = j.child_frame >= 0 ? (int32_t)f0 + j.child_frame : -1;
                jt.push_back(y);
            }
            for (const bl_shock& x : hb.shocks) {
                const bl_beam& bm = hb.beams[x.beam];
                const uint32_t gb = hb.beam0 + x.beam;
                sh.push_back(x);
                sx.push_back(make_uint4(c->a_slot[gb], b_slot[gb], hb.node0 + bm.a, hb.node0 + bm.b));
            }
            for (const bl_slide& x : hb.slides) {
                bl_slide y = x;
                y.node = hb.node0 + x.node;
                y.rail0 = (uint32_t)rail.size() + x.rail0;
                sl.push_back(y);
            }
            for (uint32_t r : hb.rail) rail.push_back(hb.node0 + r);
            xb.push_back(make_uint4(s0, (uint32_t)sh.size(), l0, (uint32_t)sl.size()));
            xb.push_back(make_uint4(j0, (uint32_t)jt.size(), f0, (uint32_t)fn.size()));
            xb.push_back(make_uint4(w0, (uint32_t)wh.size(), t0, (uint32_t)ty.size()));
        }
        std::vector<float2> wsp(xb.size() / 3);
        for (const HostBody& hb : c->bodies)
            if (hb.xi >= 0) wsp[hb.xi] = make_float2(hb.wsp[0], hb.wsp[1]);
        if ((!wh.empty() && ((rc = c->xwh.upload(wh)) || (rc = c->xctl.upload(ctl)) || (rc = c->xtsum.alloc(wh.size())))) ||
            (!wn.empty() && (rc = c->xwn.upload(wn))) || (!wsp.empty() && (rc = c->xwsp.upload(wsp))) ||
            (!ty.empty() && ((rc = c->xty.upload(ty)) || (rc = c->xtneed.alloc(ty.size())) ||
                             (rc = c->xtslide.alloc(ty.size())))))
            return rc;
        if ((!xb.empty() && (rc = c->xbody.upload(xb))) || (!sx.empty() && (rc = c->shx.upload(sx))) ||
            (!sh.empty() && (rc = c->xsh.upload(sh))) || (!sl.empty() && (rc = c->xsl.upload(sl))) ||
            (!rail.empty() && (rc = c->xrail.upload(rail))) || (!jt.empty() && (rc = c->xjt.upload(jt))) ||
            (!fn.empty() && ((rc = c->fnode.upload(fn)) || (rc = c->fq.upload(fq4)) || (rc = c->fw.upload(fw4)) || (rc = c->fref.upload(fr2)) ||
                             (rc = c->ftq.upload(std::vector<float4>(fn.size(), make_float4(0, 0, 0, 0)))))))
            return rc;
        c->xargs = XArgs{c->xbody.p, (uint32_t)(xb.size() / 3), c->xsh.p, c->shx.p, c->xsl.p, c->xrail.p,
                         c->xjt.p, c->fq.p, c->fw.p, c->ftq.p, c->fnode.p, c->fref.p, (uint32_t)fn.size(),
                         c->xwh.p, c->xwn.p, c->xctl.p, c->xwsp.p, ty.empty() ? nullptr : c->xty.p, c->xtsum.p,
                         c->xtneed.p, c->xtslide.p, c->s.tyre_grip};
    }
    c->built_block = T, c->built_per = effective_per(c);
    c->built = true;
    return BL_OK;
}

int ensure_built(bl_ctx* c) {
    if (c->built) return BL_OK;
    if (c->bodies.empty()) {
        g_last_error = "no body added";
        return BL_ERR_STATE;
    }
    return build(c);
}

// AUTO: the halo or resident form when the chip holds every chunk's workgroup at once (and the runtime has not
// refused a cooperative launch), else one launch per substep.
int resolved_mode(const bl_ctx* c) {
    if (c->tune.mode != BL_MODE_AUTO) return c->tune.mode;
    if (c->no_coop) return BL_MODE_STEPPED;
    const int T = effective_block(c), per = effective_per(c);
    bl_tuning h{BL_MODE_HALO, T, 0, 0, per};
    if (tuning_valid(h) && halo_fits(T, per, c->args.nchunks)) return BL_MODE_HALO;
    bl_tuning t{BL_MODE_RESIDENT, T, 0, 0, per};
    if (tuning_valid(t) && resident_fits(T, per, c->args.nchunks)) return BL_MODE_RESIDENT;
    return BL_MODE_STEPPED;
}

int enqueue(bl_ctx* c, int substeps) {
    if (substeps <= 0) return BL_OK;
    bl_tuning t = c->tune;
    t.block = effective_block(c);
    t.incidences_per_thread = effective_per(c);
    t.mode = resolved_mode(c);
    if (t.block != c->built_block || t.incidences_per_thread != c->built_per) {
        g_last_error = "execution form differs from the chunking the world was built for";
        return BL_ERR_STATE;
    }
    int rc;
    if (t.mode == BL_MODE_HALO) {
        if (!halo_fits(t.block, t.incidences_per_thread, c->args.nchunks)) {
            g_last_error = "the halo form needs every chunk's workgroup resident at once: too many chunks";
            return BL_ERR_UNSUPPORTED;
        }
        const int per = t.substeps_per_launch > 0 ? t.substeps_per_launch : substeps;
        for (int s = 0; s < substeps; s += per) {
            const int k = std::min(per, substeps - s);
            if ((rc = launch_halo(c->args, c->stream, c->parity, k, t, c->flags.p, c->err.p, kSpinLimit))) {
                g_last_error = std::string("cooperative launch of the halo form failed: ") + coop_error_name();
                if (c->tune.mode != BL_MODE_AUTO) return rc;
                // (the runtime may refuse a cooperative launch, e.g. beside another process on the device: the
                // automatic choice then takes the stepped form, the same arithmetic, for the rest)
                (void)hipGetLastError();
                c->no_coop = true;
                return enqueue(c, substeps - s);
            }
            c->parity ^= (k & 1);
        }
    } else if (t.mode == BL_MODE_RESIDENT) {
        if (!resident_fits(t.block, t.incidences_per_thread, c->args.nchunks)) {
            g_last_error = "the resident form needs every chunk's workgroup resident at once: too many chunks";
            return BL_ERR_UNSUPPORTED;
        }
        const int per = t.substeps_per_launch > 0 ? t.substeps_per_launch : substeps;
        for (int s = 0; s < substeps; s += per) {
            const int k = std::min(per, substeps - s);
            if ((rc = launch_resident(c->args, c->stream, c->parity, k, t, c->bar.p, c->err.p, kSpinLimit))) {
                g_last_error = std::string("cooperative launch of the resident form failed: ") + coop_error_name();
                if (c->tune.mode != BL_MODE_AUTO) return rc;
                (void)hipGetLastError();
                c->no_coop = true;
                return enqueue(c, substeps - s);
            }
            c->parity ^= (k & 1);
        }
    } else if (t.mode == BL_MODE_PERSISTENT) {
        const int per = t.substeps_per_launch > 0 ? t.substeps_per_launch : 64;
        for (int s = 0; s < substeps; s += per) {
            const int k = std::min(per, substeps - s);
            if ((rc = launch_substeps(c->args, c->stream, c->parity, k, t))) return rc;
            c->parity ^= (k & 1);
        }
    } else if (t.use_graph) {
        const int per = t.substeps_per_launch > 0 ? t.substeps_per_launch : 32;
        for (int s = 0; s < substeps; s += per) {
            const int k = std::min(per, substeps - s);
            auto key = std::make_pair(k, c->parity);
            auto it = c->graphs.find(key);
            if (it == c->graphs.end()) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                rc = launch_substeps(c->args, c->stream, c->parity, k, t);
                hipError_t ce = hipStreamEndCapture(c->stream, &g);
                if (rc) return rc;
                HIPCHK(ce);
                hipGraphExec_t ex;
                HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                it = c->graphs.emplace(key, ex).first;
            }
            HIPCHK(hipGraphLaunch(it->second, c->stream));
            c->parity ^= (k & 1);
        }
    } else {
        if ((rc = launch_substeps(c->args, c->stream, c->parity, substeps, t))) return rc;
        c->parity ^= (substeps & 1);
    }
    c->substeps += substeps;
    return BL_OK;
}

// World::repair_nodes (src/phys/world.cpp:4022-4111), beams only: bad nodes (not finite, faster than 360 m/s,
// further than 300 m from the median) move to the mean of their good neighbours over up to 8 passes, leftovers
// rest at the median. Returns the nodes reset, or -1 when more than half of the body is bad.
int repair_body(HostBody& b) {
    const size_t n = b.nodes.size();
    std::vector<float> xs, ys, zs;
    for (const bl_node& x : b.nodes)
        if (std::isfinite(x.p[0]) && std::isfinite(x.p[1]) && std::isfinite(x.p[2]))
            xs.push_back(x.p[0]), ys.push_back(x.p[1]), zs.push_back(x.p[2]);
    if (xs.empty()) return -1;
    auto median = [](std::vector<float>& v) {
        std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
        return v[v.size() / 2];
    };
    const f3 med = mk(median(xs), median(ys), median(zs));
    const float reach = 300.0f, vmax = 0.9f * 400.0f;
    std::vector<uint8_t> bad(n, 0);
    int nbad = 0;
    for (size_t i = 0; i < n; i++) {
        const bl_node& x = b.nodes[i];
        const bool finite = std::isfinite(x.p[0]) && std::isfinite(x.p[1]) && std::isfinite(x.p[2]) &&
                            std::isfinite(x.v[0]) && std::isfinite(x.v[1]) && std::isfinite(x.v[2]);
        const f3 v = mk(x.v[0], x.v[1], x.v[2]);
        if (!finite || dot(v, v) > vmax * vmax || length(sub(mk(x.p[0], x.p[1], x.p[2]), med)) > reach) bad[i] = 1, nbad++;
    }
    if (nbad == 0) return 0;
    if (nbad * 2 > (int)n) return -1;
    const int reset = nbad;
    for (int pass = 0; pass < 8 && nbad > 0; pass++) {
        std::vector<f3> sp(n, mk(0, 0, 0)), sv(n, mk(0, 0, 0));
        std::vector<int> cnt(n, 0);
        auto link = [&](uint32_t x, uint32_t y) {
            const bl_node &nx = b.nodes[x], &ny = b.nodes[y];
            if (bad[x] && !bad[y]) {
                sp[x] = add(sp[x], mk(ny.p[0], ny.p[1], ny.p[2])), sv[x] = add(sv[x], mk(ny.v[0], ny.v[1], ny.v[2])), cnt[x]++;
            } else if (bad[y] && !bad[x]) {
                sp[y] = add(sp[y], mk(nx.p[0], nx.p[1], nx.p[2])), sv[y] = add(sv[y], mk(nx.v[0], nx.v[1], nx.v[2])), cnt[y]++;
            }
        };
        for (const bl_beam& bm : b.beams)
            if (!(bm.flags & BL_BF_BROKEN)) link(bm.a, bm.b);
        for (size_t i = 0; i < n; i++)
            if (bad[i] && cnt[i] > 0) {
                const f3 p = divs(sp[i], (float)cnt[i]), v = divs(sv[i], (float)cnt[i]);
                b.nodes[i].p[0] = p.x, b.nodes[i].p[1] = p.y, b.nodes[i].p[2] = p.z;
                b.nodes[i].v[0] = v.x, b.nodes[i].v[1] = v.y, b.nodes[i].v[2] = v.z;
                bad[i] = 0;
                nbad--;
            }
    }
    for (size_t i = 0; i < n; i++)
        if (bad[i]) {
            b.nodes[i].p[0] = med.x, b.nodes[i].p[1] = med.y, b.nodes[i].p[2] = med.z;
            b.nodes[i].v[0] = b.nodes[i].v[1] = b.nodes[i].v[2] = 0;
        }
    for (bl_frame& f : b.frames) f.w[0] = f.w[1] = f.w[2] = 0;
    return reset;
}

// The substeps from a pair rebuild up to the next: the first one's narrow phase, response and integration, then
// `len - 1` whole substeps, then (`tail`) the force phase of the substep whose rebuild comes next.
int enqueue_segment(bl_ctx* c, const bl_tuning& t, const PArgs& p, uint32_t nblk, int len, int parity, int tail) {
    const KArgs& a = c->args;
    int rc;
    for (int q = 0; q < len; q++) {
        if (q > 0 && t.block == 256) {
            if ((rc = launch_forces_narrow(a, c->stream, parity, t, p, nblk))) return rc;
            if ((rc = launch_extras(a, c->xargs, c->stream, parity))) return rc;
            if ((rc = c->contacts.launch_response(a, p, nblk, c->stream, parity))) return rc;
        } else {
            if (q > 0 && ((rc = launch_forces(a, c->stream, parity, t)) || (rc = launch_extras(a, c->xargs, c->stream, parity))))
                return rc;
            if ((rc = c->contacts.launch_collide(a, p, nblk, c->stream, parity))) return rc;
        }
        if ((rc = launch_tyres(a, c->xargs, c->stream, parity)) || (rc = launch_integrate(a, c->stream, parity, c->wstat.p)) ||
            (rc = launch_frames(a, c->xargs, c->stream, parity ^ 1)))
            return rc;
        parity ^= 1;
    }
    if (tail && ((rc = launch_forces(a, c->stream, parity, t)) || (rc = launch_extras(a, c->xargs, c->stream, parity))))
        return rc;
    return BL_OK;
}

// One segment as a graph launch: captured once per (length, parity, tail) for the lists' buffers of the moment;
// the narrow phase covers the lists' capacities and reads the rebuild's counts on the device.
int launch_segment(bl_ctx* c, const bl_tuning& t, int len, int tail) {
    PArgs p;
    uint32_t nblk;
    int rc;
    if ((rc = c->contacts.collide_fixed(p, nblk))) return rc;
    if (nblk != c->seg_nblk || std::memcmp(&p, &c->seg_p, sizeof(PArgs)) != 0) {
        c->drop_segments();
        c->seg_p = p, c->seg_nblk = nblk;
    }
    hipGraphExec_t ex = nullptr;
    for (const auto& g : c->segs)
        if (g.len == len && g.parity == c->parity && g.tail == tail) ex = g.ex;
    if (!ex) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_segment(c, t, p, nblk, len, c->parity, tail);
        const hipError_t ce = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        HIPCHK(ce);
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
        c->segs.push_back({len, c->parity, tail, ex});
    }
    HIPCHK(hipGraphLaunch(ex, c->stream));
    c->contacts.count_tests(len);
    c->parity ^= (len & 1);
    return BL_OK;
}

// A frame with inter-body contacts (World::simulate_island, src/phys/world.cpp:3211-3591): per substep the forces
// into F, the pair lists rebuilt at the first substep and every `interval` after (from the state at that point),
// the narrow phase and the response on F, then the static contacts and the integration. The host waits at each
// rebuild (list sizes) - about one substep in 17 at the default 120 Hz pair rate. With the graph form the substeps
// between two rebuilds go as one graph launch; `kev` (per-launch events around the force phase) takes the
// launches one by one.
// The frame just enqueued is pending its post-frame check; its error word and body frame words are copied into
// pinned memory behind it.
int frame_enqueued(bl_ctx* c) {
    c->frame_pending = true;
    c->fin_queued = false;
    const size_t nb = c->bodies.size();
    if (!c->fin.ensure(256 + sizeof(unsigned int) * nb * BF_N)) return BL_ERR_HIP;
    HIPCHK(hipMemcpyAsync(c->fin.p, c->err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (nb)
        HIPCHK(hipMemcpyAsync((char*)c->fin.p + 256, c->body_frame.p, sizeof(unsigned int) * nb * BF_N,
                              hipMemcpyDeviceToHost, c->stream));
    c->fin_queued = true;
    return BL_OK;
}

int contact_frame(bl_ctx* c, int substeps, const hipEvent_t* kev = nullptr) {
    bl_tuning t = c->tune;
    t.block = c->built_block, t.incidences_per_thread = c->built_per;
    const double hz = (double)smax(1.0f, c->s.collision_hz);
    int interval = (int)std::ceil(1.0 / (hz * (double)c->s.dt) - 1e-6);
    if (interval < 1) interval = 1;
    c->contacts.interval = interval;
    const size_t nb = c->bodies.size();
    KArgs& a = c->args;
    KArgs ab = a;
    ab.body_frame = c->cbox.p;
    int rc;
    if ((rc = c->contacts.begin_frame(c->stream))) return rc;
    // the bounds the first rebuild reads: as the last frame left them
    if (c->w0_ev) HIPCHK(hipEventSynchronize(c->w0_ev)); // (the last upload from the staging buffer done)
    else HIPCHK(hipEventCreateWithFlags(&c->w0_ev, hipEventDisableTiming));
    if (!c->w0.ensure(sizeof(unsigned int) * nb * BF_N + 4)) return BL_ERR_HIP;
    unsigned int* w0 = (unsigned int*)c->w0.p;
    std::memset(w0, 0, sizeof(unsigned int) * nb * BF_N);
    for (size_t k = 0; k < nb; k++) {
        const HostBody& hb = c->bodies[k];
        const float mn[3] = {hb.mn.x, hb.mn.y, hb.mn.z}, mx[3] = {hb.mx.x, hb.mx.y, hb.mx.z};
        for (int q = 0; q < 3; q++) w0[k * BF_N + BF_MN + q] = fkey(mn[q]), w0[k * BF_N + BF_MX + q] = fkey(mx[q]);
    }
    HIPCHK(hipMemcpyAsync(c->cbox.p, w0, sizeof(unsigned int) * nb * BF_N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->w0_ev, c->stream));
    if ((rc = launch_frame_begin(a, c->stream, c->parity))) return rc;
    const bool graph = !kev && t.use_graph;
    if (kev) HIPCHK(hipEventRecord(kev[0], c->stream));
    if ((rc = launch_forces(a, c->stream, c->parity, t))) return rc;
    if (kev) HIPCHK(hipEventRecord(kev[1], c->stream));
    if ((rc = launch_extras(a, c->xargs, c->stream, c->parity))) return rc;
    for (int s = 0; s < substeps;) {
        // each body's top speed so far (the frame's start included) and, after the first substep, its bounds now
        if (s > 0 && (rc = launch_box(ab, c->stream, c->parity))) return rc;
        if ((rc = c->contacts.rebuild(a, c->cbox.p, c->stream, c->parity))) {
            g_last_error = "pair rebuild failed";
            return rc;
        }
        if (c->contacts.last.unsupported & 4) {
            g_last_error = "a scan of the pair rebuild timed out waiting for its predecessors";
            return BL_ERR_HIP;
        }
        if (c->contacts.last.unsupported) {
            g_last_error = "a capsule meets more than 256 triangles in one pair rebuild";
            return BL_ERR_UNSUPPORTED;
        }
        const int next = std::min(s + interval, substeps);
        if (graph) {
            if ((rc = launch_segment(c, t, next - s, next < substeps))) {
                g_last_error = "contact segment capture or launch failed";
                return rc;
            }
            s = next;
            continue;
        }
        // (after the segment's first substep the narrow phase runs in the force phase's launch - it reads the
        // positions and the lists only - unless the force phase is timed alone)
        PArgs np{};
        uint32_t nnb = 0;
        for (; s < next; s++) {
            rc = nnb ? c->contacts.launch_response(a, np, nnb, c->stream, c->parity) : c->contacts.collide(a, c->stream, c->parity);
            if (rc) {
                g_last_error = "contact narrow phase or response failed";
                return rc;
            }
            if ((rc = launch_tyres(a, c->xargs, c->stream, c->parity)) ||
                (rc = launch_integrate(a, c->stream, c->parity, c->wstat.p)) ||
                (rc = launch_frames(a, c->xargs, c->stream, c->parity ^ 1)))
                return rc;
            c->parity ^= 1;
            nnb = 0;
            if (s + 1 < substeps) {
                if (!kev && s + 1 < next && t.block == 256) {
                    if ((rc = c->contacts.narrow_args(np, nnb))) return rc;
                    if (nnb) {
                        c->contacts.count_tests(1);
                        if ((rc = launch_forces_narrow(a, c->stream, c->parity, t, np, nnb)) ||
                            (rc = launch_extras(a, c->xargs, c->stream, c->parity)))
                            return rc;
                        continue;
                    }
                }
                if (kev) HIPCHK(hipEventRecord(kev[2 * (s + 1)], c->stream));
                if ((rc = launch_forces(a, c->stream, c->parity, t))) return rc;
                if (kev) HIPCHK(hipEventRecord(kev[2 * (s + 1) + 1], c->stream));
                if ((rc = launch_extras(a, c->xargs, c->stream, c->parity))) return rc;
            }
        }
    }
    c->substeps += substeps;
    if ((rc = launch_wstat_fold(a, c->stream, c->wstat.p)) || (rc = launch_frame_end(a, c->stream, c->parity))) return rc;
    return frame_enqueued(c);
}

// Whether this frame has an island that searches for pairs (World::build_islands on the bodies as the last frame
// left them).
bool needs_contacts(bl_ctx* c, int substeps) {
    if (!c->contact_geometry || !c->s.inter_body_collisions) return false;
    for (size_t k = 0; k < c->bodies.size(); k++) {
        c->hv.max_speed[k] = c->bodies[k].max_speed;
        c->hv.mn[k] = c->bodies[k].mn, c->hv.mx[k] = c->bodies[k].mx;
    }
    c->contacts.build_islands(c->hv, (float)substeps * c->s.dt);
    return c->contacts.npair_islands() > 0;
}

// The split form without contacts, for worlds with shocks or slide nodes (their forces join the beams' between the
// force phase and the integration): per substep the force phase into F, the shocks and slides, the static contacts
// and the integration; up to 32 substeps a captured graph.
int enqueue_split(bl_ctx* c, int substeps) {
    if (c->F.n < c->args.nnodes || c->wstat.n < (c->args.nnodes + 255) / 256) {
        g_last_error = "internal: the split form's force and count buffers are not allocated";
        return BL_ERR_STATE;
    }
    bl_tuning t = c->tune;
    t.block = c->built_block, t.incidences_per_thread = c->built_per;
    const KArgs& a = c->args;
    int rc;
    auto one = [&](int parity) -> int {
        if ((rc = launch_forces(a, c->stream, parity, t)) || (rc = launch_extras(a, c->xargs, c->stream, parity)) ||
            (rc = launch_tyres(a, c->xargs, c->stream, parity)) || (rc = launch_integrate(a, c->stream, parity, c->wstat.p)) ||
            (rc = launch_frames(a, c->xargs, c->stream, parity ^ 1)))
            return rc;
        return BL_OK;
    };
    for (int s = 0; s < substeps; s += 32) {
        const int k = std::min(32, substeps - s);
        if (!t.use_graph) {
            for (int q = 0; q < k; q++)
                if ((rc = one(c->parity ^ (q & 1)))) return rc;
        } else {
            auto key = std::make_pair(k, c->parity);
            auto it = c->sgraphs.find(key);
            if (it == c->sgraphs.end()) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int r2 = BL_OK;
                for (int q = 0; q < k && !r2; q++) r2 = one(c->parity ^ (q & 1));
                const hipError_t ce = hipStreamEndCapture(c->stream, &g);
                if (r2) return r2;
                HIPCHK(ce);
                hipGraphExec_t ex;
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHK(ie);
                it = c->sgraphs.emplace(key, ex).first;
            }
            HIPCHK(hipGraphLaunch(it->second, c->stream));
        }
        c->parity ^= (k & 1);
    }
    c->substeps += substeps;
    return launch_wstat_fold(a, c->stream, c->wstat.p);
}

// Enqueues one frame: the frame-start speeds, the substeps, the frame-end bounds.
void motion_energy(const HostBody& hb, const float g[3], double& ke, double& pe);

// Heightfield::max_height (src/phys/static_world.cpp:220-231) over the terrain's vertex blocks
float terrain_max_height(const bl_ctx* c, float x0, float z0, float x1, float z1) {
    int bx0 = (int)std::floor((x0 - c->tox) / c->tcell) / 16, bx1 = (int)std::floor((x1 - c->tox) / c->tcell) / 16;
    int bz0 = (int)std::floor((z0 - c->toz) / c->tcell) / 16, bz1 = (int)std::floor((z1 - c->toz) / c->tcell) / 16;
    bx0 = std::max(0, bx0), bz0 = std::max(0, bz0);
    bx1 = std::min(c->tbx - 1, bx1), bz1 = std::min(c->tbz - 1, bz1);
    float m = -1e30f;
    for (int bz = bz0; bz <= bz1; bz++)
        for (int bx = bx0; bx <= bx1; bx++) m = smax(m, c->tblock[(size_t)bz * c->tbx + bx]);
    return m;
}

int enqueue_frame(bl_ctx* c, int substeps) {
    if (substeps <= 0) return BL_OK;
    int rc;
    // (the energy guard's start: src/phys/world.cpp:3236; the guarded bodies' host nodes are current)
    for (HostBody& hb : c->bodies)
        if (hb.opts.energy_guard) motion_energy(hb, c->s.gravity, hb.guard_ke, hb.guard_pe);
    if (c->tnx) {
        // each body's terrain top over its reach this frame: its bounds widened by twice what its top speed covers
        // in the frame and half a metre (World::build_islands, src/phys/world.cpp:743-756)
        const float ft = (float)substeps * c->s.dt;
        c->htmax.resize(c->bodies.size());
        for (size_t k = 0; k < c->bodies.size(); k++) {
            const HostBody& hb = c->bodies[k];
            const float m = smin(hb.max_speed, 200.0f) * ft * 2.0f + 0.5f;
            c->htmax[k] = terrain_max_height(c, hb.mn.x - m, hb.mn.z - m, hb.mx.x + m, hb.mx.z + m);
        }
        HIPCHK(hipMemcpyAsync(c->dtmax.p, c->htmax.data(), sizeof(float) * c->htmax.size(), hipMemcpyHostToDevice, c->stream));
    }
    if (needs_contacts(c, substeps)) return contact_frame(c, substeps);
    if ((rc = launch_frame_begin(c->args, c->stream, c->parity))) return rc;
    if ((rc = c->xargs.nbody ? enqueue_split(c, substeps) : enqueue(c, substeps))) return rc;
    if ((rc = launch_frame_end(c->args, c->stream, c->parity))) return rc;
    return frame_enqueued(c);
}

// The post-frame check of World::simulate_island (src/phys/world.cpp:3962-3979) on the frame just run: a body whose
// bounds are not finite or wider than 800 m, or whose top speed passed 360 m/s, is repaired on the host, or frozen
// when it cannot be.
// SoftBody::motion_energy (src/phys/softbody.cpp:798-803), the nodes' part: in double, in node order
void motion_energy(const HostBody& hb, const float g[3], double& ke, double& pe) {
    ke = pe = 0;
    for (const bl_node& x : hb.nodes)
        if (x.inv_mass > 0) {
            ke += 0.5 * x.mass * (x.v[0] * x.v[0] + x.v[1] * x.v[1] + x.v[2] * x.v[2]);
            pe -= x.mass * (double)(g[0] * x.p[0] + g[1] * x.p[1] + g[2] * x.p[2]);
        }
}

// The energy guard after a frame (World::simulate_island, src/phys/world.cpp:3995-3996; SoftBody::guard_energy,
// src/phys/softbody.cpp:805-815): a body quiet for more than 15 frames (no collision volume or sphere contact - none
// here) that gained more than 2 J and 5 % of its kinetic energy at the frame's start has its velocities scaled back.
// The body's nodes on the host are current. True when it scaled them.
bool guard_body(bl_ctx* c, HostBody& hb) {
    hb.guard_quiet++;
    if (!hb.opts.energy_guard || hb.guard_quiet <= 15) return false;
    double ke, pe;
    motion_energy(hb, c->s.gravity, ke, pe);
    const double excess = (ke + pe) - (hb.guard_ke + hb.guard_pe) - (2.0 + 0.05 * hb.guard_ke);
    if (!(excess > 0) || !(ke > 0)) return false;
    const float s = (float)std::sqrt(std::max(0.0, ke - excess) / ke);
    for (bl_node& x : hb.nodes) x.v[0] *= s, x.v[1] *= s, x.v[2] *= s;
    hb.guard_cuts++;
    return true;
}

// A body's nodes from the device's current state into the host copy, or back (p, inv_mass; v, mass)
int body_nodes(bl_ctx* c, HostBody& hb, bool to_device) {
    const size_t n = hb.nodes.size();
    if (!n) return BL_OK;
    std::vector<float4> P(n), V(n);
    if (to_device) {
        for (size_t i = 0; i < n; i++) {
            const bl_node& x = hb.nodes[i];
            P[i] = make_float4(x.p[0], x.p[1], x.p[2], x.inv_mass), V[i] = make_float4(x.v[0], x.v[1], x.v[2], x.mass);
        }
        HIPCHK(hipMemcpy(c->pm[c->parity].p + hb.node0, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->vm[c->parity].p + hb.node0, V.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        return BL_OK;
    }
    HIPCHK(hipMemcpy(P.data(), c->pm[c->parity].p + hb.node0, sizeof(float4) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(V.data(), c->vm[c->parity].p + hb.node0, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        bl_node& x = hb.nodes[i];
        x.p[0] = P[i].x, x.p[1] = P[i].y, x.p[2] = P[i].z, x.inv_mass = P[i].w;
        x.v[0] = V[i].x, x.v[1] = V[i].y, x.v[2] = V[i].z, x.mass = V[i].w;
    }
    return BL_OK;
}

int post_frame(bl_ctx* c) {
    if (!c->frame_pending) return BL_OK;
    c->frame_pending = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t nb = c->bodies.size();
    const bool queued = c->fin_queued;
    c->fin_queued = false;
    int errw = 0;
    if (queued) std::memcpy(&errw, c->fin.p, sizeof(int));
    else HIPCHK(hipMemcpy(&errw, c->err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (errw) {
        HIPCHK(hipMemset(c->err.p, 0, sizeof(int)));
    HIPCHK(hipMemset(c->cstat.p, 0, sizeof(uint4) * c->cstat.n));
        g_last_error = (errw & 2) ? "a contact response's dependency wait gave up: the state of that frame is not valid"
                                  : "a resident launch's grid barrier timed out: the state of that frame is not valid";
        return BL_ERR_HIP;
    }
    std::vector<unsigned int> w(nb * BF_N);
    if (queued) {
        if (nb) std::memcpy(w.data(), (const char*)c->fin.p + 256, sizeof(unsigned int) * w.size());
    } else {
        HIPCHK(hipMemcpy(w.data(), c->body_frame.p, sizeof(unsigned int) * w.size(), hipMemcpyDeviceToHost));
    }
    std::vector<size_t> trouble;
    const float lim = 400.0f * 0.9f;
    for (size_t k = 0; k < nb; k++) {
        HostBody& hb = c->bodies[k];
        if (hb.nodes.empty()) continue;
        const unsigned int* x = &w[k * BF_N];
        const f3 mn = mk(fkey_inv(x[BF_MN]), fkey_inv(x[BF_MN + 1]), fkey_inv(x[BF_MN + 2]));
        const f3 mx = mk(fkey_inv(x[BF_MX]), fkey_inv(x[BF_MX + 1]), fkey_inv(x[BF_MX + 2]));
        const f3 e = sub(mx, mn);
        const float ms = std::sqrt(u2f(x[BF_V2]));
        hb.max_speed = ms, hb.mn = mn, hb.mx = mx; // (the next frame's islands start from these)
        if (!std::isfinite(mn.x) || !std::isfinite(mx.y) || smax(e.x, smax(e.y, e.z)) > 800.0f || ms > lim)
            trouble.push_back(k);
    }
    int rc;
    if (trouble.empty()) {
        // (the guarded bodies' nodes come back each frame: the next frame's start needs their energy too)
        for (HostBody& hb : c->bodies)
            if (hb.opts.energy_guard) {
                if ((rc = body_nodes(c, hb, false))) return rc;
                if (guard_body(c, hb) && (rc = body_nodes(c, hb, true))) return rc;
            } else {
                hb.guard_quiet++;
            }
        return BL_OK;
    }
    if ((rc = pull_state(c))) return rc;
    std::vector<char> frozen(nb, 0);
    for (size_t k : trouble) {
        HostBody& hb = c->bodies[k];
        const int fixed = repair_body(hb);
        frozen[k] = fixed < 0;
        if (fixed > 0) {
            // bounds and top speed of the repaired body (src/phys/world.cpp:4106-4109)
            hb.compute_bounds();
            hb.max_speed = 0;
            for (const bl_node& x : hb.nodes) hb.max_speed = smax(hb.max_speed, length(mk(x.v[0], x.v[1], x.v[2])));
        }
        if (fixed < 0) {
            for (bl_node& x : hb.nodes) {
                if (!std::isfinite(x.p[0]) || !std::isfinite(x.p[1]) || !std::isfinite(x.p[2])) x.p[0] = 0, x.p[1] = -1000, x.p[2] = 0;
                x.v[0] = x.v[1] = x.v[2] = 0;
                x.inv_mass = 0;
            }
            hb.compute_bounds();
        }
        c->repairs++;
    }
    // (a frozen body skips the rest of the frame's end: src/phys/world.cpp:3967-3976)
    for (size_t k = 0; k < nb; k++)
        if (!frozen[k]) guard_body(c, c->bodies[k]);
    // the repaired state goes back; the counters survive the rebuild
    std::vector<unsigned long long> cnt(CNT_N, 0), brk(nb, 0);
    HIPCHK(hipMemcpy(cnt.data(), c->counters.p, sizeof(unsigned long long) * CNT_N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(brk.data(), c->broken.p, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost));
    c->release();
    if ((rc = build(c))) return rc;
    HIPCHK(hipMemcpy(c->counters.p, cnt.data(), sizeof(unsigned long long) * CNT_N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->broken.p, brk.data(), sizeof(unsigned long long) * nb, hipMemcpyHostToDevice));
    return BL_OK;
}

} // namespace

void blg::set_last_error(const char* msg) { g_last_error = msg; }

extern "C" {

int bl_abi_version(void) { return BL_ABI_VERSION; }

const char* bl_error_string(int status) {
    if (!g_last_error.empty() && status != BL_OK) return g_last_error.c_str();
    switch (status) {
    case BL_OK: return "ok";
    case BL_ERR_ARG: return "invalid argument";
    case BL_ERR_HIP: return "HIP runtime error";
    case BL_ERR_UNSUPPORTED: return "unsupported feature";
    case BL_ERR_STATE: return "call out of order";
    default: return "unknown status";
    }
}

void bl_default_settings(bl_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->gravity[1] = -9.81f;
    s->dt = 0.0005f;
    s->collision_hz = 120.0f;
    s->frame_theta = 0.5f;
    s->frame_dissipation = 0.02f;
    s->inter_body_collisions = 1;
    s->sleeping = 0;
    s->kill_y = -200.0f;
    s->tyre_grip = 1.0f;
}

void bl_default_body_opts(bl_body_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->collision_radius = 0.05f;
    o->ground_friction = 1.0f;
    o->bounce = -1.0f;
    o->contact_push_max = 2.0f;
    o->contact_slop = 0.0f;
    o->hull_depth = 0.3f;
    o->face_skin = 0.002f;
    o->contact_friction = 1.0f;
    o->allow_break = 1;
    o->allow_deform = 1;
}

// SoftBody::finalize (src/phys/softbody.cpp:76-83): L = L0 = max(1e-3, |pa - pb|), triangles' longest rest edge.
int bl_finalize(const bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb, bl_triangle* tris, uint32_t nt) {
    if ((nb && (!nodes || !beams)) || (nt && (!nodes || !tris))) return BL_ERR_ARG;
    for (uint32_t j = 0; j < nb; j++) {
        if (beams[j].a >= n || beams[j].b >= n) return BL_ERR_ARG;
        const bl_node &pa = nodes[beams[j].a], &pb = nodes[beams[j].b];
        const float d = length(sub(mk(pa.p[0], pa.p[1], pa.p[2]), mk(pb.p[0], pb.p[1], pb.p[2])));
        beams[j].L = beams[j].L0 = smax(1e-3f, d);
    }
    for (uint32_t t = 0; t < nt; t++) {
        if (tris[t].a >= n || tris[t].b >= n || tris[t].c >= n) return BL_ERR_ARG;
        const f3 a = mk(nodes[tris[t].a].p[0], nodes[tris[t].a].p[1], nodes[tris[t].a].p[2]);
        const f3 b = mk(nodes[tris[t].b].p[0], nodes[tris[t].b].p[1], nodes[tris[t].b].p[2]);
        const f3 cc = mk(nodes[tris[t].c].p[0], nodes[tris[t].c].p[1], nodes[tris[t].c].p[2]);
        const f3 ba = sub(b, a), cb = sub(cc, b), ac = sub(a, cc);
        // max_edge2 (src/phys/softbody.h:154)
        tris[t].rest_edge2 = smax(dot(ba, ba), smax(dot(cb, cb), dot(ac, ac)));
    }
    return BL_OK;
}

// SoftBody::stabilize's beam part (src/phys/softbody.cpp:85-110).
int bl_stabilize(const bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb, float dt, float k_budget,
                 float d_budget, int32_t* changed) {
    if (nb && (!nodes || !beams)) return BL_ERR_ARG;
    std::vector<float> ksum(n, 0.0f), dsum(n, 0.0f), ks(n, 1.0f), ds(n, 1.0f);
    for (uint32_t j = 0; j < nb; j++) {
        if (beams[j].a >= n || beams[j].b >= n) return BL_ERR_ARG;
        ksum[beams[j].a] += beams[j].k;
        ksum[beams[j].b] += beams[j].k;
        dsum[beams[j].a] += beams[j].d;
        dsum[beams[j].b] += beams[j].d;
    }
    for (uint32_t i = 0; i < n; i++) {
        const float m = nodes[i].mass;
        if (m <= 0) continue;
        const float kl = ksum[i] * dt * dt / m;
        if (kl > k_budget) ks[i] = k_budget / kl;
        const float dl = dsum[i] * dt / m;
        if (dl > d_budget) ds[i] = d_budget / dl;
    }
    int32_t ch = 0;
    for (uint32_t j = 0; j < nb; j++) {
        const float a = smin(ks[beams[j].a], ks[beams[j].b]);
        const float d = smin(ds[beams[j].a], ds[beams[j].b]);
        if (a < 1.0f || d < 1.0f) ch++;
        beams[j].k *= a;
        beams[j].d *= d;
    }
    if (changed) *changed = ch;
    return BL_OK;
}

// StaticWorld::add_box (src/phys/static_world.cpp:298-306): to_mat3 (src/core/math.h:163-170) and
// StaticBox::update_aabb (src/phys/static_world.cpp:242-246).
int bl_box_from_quat(const float center[3], const float half[3], const float q[4], uint32_t surface, bl_box* out) {
    if (!center || !half || !q || !out) return BL_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
    const float m[9] = {1 - 2 * (yy + zz), 2 * (xy + wz), 2 * (xz - wy), 2 * (xy - wz), 1 - 2 * (xx + zz),
                        2 * (yz + wx), 2 * (xz + wy), 2 * (yz - wx), 1 - 2 * (xx + yy)};
    std::memcpy(out->rot, m, sizeof(m));
    const f3 e = add(add(mul(mk(std::fabs(m[0]), std::fabs(m[1]), std::fabs(m[2])), half[0]),
                         mul(mk(std::fabs(m[3]), std::fabs(m[4]), std::fabs(m[5])), half[1])),
                     mul(mk(std::fabs(m[6]), std::fabs(m[7]), std::fabs(m[8])), half[2]));
    for (int k = 0; k < 3; k++) out->center[k] = center[k], out->half[k] = half[k];
    out->aabb_mn[0] = center[0] - e.x, out->aabb_mn[1] = center[1] - e.y, out->aabb_mn[2] = center[2] - e.z;
    out->aabb_mx[0] = center[0] + e.x, out->aabb_mx[1] = center[1] + e.y, out->aabb_mx[2] = center[2] + e.z;
    out->surface = surface;
    out->max_force = 0;
    return BL_OK;
}

int bl_create(bl_ctx** out, const bl_settings* s) {
    if (!out || !s) return BL_ERR_ARG;
    *out = nullptr;
    if (s->sleeping) {
        g_last_error = "sleeping bodies are not stepped on the GPU (settings.sleeping must be 0)";
        return BL_ERR_UNSUPPORTED;
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) {
        g_last_error = "no HIP device";
        return BL_ERR_HIP;
    }
    bl_ctx* c = new bl_ctx();
    c->s = *s;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        g_last_error = "stream/event creation failed";
        delete c;
        return BL_ERR_HIP;
    }
    c->tune.mode = BL_MODE_AUTO;
    c->tune.block = 256;
    c->tune.substeps_per_launch = 0;
    c->tune.use_graph = 1;
    c->tune.incidences_per_thread = 4;
    *out = c;
    return BL_OK;
}

void bl_destroy(bl_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->release();
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->w0_ev) (void)hipEventDestroy(c->w0_ev);
    c->fin.release(), c->w0.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int bl_set_statics(bl_ctx* c, const bl_box* boxes, int32_t n) {
    if (!c || n < 0 || (n > 0 && !boxes)) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    c->boxes.assign(boxes, boxes + n);
    for (const bl_box& b : c->boxes)
        if (b.surface >= BL_SURF_COUNT) return BL_ERR_ARG;
    c->release();
    return BL_OK;
}

int bl_set_cylinders(bl_ctx* c, const bl_cylinder* cyl, int32_t n) {
    if (!c || n < 0 || (n > 0 && !cyl)) return BL_ERR_ARG;
    for (int32_t k = 0; k < n; k++)
        if (cyl[k].surface >= BL_SURF_COUNT) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    c->cyls.assign(cyl, cyl + n);
    c->release();
    return BL_OK;
}

int bl_set_terrain(bl_ctx* c, int32_t nx, int32_t nz, float cell, const float origin[2], const float* heights,
                   const uint8_t* surfaces) {
    if (!c || nx < 0 || nz < 0) return BL_ERR_ARG;
    if (nx > 0 && (nx < 2 || nz < 2 || !origin || !heights || !(cell > 0) || (size_t)nx * nz > (1u << 28))) return BL_ERR_ARG;
    const size_t n = (size_t)nx * nz;
    for (size_t q = 0; surfaces && q < n; q++)
        if (surfaces[q] >= BL_SURF_COUNT) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    c->tnx = nx > 0 ? nx : 0, c->tnz = nx > 0 ? nz : 0;
    c->th.clear(), c->ts.clear(), c->tblock.clear();
    if (c->tnx) {
        c->tcell = cell, c->tox = origin[0], c->toz = origin[1];
        c->th.assign(heights, heights + n);
        c->ts.assign(n, 0);
        if (surfaces) c->ts.assign(surfaces, surfaces + n);
        // Heightfield::update_bounds (src/phys/static_world.cpp:160-181): a vertex counts for up to 4 blocks
        c->tbx = (nx - 1 + 15) / 16, c->tbz = (nz - 1 + 15) / 16;
        c->tblock.assign((size_t)c->tbx * c->tbz, -1e30f);
        for (int z = 0; z < nz; z++)
            for (int x = 0; x < nx; x++) {
                const float v = c->th[(size_t)z * nx + x];
                const int bx0 = std::max(0, (x - 1) / 16), bx1 = std::min(c->tbx - 1, x / 16);
                const int bz0 = std::max(0, (z - 1) / 16), bz1 = std::min(c->tbz - 1, z / 16);
                for (int bz = bz0; bz <= bz1; bz++)
                    for (int bx = bx0; bx <= bx1; bx++) {
                        float& m = c->tblock[(size_t)bz * c->tbx + bx];
                        m = std::max(m, v);
                    }
            }
    }
    c->release();
    return BL_OK;
}

int bl_add_body(bl_ctx* c, const bl_body_desc* d, uint32_t* id) {
    if (!c || !d) return BL_ERR_ARG;
    if ((d->n_nodes && (!d->nodes || !d->info)) || (d->n_beams && !d->beams) || (d->n_tris && !d->tris) ||
        (d->n_capsules && !d->capsules))
        return BL_ERR_ARG;
    if (c->bodies.size() >= 65535) return BL_ERR_ARG;
    if (unsupported_opts(d->opts)) {
        g_last_error = "air drag, aerodynamic drag, rest damping and the energy guard are not stepped on the GPU";
        return BL_ERR_UNSUPPORTED;
    }
    for (uint32_t j = 0; j < d->n_beams; j++)
        if (d->beams[j].a >= d->n_nodes || d->beams[j].b >= d->n_nodes || d->beams[j].type > BL_BT_SUPPORT) return BL_ERR_ARG;
    if ((d->n_wheels && !d->wheels) || (d->n_wheel_nodes && !d->wheel_nodes)) return BL_ERR_ARG;
    for (uint32_t q = 0; q < d->n_wheel_nodes; q++)
        if (d->wheel_nodes[q] >= d->n_nodes) return BL_ERR_ARG;
    for (uint32_t q = 0; q < d->n_wheels; q++) {
        const bl_wheel& w = d->wheels[q];
        if (w.ring) {
            g_last_error = "ring tyres are not stepped on the GPU";
            return BL_ERR_UNSUPPORTED;
        }
        if (w.axle0 >= d->n_nodes || w.axle1 >= d->n_nodes || (uint64_t)w.node0 + w.n_nodes > d->n_wheel_nodes ||
            w.arm >= (int32_t)d->n_nodes || w.near_attach >= (int32_t)d->n_nodes || d->n_wheels > 32767)
            return BL_ERR_ARG;
    }
    for (uint32_t t = 0; t < d->n_tris; t++) {
        const bl_triangle& tr = d->tris[t];
        if (tr.a >= d->n_nodes || tr.b >= d->n_nodes || tr.c >= d->n_nodes) return BL_ERR_ARG;
    }
    if ((d->n_shocks && !d->shocks) || (d->n_slides && !d->slides) || (d->n_rail && !d->rail) ||
        (d->n_frames && !d->frames) || (d->n_joints && !d->joints))
        return BL_ERR_ARG;
    for (uint32_t q = 0; q < d->n_frames; q++) {
        if (d->frames[q].node >= d->n_nodes) return BL_ERR_ARG;
        // (a triad's nodes out of range make a dynamic frame, as in SoftBody::integrate_frames)
    }
    for (uint32_t q = 0; q < d->n_joints; q++) {
        const bl_joint& j = d->joints[q];
        if (j.parent_frame >= d->n_frames || j.child_node >= d->n_nodes || j.child_frame >= (int32_t)d->n_frames ||
            j.child_frame < -1)
            return BL_ERR_ARG;
    }
    for (uint32_t q = 0; q < d->n_shocks; q++)
        if (d->shocks[q].beam >= d->n_beams || d->shocks[q].type < 1 || d->shocks[q].type > 3) return BL_ERR_ARG;
    for (uint32_t r = 0; r < d->n_rail; r++)
        if (d->rail[r] >= d->n_nodes) return BL_ERR_ARG;
    for (uint32_t q = 0; q < d->n_slides; q++) {
        const bl_slide& s = d->slides[q];
        if (s.node >= d->n_nodes || (uint64_t)s.rail0 + s.nrail > d->n_rail || (s.nrail >= 2 && s.seg >= (int32_t)s.nrail - 1))
            return BL_ERR_ARG;
    }
    for (uint32_t q = 0; q < d->n_capsules; q++) {
        const bl_capsule& cp = d->capsules[q];
        if (cp.a >= d->n_nodes || cp.b >= d->n_nodes) return BL_ERR_ARG;
        if (cp.joint >= 0) {
            g_last_error = "capsules tied to joints are not stepped on the GPU";
            return BL_ERR_UNSUPPORTED;
        }
    }
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    HostBody hb;
    hb.nodes.assign(d->nodes, d->nodes + d->n_nodes);
    hb.info.assign(d->info, d->info + d->n_nodes);
    hb.beams.assign(d->beams, d->beams + d->n_beams);
    if (d->n_tris) hb.tris.assign(d->tris, d->tris + d->n_tris);
    if (d->n_capsules) hb.caps.assign(d->capsules, d->capsules + d->n_capsules);
    if (d->n_shocks) hb.shocks.assign(d->shocks, d->shocks + d->n_shocks);
    if (d->n_slides) hb.slides.assign(d->slides, d->slides + d->n_slides);
    if (d->n_rail) hb.rail.assign(d->rail, d->rail + d->n_rail);
    if (d->n_frames) hb.frames.assign(d->frames, d->frames + d->n_frames);
    if (d->n_joints) hb.joints.assign(d->joints, d->joints + d->n_joints);
    if (d->n_wheels) hb.wheels.assign(d->wheels, d->wheels + d->n_wheels);
    if (d->n_wheel_nodes) hb.wnodes.assign(d->wheel_nodes, d->wheel_nodes + d->n_wheel_nodes);
    hb.ctl.assign(d->n_wheels, make_float4(0, 0, 0, 0));
    for (uint32_t i = 0; i < d->n_nodes; i++) hb.tyres |= (d->info[i].flags & BL_NF_TYRE) != 0;
    hb.opts = d->opts;
    hb.compute_bounds(); // (SoftBody::compute_aabb at World::add_body, src/phys/world.cpp:283-289)
    c->bodies.push_back(std::move(hb));
    c->release();
    if (id) *id = (uint32_t)c->bodies.size() - 1;
    return BL_OK;
}

int bl_step_async(bl_ctx* c, int32_t substeps) {
    if (!c || substeps < 0) return BL_ERR_ARG;
    int rc = ensure_built(c);
    if (rc) return rc;
    if ((rc = post_frame(c))) return rc; // (the previous frame's check first)
    return enqueue_frame(c, substeps);
}

int bl_sync(bl_ctx* c) {
    if (!c) return BL_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    return post_frame(c);
}

int bl_step(bl_ctx* c, int32_t substeps) {
    int rc = bl_step_async(c, substeps);
    if (rc) return rc;
    return bl_sync(c);
}

int bl_read_body(bl_ctx* c, uint32_t id, bl_node* nodes_out, bl_beam* beams_out) {
    if (!c || id >= c->bodies.size()) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    const HostBody& hb = c->bodies[id];
    if (nodes_out && !hb.nodes.empty()) std::memcpy(nodes_out, hb.nodes.data(), sizeof(bl_node) * hb.nodes.size());
    if (beams_out && !hb.beams.empty()) std::memcpy(beams_out, hb.beams.data(), sizeof(bl_beam) * hb.beams.size());
    return BL_OK;
}

int bl_write_body(bl_ctx* c, uint32_t id, const bl_node* nodes, const bl_beam* beams) {
    if (!c || id >= c->bodies.size()) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    HostBody& hb = c->bodies[id];
    if (beams)
        for (size_t j = 0; j < hb.beams.size(); j++)
            if (beams[j].a != hb.beams[j].a || beams[j].b != hb.beams[j].b || beams[j].type > BL_BT_SUPPORT) return BL_ERR_ARG;
    if (nodes && !hb.nodes.empty()) std::memcpy(hb.nodes.data(), nodes, sizeof(bl_node) * hb.nodes.size());
    if (beams && !hb.beams.empty()) std::memcpy(hb.beams.data(), beams, sizeof(bl_beam) * hb.beams.size());
    // counters survive the rebuild
    std::vector<unsigned long long> cnt(CNT_N, 0), brk(c->bodies.size(), 0);
    if (c->built) {
        HIPCHK(hipMemcpy(cnt.data(), c->counters.p, sizeof(unsigned long long) * CNT_N, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(brk.data(), c->broken.p, sizeof(unsigned long long) * brk.size(), hipMemcpyDeviceToHost));
    }
    c->release();
    if ((rc = build(c))) return rc;
    HIPCHK(hipMemcpy(c->counters.p, cnt.data(), sizeof(unsigned long long) * CNT_N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->broken.p, brk.data(), sizeof(unsigned long long) * brk.size(), hipMemcpyHostToDevice));
    return BL_OK;
}

int bl_read_slides(bl_ctx* c, uint32_t id, bl_slide* out) {
    if (!c || id >= c->bodies.size()) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    const HostBody& hb = c->bodies[id];
    if (out && !hb.slides.empty()) std::memcpy(out, hb.slides.data(), sizeof(bl_slide) * hb.slides.size());
    return BL_OK;
}

int bl_read_wheels(bl_ctx* c, uint32_t id, bl_wheel* wheels_out, float speed_spin[2]) {
    if (!c || id >= c->bodies.size()) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    const HostBody& hb = c->bodies[id];
    if (wheels_out && !hb.wheels.empty()) std::memcpy(wheels_out, hb.wheels.data(), sizeof(bl_wheel) * hb.wheels.size());
    if (speed_spin) speed_spin[0] = hb.wsp[0], speed_spin[1] = hb.wsp[1];
    return BL_OK;
}

int bl_set_wheel_controls(bl_ctx* c, uint32_t id, const float* torque, const float* brake) {
    if (!c || id >= c->bodies.size()) return BL_ERR_ARG;
    HostBody& hb = c->bodies[id];
    const bool on = torque && brake;
    for (size_t k = 0; k < hb.wheels.size(); k++) hb.ctl[k] = on ? make_float4(torque[k], brake[k], 1.0f, 0) : make_float4(0, 0, 0, 0);
    if (c->built && !hb.wheels.empty() && c->xctl.n >= hb.wheel0 + hb.wheels.size()) {
        // (the step in flight reads the controls: it ends first)
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->xctl.p + hb.wheel0, hb.ctl.data(), sizeof(float4) * hb.ctl.size(), hipMemcpyHostToDevice));
    }
    return BL_OK;
}

int bl_read_joints(bl_ctx* c, uint32_t id, bl_frame* frames_out, bl_joint* joints_out) {
    if (!c || id >= c->bodies.size()) return BL_ERR_ARG;
    int rc = post_frame(c);
    if (rc) return rc;
    rc = pull_state(c);
    if (rc) return rc;
    const HostBody& hb = c->bodies[id];
    if (frames_out && !hb.frames.empty()) std::memcpy(frames_out, hb.frames.data(), sizeof(bl_frame) * hb.frames.size());
    if (joints_out && !hb.joints.empty()) std::memcpy(joints_out, hb.joints.data(), sizeof(bl_joint) * hb.joints.size());
    return BL_OK;
}

int bl_body_broken_beams(bl_ctx* c, uint32_t id, int64_t* out) {
    if (!c || !out || id >= c->bodies.size()) return BL_ERR_ARG;
    *out = 0;
    if (!c->built) return BL_OK;
    unsigned long long v = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&v, c->broken.p + id, sizeof(v), hipMemcpyDeviceToHost));
    *out = (int64_t)v;
    return BL_OK;
}

int bl_get_stats(bl_ctx* c, bl_stats* out) {
    if (!c || !out) return BL_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->substeps = c->substeps;
    if (!c->built) return BL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long cnt[CNT_N];
    HIPCHK(hipMemcpy(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> brk(c->broken.n);
    HIPCHK(hipMemcpy(brk.data(), c->broken.p, sizeof(unsigned long long) * brk.size(), hipMemcpyDeviceToHost));
    out->static_contacts = (int64_t)cnt[CNT_STATIC];
    out->vel_clamps = (int64_t)cnt[CNT_CLAMPS];
    out->nan_resets = (int64_t)cnt[CNT_NANS];
    out->contacts = (int64_t)cnt[CNT_CONTACTS];
    out->pair_rebuilds = c->contacts.pair_rebuilds;
    out->contact_pairs = c->contacts.contact_pairs;
    out->narrow_tests = c->contacts.narrow_tests;
    for (unsigned long long b : brk) out->broken_beams += (int64_t)b;
    unsigned int init[8] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0, 0, 0};
    HIPCHK(hipMemcpy(c->bounds.p, init, sizeof(init), hipMemcpyHostToDevice));
    int rc = launch_bounds(c->args, c->stream, c->parity, c->bounds.p);
    if (rc) return rc;
    unsigned int res[8];
    HIPCHK(hipMemcpy(res, c->bounds.p, sizeof(res), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; k++) {
        const uint32_t lo = res[k], hi = res[3 + k];
        out->aabb_mn[k] = u2f((lo & 0x80000000u) ? (lo & 0x7fffffffu) : ~lo);
        out->aabb_mx[k] = u2f((hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi);
    }
    out->max_speed = std::sqrt(u2f(res[6]));
    return BL_OK;
}

int bl_reset_stats(bl_ctx* c) {
    if (!c) return BL_ERR_ARG;
    c->substeps = 0;
    c->contacts.pair_rebuilds = c->contacts.contact_pairs = c->contacts.narrow_tests = 0;
    if (!c->built) return BL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemset(c->counters.p, 0, sizeof(unsigned long long) * CNT_N));
    HIPCHK(hipMemset(c->broken.p, 0, sizeof(unsigned long long) * c->broken.n));
    return BL_OK;
}

int bl_set_tuning(bl_ctx* c, const bl_tuning* t) {
    if (!c || !t) return BL_ERR_ARG;
    if (t->substeps_per_launch < 0 || !tuning_valid(*t)) {
        g_last_error = "no such execution form (mode, block, incidences per thread)";
        return BL_ERR_ARG;
    }
    bl_tuning nt = *t;
    if (!nt.block) nt.block = 256;
    if (!nt.incidences_per_thread) nt.incidences_per_thread = 4;
    const bool rechunk = nt.block != effective_block(c) || nt.incidences_per_thread != effective_per(c);
    int rc = post_frame(c);
    if (rc) return rc;
    if (rechunk && (rc = pull_state(c))) return rc;
    c->tune = nt;
    c->drop_graphs();
    if (rechunk && c->built) {
        c->release();
        return build(c);
    }
    return BL_OK;
}

int bl_get_tuning(bl_ctx* c, bl_tuning* t) {
    if (!c || !t) return BL_ERR_ARG;
    *t = c->tune;
    if (c->built) t->mode = resolved_mode(c);
    return BL_OK;
}

int bl_time_step(bl_ctx* c, int32_t substeps, float* ms) {
    if (!c || !ms || substeps < 0) return BL_ERR_ARG;
    int rc = ensure_built(c);
    if (rc) return rc;
    if ((rc = post_frame(c))) return rc;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if ((rc = enqueue_frame(c, substeps))) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
    return post_frame(c);
}

// Times one frame of `substeps` with an event pair around every kernel launch of the substeps (no graph), so that
// *kernel_ms is the summed duration of the substep kernels alone and *total_ms the frame's wall time on the stream.
int bl_time_kernels(bl_ctx* c, int32_t substeps, float* total_ms, float* kernel_ms, int32_t* launches) {
    if (!c || !total_ms || !kernel_ms || !launches || substeps <= 0) return BL_ERR_ARG;
    int rc = ensure_built(c);
    if (rc) return rc;
    if ((rc = post_frame(c))) return rc;
    bl_tuning t = c->tune;
    t.block = effective_block(c);
    t.mode = resolved_mode(c);
    t.incidences_per_thread = effective_per(c);
    if (t.block != c->built_block || t.incidences_per_thread != c->built_per) return BL_ERR_STATE;
    if (needs_contacts(c, substeps)) {
        // a contact frame: the beam kernel (the force phase) of every substep between events
        while ((int)c->kev.size() < 2 * substeps) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c->kev.push_back(e);
        }
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        if ((rc = contact_frame(c, substeps, c->kev.data()))) return rc;
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipEventSynchronize(c->ev1));
        HIPCHK(hipEventElapsedTime(total_ms, c->ev0, c->ev1));
        float sum = 0;
        for (int l = 0; l < substeps; l++) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, c->kev[2 * l], c->kev[2 * l + 1]));
            sum += ms;
        }
        *kernel_ms = sum;
        *launches = substeps;
        return post_frame(c);
    }
    if (c->xargs.nbody) {
        g_last_error = "the forms' kernel timing covers worlds without shocks or slide nodes (those step in the split form)";
        return BL_ERR_UNSUPPORTED;
    }
    const int per = t.mode == BL_MODE_PERSISTENT ? (t.substeps_per_launch > 0 ? t.substeps_per_launch : 64)
                    : (t.mode == BL_MODE_RESIDENT || t.mode == BL_MODE_HALO)
                        ? (t.substeps_per_launch > 0 ? t.substeps_per_launch : substeps) : 1;
    const int nl = (substeps + per - 1) / per;
    while ((int)c->kev.size() < 2 * nl) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->kev.push_back(e);
    }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if ((rc = launch_frame_begin(c->args, c->stream, c->parity))) return rc;
    for (int l = 0; l < nl; l++) {
        const int k = std::min(per, substeps - l * per);
        HIPCHK(hipEventRecord(c->kev[2 * l], c->stream));
        if (t.mode == BL_MODE_RESIDENT) rc = launch_resident(c->args, c->stream, c->parity, k, t, c->bar.p, c->err.p, kSpinLimit);
        else if (t.mode == BL_MODE_HALO) rc = launch_halo(c->args, c->stream, c->parity, k, t, c->flags.p, c->err.p, kSpinLimit);
        else rc = launch_substeps(c->args, c->stream, c->parity, k, t);
        if (rc) return rc;
        HIPCHK(hipEventRecord(c->kev[2 * l + 1], c->stream));
        c->parity ^= (k & 1);
    }
    c->substeps += substeps;
    if ((rc = launch_frame_end(c->args, c->stream, c->parity))) return rc;
    if ((rc = frame_enqueued(c))) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(total_ms, c->ev0, c->ev1));
    float sum = 0;
    for (int l = 0; l < nl; l++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->kev[2 * l], c->kev[2 * l + 1]));
        sum += ms;
    }
    *kernel_ms = sum;
    *launches = nl;
    return post_frame(c);
}

} // extern "C"
