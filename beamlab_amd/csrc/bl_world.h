// Device-side world layout shared by the kernels and the host code of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/beamlab_gpu.h"
#include "bl_devmath.h"

namespace blg {

// Per-incidence flags (iM[e].w): the beam's flags that the force loop reads, folded with its body's settings.
enum : uint32_t {
    TF_BROKEN = 1u << 0,   // BF_BROKEN
    TF_SKIP = 1u << 1,     // BF_SHOCK: left to the shock loop, skipped by the beam loop
    TF_NODEFORM = 1u << 2, // BF_NO_DEFORM, or the body's allow_deform off
    TF_NOBREAK = 1u << 3,  // BF_NO_BREAK
    TF_BRK = 1u << 4,      // the body's allow_break
    TF_DUCTILE = 1u << 5,  // the body's ductile
    TF_ISB = 1u << 8,      // this incidence is the beam's b end
};
constexpr int TF_TYPE_SHIFT = 6;
// bits 16 and up: the incidence's own node less its chunk's first node (the other end is in iOth)
constexpr int TF_OWN_SHIFT = 16;

enum { CNT_STATIC = 0, CNT_CLAMPS = 1, CNT_NANS = 2, CNT_CONTACTS = 3, CNT_N = 4 };
// per-body frame words: [0] top speed squared (float bits), [1..3] min key x y z, [4..6] max key x y z
enum { BF_V2 = 0, BF_MN = 1, BF_MX = 4, BF_N = 8 };

// order-preserving map of a float onto unsigned integers (for atomic min / max of floats)
BL_HD uint32_t fkey(float f) {
    const uint32_t u = f2u(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
BL_HD float fkey_inv(uint32_t k) { return u2f((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

struct BodyParam {
    float bounce, push_max, slop;
    float collision_radius, contact_friction; // inter-body contacts
    float hull_depth, face_skin;              // its hull (one-sided) triangles: depth behind, faces-only skin
    float hull_reach;                         // hull_depth when it has a hull triangle, else 0
    int32_t group, self_coll, faces_only, pad;
    int32_t deform, brk, pad2, pad3; // its allow_deform, allow_break (the joints read them)
};

struct KArgs {
    float4* pm[2];
    float4* vm[2];
    const uint32_t* ninfo; // node flags (low 16 bits) | body index (high 16 bits)
    const float2* naux;    // radius, friction x ground friction
    const uint32_t* noff;  // CSR offsets of the incidences, [N + 1]
    float4* iA;            // L, k, d, strength
    float4* iM;            // max_pos, max_neg, beam index (bits), flags (bits)
    const uint32_t* iOth;  // the beam's other end (its a end when this is its b end, else its b end)
    const float4* bcold;   // per beam: plastic, L0, support_limit, 0
    float* stress;         // per beam
    const uint32_t* ch_n0; // chunk node starts, [chunks + 1]
    const uint4* ch_rng;   // per chunk: first node, end node, first incidence, end incidence (one load)
    const uint32_t* body_c0; // first chunk of each body, [bodies + 1]
    const BodyParam* bodies;
    const bl_box* boxes;
    unsigned long long* counters; // CNT_*
    unsigned long long* broken;   // per body
    unsigned int* body_frame;     // per body, BF_N words: top |v|^2 of the frame (float bits), bounds keys
    uint4* cstat;                 // per chunk: static contacts, clamps, NaN resets, top |v|^2 (float bits) of the frame
    const uint32_t* chunk_body;   // per chunk: its body
    // the halo form: per chunk, the nodes outside it that its incidences read (halo[halo_off[c] ..]), the chunks
    // owning them (nbr[nbr_off[c] ..]), and per incidence the LDS slots of its two nodes (a | b << 16; the chunk's
    // own nodes first, then its halo)
    const uint32_t* halo_off;
    const uint32_t* halo;
    const uint32_t* nbr_off;
    const uint32_t* nbr;
    const uint32_t* iloc;
    // inter-body contacts: the node forces between the force phase and the integration, the collision triangles
    // and capsules with global node indices
    float4* F;
    const uint4* tri_abc;  // a, b, c, body
    const uint8_t* tri_hull; // 1: a hull (one-sided) triangle
    const float* tri_rest2;
    uint32_t* tri_torn;
    const uint4* cap_data; // a, b, body, radius (bits)
    uint32_t ntris, ncaps;
    f3 g;
    float dt;
    float kill_y;
    int nbox;
    uint32_t nnodes;
    uint32_t nchunks;
    uint32_t nbodies;
    const uint32_t* wg_chunk; // workgroup -> chunk (XCD placement), or null for the identity
    // the force phase's own chunks (k_forces): the chunks with beams, then the nodes of the beamless chunks packed
    // into runs of at most a workgroup (bodies without beams - balls - would each take a workgroup of their own)
    const uint4* fch_rng;     // (first node, end node, first incidence, end incidence)
    const uint32_t* fwg;      // workgroup -> force chunk (contiguous ranges over the XCDs)
    uint32_t nfchunks;
    // the ranges in workgroup order: wg_rng[b] = ch_rng[wg_chunk[b]], fwg_rng[b] = fch_rng[fwg[b]]
    const uint4* wg_rng;
    const uint4* fwg_rng;
    // the static world beyond the boxes: vertical cylinders; the terrain (dense heights and surfaces, z-major) and
    // each body's terrain top over its reach this frame (null without a terrain)
    const bl_cylinder* cyls;
    int ncyl;
    const float* th;
    const uint8_t* ts;
    int tnx, tnz;
    float tcell, tox, toz;
    const float* tmax;
};

// The bodies' shocks and slide nodes (the forces that follow the beams in World::simulate_island's base_forces,
// src/phys/world.cpp:3255-3273): run after the force phase, into F, a thread per body with any, in the body's order.
struct XArgs {
    const uint4* body;    // per such body, three records: first shock, end shock, first slide, end slide; first joint,
                          // end joint, first frame, end frame; first wheel, end wheel, first tyre node, end tyre node
    uint32_t nbody;
    const bl_shock* sh;   // the shocks (beam: the body's beam index, unused here)
    const uint4* shx;     // per shock: its beam's a-end and b-end incidences, its a and b nodes
    bl_slide* sl;         // the slides (node, rail0 global); seg and broken updated here
    const uint32_t* rail; // the rails' nodes (global)
    bl_joint* jt;         // the joints (frames and child node global); their plastic state, stress, broken updated
    float4* fq;           // per frame: orientation (x, y, z, w)
    float4* fw;           // per frame: angular velocity, inverse inertia
    float4* ftq;          // per frame: torque of the substep
    const uint32_t* fnode;
    const int2* fref;     // per frame: a kinematic frame's two reference nodes (global), or -1
    uint32_t nframes;
    // wheels: the records (axles, arm and near_attach global; node0 into wn), their tyre nodes, a controller's
    // (torque, brake, on) per wheel, and per such body its wheel_speed, wheel_spin
    bl_wheel* wh;
    const uint32_t* wn;
    const float4* ctl;
    float2* wsp;
    // per body, its tyre nodes in node order (node, wheel or ~0), and the tyre patches' scratch: per wheel the need
    // and cap sums, per tyre node its need (cap in w) and slide (touching in w)
    const uint2* ty;
    float4* tsum;
    float4* tneed;
    float4* tslide;
    float tyre_grip;
};
int launch_extras(const KArgs& a, const XArgs& x, hipStream_t st, int parity);
// the tyre nodes' static contacts and their wheels' shared grip (World::collide_static's tyre part), into F before the
// integration, whose static contact then leaves those nodes' forces as they are
int launch_tyres(const KArgs& a, const XArgs& x, hipStream_t st, int parity);
// the frames' integration after the nodes' (SoftBody::integrate_frames)
int launch_frames(const KArgs& a, const XArgs& x, hipStream_t st, int parity_new); // parity_new: the integrated state

int launch_substeps(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t);
int launch_resident(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t, unsigned int* bar,
                    int* err, uint32_t spin_limit);
int resident_fits(int block, int per, uint32_t nchunks);
int launch_halo(const KArgs& a, hipStream_t st, int parity, int substeps, const bl_tuning& t, unsigned int* flags,
                int* err, uint32_t spin_limit);
int halo_fits(int block, int per, uint32_t nchunks);
int halo_capacity(int block, int per);
// the name of the last failed cooperative launch's status (resident and halo forms)
const char* coop_error_name();
int launch_bounds(const KArgs& a, hipStream_t st, int parity, unsigned int* out);
int launch_frame_begin(const KArgs& a, hipStream_t st, int parity);
int launch_frame_end(const KArgs& a, hipStream_t st, int parity);
int launch_chunk_clear(const KArgs& a, hipStream_t st);
// each body's bounds of the current state into a.body_frame's bound words (the caller points it at its own buffer)
int launch_box(const KArgs& a, hipStream_t st, int parity);
// the substep split for inter-body contacts: forces into F, then the static contacts and the integration from F
int launch_forces(const KArgs& a, hipStream_t st, int parity, const bl_tuning& t);
int launch_integrate(const KArgs& a, hipStream_t st, int parity, uint4* wstat); // wstat: ceil(nodes / 256) slots
int launch_wstat_fold(const KArgs& a, hipStream_t st, uint4* wstat);
int chunk_capacity(int block, int per);
// the message bl_error_string returns for the calling thread's next error status
void set_last_error(const char* msg);

} // namespace blg
