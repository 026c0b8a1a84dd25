// Inter-body and self contacts of slow bodies: the device records of the pair search and the narrow phase, and the
// host engine that drives them once per pair rebuild and once per substep.
#pragma once

#include "bl_world.h"

#include <vector>

namespace blg {

// A narrow-phase hit (src/phys/world.cpp:2807-2962): the candidate's index in the concatenated lists, the capsule
// parameter, the triangle's barycentric weights, the normal (from the other side towards the first) and the depth.
struct Hit {
    uint32_t pair;
    float s;
    f3 bary, nrm;
    float pen;
};

// The device tables of one island's pair rebuild.
struct CArgs {
    int parity;
    float margin, inv_cell, rmax, horizon;
    const int* island_local;     // per body: its index in the island, -1 outside
    const uint32_t* query;       // per island body: has a partner or collides with itself
    const uint32_t* self;        // per island body: self collision
    const uint32_t* island_body; // per island body: its body
    const float* boxes;          // per island body: the partner-search box (mn xyz, mx xyz)
    const uint32_t* part_off;    // per island body: its partners (island indices), CSR
    const uint32_t* part;
    const uint32_t* body_tri0;   // per body: its first triangle
    const uint32_t* sample_base; // per body: its first size sample
    float* samples;              // every 7th triangle's longest edge
    uint32_t* overflow;          // set when a capsule meets more triangles than it keeps
    uint32_t *flag, *scan, *scan2, *count;
    uint32_t npts, nprims, mask;
    uint64_t* keys;
    int4* pt_cell;               // hashed nodes: cell x y z, node
    float4* pt_pos;              // hashed nodes: position, island body (bits)
    uint32_t *bstart, *items, *run_end;
    uint32_t *entry_pid, *prim_tri;
    uint4 *nt, *nc, *ct;
};

// The candidate lists of a frame (all pair islands) and the narrow phase's buffers.
struct PArgs {
    const uint4* nt; // node body, triangle body, node, triangle
    const uint4* nc; // node body, capsule body, node, capsule
    const uint4* ct; // capsule body, triangle body, capsule, triangle
    uint32_t nnt, nnc, nct;
    uint32_t *hflag, *hscan;
    Hit *hraw, *hits;
    uint32_t* nhits;
};

// What the island and pair building read of a body on the host.
struct BodyView {
    int32_t group;
    uint32_t self, contacters, ntri, ncap, tri0;
    float collision_radius;
};

struct HostView {
    std::vector<BodyView> bodies;
    std::vector<float> max_speed; // per body, as of the frame's start
    std::vector<f3> mn, mx;       // per body bounds, as of the frame's start
    float dt = 0;
    int inter_body = 0;
};

template <class T>
struct Buf {
    T* p = nullptr;
    size_t cap = 0;
    int ensure(size_t n) {
        if (p && n <= cap) return BL_OK;
        if (p && hipDeviceSynchronize() != hipSuccess) return BL_ERR_HIP; // (work queued on the old buffer ends first)
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = (n < 1024 ? 1024 : n) * 3 / 2;
        return hipMalloc((void**)&p, sizeof(T) * cap) == hipSuccess ? BL_OK : BL_ERR_HIP;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

class ContactEngine {
public:
    // per frame
    std::vector<std::vector<uint32_t>> islands, pair_islands;
    int interval = 1;
    bool unsupported = false;
    // counters (bl_stats)
    int64_t pair_rebuilds = 0, contact_pairs = 0, narrow_tests = 0;

    int setup(const HostView& hv);
    void build_islands(const HostView& hv, float frame_time);
    // the pair lists anew; body_v2 / body_box: each body's top speed squared so far and its current bounds
    int rebuild(const KArgs& a, const HostView& hv, const std::vector<float>& body_v2, const std::vector<float>& body_box,
                hipStream_t st, int parity);
    int collide(const KArgs& a, hipStream_t st, int parity);
    void release();
    ~ContactEngine() { release(); }

private:
    size_t nnt = 0, nnc = 0, nct = 0;
    uint32_t nsamples = 0;
    Buf<uint8_t> temp;
    Buf<uint32_t> flag, scanv, scan2, count, bstart, items, run_end, entry_pid, prim_tri, hflag, hscan, nhits, overflow;
    Buf<uint32_t> isl_query, isl_self, isl_body, part_off, part, sample_base, body_tri0;
    Buf<int> isl_local;
    Buf<float> boxes, samples;
    Buf<uint64_t> keys, sorted;
    Buf<int4> pt_cell;
    Buf<float4> pt_pos;
    Buf<uint4> nt, nt2, nc, ct, all_nt, all_nc, all_ct;
    Buf<Hit> hraw, hits;

    int scan(const uint32_t* in, uint32_t* out, size_t n, hipStream_t st);
    int sort(const uint64_t* in, uint64_t* out, size_t n, int end_bit, hipStream_t st);
    int total(const uint32_t* scan_out, const uint32_t* in, size_t n, uint32_t& tot, hipStream_t st);
    int rebuild_island(const KArgs& a, const HostView& hv, const std::vector<uint32_t>& isl,
                       const std::vector<float>& body_v2, const std::vector<float>& body_box,
                       const std::vector<float>& samples, hipStream_t st, int parity, int64_t& island_pairs);
};

} // namespace blg
