// Inter-body and self contacts of slow bodies: the device records of the pair search and the narrow phase, and the
// host engine that drives them once per pair rebuild and once per substep.
#pragma once

#include "bl_islands.h"
#include "bl_world.h"

#include <vector>

namespace blg {

// A narrow-phase hit (src/phys/world.cpp:2807-2962): the candidate's index in the concatenated lists, the capsule
// parameter, the triangle's barycentric weights, the normal (from the other side towards the first) and the depth;
// and what the response reads of its candidate, written by the narrow phase, which has it at hand: the contact's
// nodes (side a's, then side b's), its friction and its shape (nodes on side a | on side b << 4 | hull push-out << 8).
struct Hit {
    uint32_t pair;
    float s;
    f3 bary, nrm;
    float pen;
    uint32_t n[5];
    float fric;
    uint32_t shape;
};

// A pair island's search parameters of the current rebuild (rebuild_pairs, src/phys/world.cpp:1879-1935).
struct IslandPar {
    float margin, inv_cell, rmax, reach;
    float speed; // the slow bodies' top speed + 0.5 m/s (the margins' speed)
};

// A fast body's own grid (World::fast_pairs, src/phys/world.cpp:2330-2344): origin, inverse cell size, cells an axis.
struct FastGrid {
    float ox, oy, oz, inv;
    int dx, dy, dz, pad;
};

// What a rebuild leaves for the host: list sizes (needed sizes when a list overflowed its buffer) and flags.
struct RStat {
    uint32_t npts, nnt_raw, nnc, nct, nent, nnt;
    uint32_t overflow;    // bit 0 NT, 1 NC, 2 CT, 3 entries, 4 partners, 5 fast queries, 6 fast NT, 7 fast NC
    uint32_t unsupported; // bit 1 a capsule meeting too many triangles, 2 a scan's wait timed out, 3 lists out of order
    uint32_t last_pairs;  // candidate pairs kept in the last pair island
    uint32_t npart;       // partner entries (when they overflowed their buffer)
    uint32_t nisl;        // response groups of the rebuild (the response's workgroups read it)
    uint32_t nfq;         // the fast bodies' queries (their partners' triangles and capsules, summed)
    uint32_t nfnt, nfnc;  // the fast bodies' node-triangle and node-capsule candidates
    uint32_t nnc_slow;    // the node-capsule candidates of the node hash (nnc counts the fast ones too)
};

// A one-pass launch's look-back: the workgroups' published words, the ticket counter (null: workgroups in index
// order), the tickets taken before the launch and its epoch.
struct Lookback {
    uint64_t* words;
    uint32_t* ticket;
    uint32_t base, epoch;
};

// The device tables of a rebuild (all pair islands at once: one hash table, its keys salted by island).
struct CArgs {
    int parity;
    float horizon;
    const int* bisl;          // per body: its pair island, -1 when none
    const uint32_t* imem_off; // per pair island: its bodies (CSR, ascending)
    const uint32_t* imem;
    uint32_t nisl, nmem;
    const int* bmem;          // per body: its index in imem, -1 outside the pair islands
    uint32_t *pcnt, *poff, *plist; // per island body: its partners (bodies), CSR
    uint32_t cap_part;
    int interval;
    float dt;
    IslandPar* ipar;
    float* bbox;              // per body: the partner box (bounds +- the island's reach)
    float* extra;             // per body: a fast body's extra margin (its bulk speed x the interval), else 0
    uint32_t* bquery;         // per body: has a partner or collides with itself
    const uint32_t* body_box; // per body: bounds keys (BF_MN, BF_MX words of a frame record)
    const uint32_t* body_tri0;
    const uint32_t* body_cap0;
    const uint32_t* itri;     // the triangles of the pair islands' bodies, island by island, each in body order
    uint32_t n_itri;
    const uint32_t* icap;     // the capsules of the pair islands' bodies, in the same order
    uint32_t n_icap;
    const int* igroup;        // per pair island: its response group (consecutive islands, balanced by size)
    uint32_t ngroup;
    uint32_t* irange;         // per response group: its runs [start, end) in the NT, NC and CT lists (6 words)
    const uint32_t* sample_base;
    float* samples;           // every 7th triangle's longest edge, in body order
    uint32_t nsamples;
    RStat* st;
    uint32_t* icount;         // per pair island: candidate pairs kept (NT after the filter, NC, CT)
    // hashed nodes
    uint32_t* pbucket;
    int4* pcell;   // cell x y z, node
    float4* ppos;  // position, body (bits)
    uint32_t tmask;
    uint32_t *tcount, *tstart;
    int4* scell;   // the hashed nodes' records by bucket
    float4* spos;
    // node-triangle, node-capsule candidates
    uint4 *ntraw, *nc;
    uint32_t cap_nt, cap_nc;
    // capsule-triangle: triangle cells, their table, the candidates
    uint32_t* eb;
    uint4* ea;     // per entry: triangle, body, group, island
    float4* ebb;   // per entry: raw bounds (min x y z, max x)
    float2* ec;    // per entry: raw bounds (max y z)
    uint4* sa;     // the entries in bucket order
    float4* sb;
    float2* sc;
    uint32_t cap_ent, emask;
    uint32_t *ecount, *estart;
    uint4* ct;
    uint32_t cap_ct;
    // the fast bodies' pairs: per island member its grid and its first query (its partners' triangles, then their
    // capsules, in partner order), the fast candidates, and the lists merged island by island (the slow candidates of
    // an island, then its fast ones) when there are fast candidates
    FastGrid* fgrid;
    uint32_t* fqoff; // [nmem + 1]
    uint4 *fnt, *fnc, *ntm, *ncm;
    uint32_t cap_fq, cap_fnt, cap_fnc;
    // the near-pair list
    uint4* nt;
    Lookback lk; // the current one-pass launch's
};

// The candidate lists of a frame (all pair islands) and the narrow phase's buffers.
struct PArgs {
    const uint4* nt; // node body, triangle body, node, triangle
    const uint4* nc; // node body, capsule body, node, capsule
    const uint4* ct; // capsule body, triangle body, capsule, triangle
    const uint4* ncm; // the node-capsule list merged with the fast bodies' (the list when st->nfnc is not 0)
    uint32_t nnt, nnc, nct; // (the kernels take them from st, the last rebuild's counts on the device)
    const RStat* st;
    Hit* hraw;       // per block of 256 candidates: its hits first, in order
    uint32_t* bcnt;  // per block: its hit count
    const uint32_t* irange; // per pair island: its runs [start, end) in the NT, NC and CT lists
    uint32_t gresp;  // the response's workgroups (at least the pair islands: st->nisl)
    int* err;        // the context's error word (bit 1: a response's dependency wait gave up)
};

// What the island building reads of a body on the host.
struct BodyView {
    int32_t group;
    uint32_t self, contacters, ntri, ncap, tri0, nnode;
    float collision_radius;
    float hull_reach; // SoftBody::hull_reach: the hull depth when it has a hull triangle
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
    // this frame's pair islands (islands of several bodies, or of one self-colliding body), by their first body,
    // members in body order: CSR
    std::vector<uint32_t> pair_off, pair_mem;
    size_t npair_islands() const { return pair_off.empty() ? 0 : pair_off.size() - 1; }
    int interval = 1;
    int* err_word = nullptr; // the context's device error word (the response raises bit 1 on a wait that gives up)
    RStat last{};
    // counters (bl_stats)
    int64_t pair_rebuilds = 0, contact_pairs = 0, narrow_tests = 0;

    int setup(const HostView& hv, uint32_t nnodes, uint32_t ntris, uint32_t ncaps);
    void build_islands(const HostView& hv, float frame_time);
    // the frame's pair-island tables to the device (after build_islands)
    int begin_frame(hipStream_t st);
    // the pair lists anew from the state of `parity`: bounds keys per body in `body_box`, top speeds in
    // a.body_frame; waits for the device once (and again when a list outgrew its buffer)
    int rebuild(const KArgs& a, const uint32_t* body_box, hipStream_t st, int parity);
    int collide(const KArgs& a, hipStream_t st, int parity);
    // The narrow phase and response sized for the lists' capacities, the counts read on the device: launches that
    // stay valid from one rebuild to the next (for a captured graph) until a list grows. `p` and `nblk` identify
    // them; the host's test counter is advanced by `count_tests`.
    int collide_fixed(PArgs& p, uint32_t& nblk);
    int launch_collide(const KArgs& a, const PArgs& p, uint32_t nblk, hipStream_t st, int parity);
    // The narrow phase's arguments for the current lists (nblk = 0: no candidates), and the response alone (after a
    // narrow phase run by the force phase's launch, launch_forces_narrow)
    int narrow_args(PArgs& p, uint32_t& nblk);
    int launch_response(const KArgs& a, const PArgs& p, uint32_t nblk, hipStream_t st, int parity);
    void count_tests(int64_t substeps) { narrow_tests += substeps * (int64_t)(nnt + nnc + nct); }
    void release();
    ~ContactEngine() { release(); }

private:
    uint32_t nnodes = 0, ntris = 0, ncaps = 0, nbodies = 0, nsamples = 0;
    size_t nnt = 0, nnc = 0, nct = 0;
    // the island building's scratch, and the tables last uploaded (a frame whose tables are the same uploads nothing)
    std::vector<int> parent_, root_isl_, body_isl_, pair_id_, isl_first_, isl_;
    std::vector<uint32_t> isl_size_, cursor_;
    std::vector<float> bx_;
    PairScratch pair_ws_;
    std::vector<uint32_t> h_tri0_, itri_h_; // per body: first triangle; the pair islands' triangles
    std::vector<uint32_t> h_cap0_, icap_h_; // per body: first capsule; the pair islands' capsules
    Buf<uint32_t> itri, icap, irange;
    Buf<int> igroup;
    std::vector<int> igroup_h_;
    uint32_t n_itri = 0, n_icap = 0, gresp_ = 0, ngroup_ = 0;
    std::vector<int> up_isl_, up_mi_, up_group_;
    std::vector<uint32_t> up_off_, up_mem_;
    bool up_valid_ = false;
    bool counts_dirty_ = true; // the bucket counts need clearing before the next rebuild
    uint32_t tsize = 0;   // node table buckets allocated
    uint32_t tsize_f = 0; // this frame's buckets: about four per node of the pair islands' bodies
    std::vector<uint32_t> h_nnode_;
    // the one-pass launches' published words and workgroup tickets (tickets taken so far: lk_base_)
    Buf<uint64_t> lk_words;
    Buf<uint32_t> lk_ticket;
    uint32_t lk_epoch_ = 0, lk_base_ = 0;
    bool lk_tickets_ = false; // order the one-pass launches' workgroups by tickets (else by index)
    RStat* hstat = nullptr; // pinned host copy of the rebuild's status
    Buf<int> bisl, bmem;
    Buf<uint32_t> pcnt, poff, plist;
    uint32_t nmem = 0, cap_part = 0;
    Buf<uint32_t> imem_off, imem, bquery, body_tri0, sample_base, icount;
    Buf<float> bbox, samples, extra;
    Buf<uint32_t> body_cap0, fqoff;
    Buf<FastGrid> fgrid;
    Buf<uint4> fnt, fnc, ntm, ncm;
    uint32_t cap_fq = 0, cap_fnt = 0, cap_fnc = 0;
    Buf<IslandPar> ipar;
    Buf<RStat> stat;
    Buf<uint32_t> pbucket, tcount, tstart;
    Buf<int4> pcell;
    Buf<float4> ppos;
    Buf<uint32_t> eb, ecount, estart;
    Buf<int4> scell;
    Buf<float4> spos, ebb, sb;
    Buf<uint4> ea, sa;
    Buf<float2> ec, sc;
    Buf<uint32_t> bcnt;
    Buf<uint4> ntraw, nc, ct, nt;
    Buf<Hit> hraw;
    uint32_t cap_nt = 0, cap_nc = 0, cap_ct = 0, cap_ent = 0, esize = 0;

    int scan(const uint32_t* in, uint32_t* out, size_t n, hipStream_t st);
    int lookback(Lookback& k, size_t nb, hipStream_t st);
    int ensure_lists();
    int enqueue_rebuild(const KArgs& a, CArgs& c, hipStream_t st);
};

// The force phase with the narrow phase's nblk blocks after its chunks in the same launch (256-thread workgroups only:
// BL_ERR_ARG otherwise, and the caller launches the two apart).
int launch_forces_narrow(const KArgs& a, hipStream_t st, int parity, const bl_tuning& t, const PArgs& p, uint32_t nblk);

} // namespace blg
