/*
 * TEST INFRASTRUCTURE - the oracle (see bl_oracle.h). A plain-C restatement of the reference's CPU step.
 * Compiled with -O2 -ffp-contract=off (oracle/Makefile) so that every float operation is the one written here, in
 * the order written, as the reference's C++ (compiled the same way) performs it. libm's expf/powf are the ones the
 * reference calls (std::exp / std::pow on floats).
 */
#include "bl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ vec3 (src/core/math.h:46-83) */
typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
/* operator/(float) multiplies by the reciprocal (src/core/math.h:57) */
static inline v3 vdivs(v3 a, float s) { float i = 1.0f / s; return V(a.x * i, a.y * i, a.z * i); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vabs3(v3 a) { return V(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
/* std::min / std::max (b < a ? b : a, a < b ? b : a): NaN behaviour as the reference's */
static inline float smin(float a, float b) { return b < a ? b : a; }
static inline float smax(float a, float b) { return a < b ? b : a; }
static inline float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float signf_(float v) { return v < 0.0f ? -1.0f : 1.0f; }
static inline v3 vmin3(v3 a, v3 b) { return V(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
static inline v3 vmax3(v3 a, v3 b) { return V(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }
static inline v3 ld3(const float* f) { return V(f[0], f[1], f[2]); }
static inline void st3(float* f, v3 a) { f[0] = a.x, f[1] = a.y, f[2] = a.z; }
/* mat3 * vec3 with columns c0 c1 c2 (src/core/math.h:101) */
static inline v3 mat_mul(const float* rot, v3 v) {
    return vadd(vadd(vmul(ld3(rot), v.x), vmul(ld3(rot + 3), v.y)), vmul(ld3(rot + 6), v.z));
}

/* ------------------------------------------------------------------ ground models (src/phys/static_world.cpp:85-115) */
typedef struct {
    float va, ms, mc, t2, vs, alpha, strength, soft_depth, fluid_drag, fluid_density;
} ground_model;

static const ground_model k_gm[BL_SURF_COUNT] = {
    /* va  ms     mc     t2       vs  alpha strength soft  drag  density */
    {3, 0.80f, 0.55f, 0.005f, 7, 2, 1, 0, 0, 0},         /* grass */
    {3, 0.80f, 0.58f, 0.006f, 4, 2, 1, 0, 0, 0},         /* dirt */
    {3, 1.20f, 0.75f, 0.010f, 6, 2, 1, 0, 0, 0},         /* asphalt */
    {3, 1.20f, 0.75f, 0.010f, 6, 2, 1, 0, 0, 0},         /* concrete */
    {3, 0.85f, 0.60f, 0.006f, 3, 2, 1, 0, 0, 0},         /* gravel */
    {2, 0.50f, 0.35f, 0.004f, 3, 2, 1, 0.12f, 60, 1500}, /* mud */
    {3, 0.70f, 0.55f, 0.0001f, 9, 2, 1, 0.03f, 20, 1600},/* sand */
    {3, 0.95f, 0.70f, 0.007f, 8, 2, 1, 0, 0, 0},         /* rock */
    {3, 0.90f, 0.60f, 0.005f, 5, 2, 1, 0, 0, 0},         /* wood */
    {3, 0.70f, 0.40f, 0.001f, 3, 2, 1, 0, 0, 0},         /* metal */
    {1, 0.30f, 0.20f, 0.0001f, 3, 2, 1, 0, 0, 0},        /* ice */
};

/* primitive_collision (src/phys/static_world.cpp:8-51) */
static v3 prim_coll(v3 F, v3 vel, float mass, v3 normal, float dt, const ground_model* gm, float pen, float friction_coef,
                    float push_max, float slop, float restitution) {
    v3 out = V(0, 0, 0);
    float Vn = vdot(vel, normal);
    float Fn = vdot(F, normal);
    float sgl = gm->soft_depth;
    if (sgl > 0 && pen >= 0) {
        out = vadd(out, vmul(vel, -gm->fluid_drag * smin(1.0f, pen / sgl + 0.2f)));
        float Fb = gm->fluid_density * smin(pen, sgl) * 9.807f * 0.02f;
        if (Vn >= 0 && Fn < 0 && Fb > -Fn) Fb = -Fn;
        out = vadd(out, vmul(normal, Fb));
    }
    if (pen >= sgl) {
        float R = -Fn;
        const float take = restitution >= 0 ? 1.0f + smin(restitution, 0.95f) : 0.8f;
        if (Vn < 0) R -= (take * Vn - smin(0.2f * smax(0.0f, pen - sgl - slop) / dt, push_max)) * mass / dt;
        if (R > 0) {
            v3 slipF = vsub(F, vmul(normal, Fn));
            v3 slip = vsub(vel, vmul(normal, Vn));
            float sv = vlen(slip);
            if (sv > 1e-9f) slip = vdivs(slip, sv);
            else slip = V(0, 0, 0);
            float G = R * gm->strength * friction_coef;
            const float stop = mass * sv / dt;
            if (sv < gm->va && G > 0 && vlen(slipF) <= gm->ms * G) {
                out = vadd(out, vsub(vsub(vmul(normal, R), vmul(slip, smin(gm->ms * G * (1.0f - expf(-sv / gm->va)), stop))), slipF));
            } else {
                float g = gm->mc + (gm->ms - gm->mc) * expf(-powf(sv / gm->vs, gm->alpha));
                out = vadd(out, vsub(vmul(normal, R), vmul(slip, smin((g + smin(gm->t2 * sv, 5.0f)) * G, smax(0.0f, stop + vdot(slipF, slip))))));
            }
        }
    }
    return out;
}

void orc_primitive_collision(const float F[3], const float vel[3], float mass, const float normal[3], float dt,
                             int surface, float pen, float friction_coef, float push_max, float slop,
                             float restitution, float out[3]) {
    const ground_model* gm = &k_gm[surface >= 0 && surface < BL_SURF_COUNT ? surface : 0];
    st3(out, prim_coll(ld3(F), ld3(vel), mass, ld3(normal), dt, gm, pen, friction_coef, push_max, slop, restitution));
}

/* ------------------------------------------------------------------ world */
typedef struct {
    uint32_t n, nb;
    bl_node* nodes;
    bl_node_info* info;
    bl_beam* beams;
    v3* force;
    bl_body_opts opts;
    int64_t broken;
    float max_speed;
    v3 mn, mx;
    int64_t repairs;
    bl_triangle* tris;
    uint32_t ntri;
    bl_capsule* caps;
    uint32_t ncap;
    float hull_reach; /* SoftBody::hull_reach (src/phys/softbody.h:751-758): hull_depth when any triangle is one-sided */
    bl_shock* shocks;
    uint32_t nsh;
    bl_slide* slides;
    uint32_t nsl;
    uint32_t* rail;
    bl_frame* frames;
    v3* torque; /* per frame (Frame::torque): cleared with the forces, summed by the joints */
    uint32_t nfr;
    bl_joint* joints;
    uint32_t nj;
    bl_wheel* wheels;
    uint32_t nw;
    uint32_t* wnodes;    /* the wheels' tyre nodes */
    int16_t* node_wheel; /* SoftBody::node_wheel: each node's wheel (the last wheel listing it), or -1 */
    float* ctl;          /* a controller's torque and brake per wheel (pre_substep), when `controlled` */
    int controlled;
    float wheel_speed, wheel_spin;
    float terrain_max; /* the terrain's top over the body's reach this frame (World::build_islands) */
    double guard_ke, guard_pe; /* the energy guard's motion energy at the frame's start */
    int guard_quiet, guard_cuts;
} orc_body;

struct orc_world {
    bl_settings s;
    bl_box* boxes;
    int nbox;
    bl_cylinder* cyl;
    int ncyl;
    /* the terrain (Heightfield) dense: heights and surfaces z-major, and each 16 x 16 block's top (update_bounds) */
    int tnx, tnz, tbx, tbz;
    float tcell, tox, toz;
    float* th;
    uint8_t* ts;
    float* tblock;
    orc_body* bodies;
    int nbodies;
    bl_stats st;
    int64_t fast_pairs; /* node-triangle and node-capsule candidates of fast bodies, summed over the rebuilds */
    int64_t last_rb[6]; /* summed over every rebuild: hashed nodes, triangles queried, NT before and after the near
                           filter, NC, CT (the measurement's algorithmic bytes) */
};

orc_world* orc_create(const bl_settings* s) {
    orc_world* w = (orc_world*)calloc(1, sizeof(orc_world));
    w->s = *s;
    return w;
}

void orc_destroy(orc_world* w) {
    if (!w) return;
    for (int i = 0; i < w->nbodies; i++) {
        free(w->bodies[i].nodes);
        free(w->bodies[i].info);
        free(w->bodies[i].beams);
        free(w->bodies[i].force);
        free(w->bodies[i].tris);
        free(w->bodies[i].caps);
        free(w->bodies[i].shocks);
        free(w->bodies[i].slides);
        free(w->bodies[i].rail);
        free(w->bodies[i].frames);
        free(w->bodies[i].torque);
        free(w->bodies[i].joints);
        free(w->bodies[i].wheels);
        free(w->bodies[i].wnodes);
        free(w->bodies[i].node_wheel);
        free(w->bodies[i].ctl);
    }
    free(w->bodies);
    free(w->boxes);
    free(w->cyl), free(w->th), free(w->ts), free(w->tblock);
    free(w);
}

int orc_set_statics(orc_world* w, const bl_box* boxes, int n) {
    free(w->boxes);
    w->boxes = (bl_box*)malloc(sizeof(bl_box) * (n > 0 ? n : 1));
    if (n > 0) memcpy(w->boxes, boxes, sizeof(bl_box) * n);
    w->nbox = n;
    return 0;
}

int orc_set_cylinders(orc_world* w, const bl_cylinder* c, int n) {
    free(w->cyl);
    w->cyl = (bl_cylinder*)malloc(sizeof(bl_cylinder) * (n > 0 ? n : 1));
    if (n > 0) memcpy(w->cyl, c, sizeof(bl_cylinder) * n);
    w->ncyl = n > 0 ? n : 0;
    return 0;
}

/* Heightfield::create + the heights and surfaces + update_bounds (src/phys/static_world.cpp:117-180) */
int orc_set_terrain(orc_world* w, int nx, int nz, float cell, const float* origin, const float* h, const uint8_t* surf) {
    free(w->th), free(w->ts), free(w->tblock);
    w->th = NULL, w->ts = NULL, w->tblock = NULL;
    w->tnx = w->tnz = 0;
    if (nx < 2 || nz < 2) return 0;
    const size_t n = (size_t)nx * nz;
    w->tnx = nx, w->tnz = nz, w->tcell = cell, w->tox = origin[0], w->toz = origin[1];
    w->th = (float*)malloc(sizeof(float) * n);
    w->ts = (uint8_t*)calloc(n, 1);
    memcpy(w->th, h, sizeof(float) * n);
    if (surf) memcpy(w->ts, surf, n);
    const int K = 16;
    w->tbx = (nx - 1 + K - 1) / K, w->tbz = (nz - 1 + K - 1) / K;
    w->tblock = (float*)malloc(sizeof(float) * w->tbx * w->tbz);
    for (int q = 0; q < w->tbx * w->tbz; q++) w->tblock[q] = -1e30f;
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            const float v = w->th[(size_t)z * nx + x];
            int bx0 = x - 1 < 0 ? 0 : (x - 1) / K, bx1 = x / K < w->tbx - 1 ? x / K : w->tbx - 1;
            int bz0 = z - 1 < 0 ? 0 : (z - 1) / K, bz1 = z / K < w->tbz - 1 ? z / K : w->tbz - 1;
            for (int bz = bz0; bz <= bz1; bz++)
                for (int bx = bx0; bx <= bx1; bx++) {
                    float* m = &w->tblock[(size_t)bz * w->tbx + bx];
                    *m = smax(*m, v);
                }
        }
    return 0;
}

/* Heightfield::max_height (src/phys/static_world.cpp:220-231) */
static float terrain_max_height(const orc_world* w, float x0, float z0, float x1, float z1) {
    int bx0 = (int)floorf((x0 - w->tox) / w->tcell) / 16, bx1 = (int)floorf((x1 - w->tox) / w->tcell) / 16;
    int bz0 = (int)floorf((z0 - w->toz) / w->tcell) / 16, bz1 = (int)floorf((z1 - w->toz) / w->tcell) / 16;
    bx0 = bx0 < 0 ? 0 : bx0, bz0 = bz0 < 0 ? 0 : bz0;
    bx1 = bx1 < w->tbx - 1 ? bx1 : w->tbx - 1, bz1 = bz1 < w->tbz - 1 ? bz1 : w->tbz - 1;
    float m = -1e30f;
    for (int bz = bz0; bz <= bz1; bz++)
        for (int bx = bx0; bx <= bx1; bx++) m = smax(m, w->tblock[(size_t)bz * w->tbx + bx]);
    return m;
}

static inline v3 vnormalize_(v3 a);

/* Heightfield::sample (src/phys/static_world.cpp:189-205) */
static int terrain_sample(const orc_world* w, float x, float z, float* out_h, v3* n) {
    float lx = (x - w->tox) / w->tcell, lz = (z - w->toz) / w->tcell;
    if (!(lx >= 0 && lz >= 0 && lx < w->tnx - 1 && lz < w->tnz - 1)) return 0;
    int ix = (int)lx, iz = (int)lz;
    float fx = lx - ix, fz = lz - iz;
    const float* H = w->th;
    const size_t r0 = (size_t)iz * w->tnx, r1 = r0 + w->tnx;
    float h00 = H[r0 + ix], h10 = H[r0 + ix + 1], h01 = H[r1 + ix], h11 = H[r1 + ix + 1];
    if (smin(smin(h00, h10), smin(h01, h11)) < -1e29f) return 0;
    if (fx > fz) {
        *out_h = h00 + (h10 - h00) * fx + (h11 - h10) * fz;
        *n = vnormalize_(V(-(h10 - h00) / w->tcell, 1.0f, -(h11 - h10) / w->tcell));
    } else {
        *out_h = h00 + (h11 - h01) * fx + (h01 - h00) * fz;
        *n = vnormalize_(V(-(h11 - h01) / w->tcell, 1.0f, -(h01 - h00) / w->tcell));
    }
    return 1;
}

/* Heightfield::surface_at (src/phys/static_world.cpp:213-218) */
static int terrain_surface(const orc_world* w, float x, float z) {
    int ix = (int)lroundf((x - w->tox) / w->tcell), iz = (int)lroundf((z - w->toz) / w->tcell);
    ix = ix < 0 ? 0 : (ix > w->tnx - 1 ? w->tnx - 1 : ix);
    iz = iz < 0 ? 0 : (iz > w->tnz - 1 ? w->tnz - 1 : iz);
    return w->ts[(size_t)iz * w->tnx + ix];
}

int orc_add_body(orc_world* w, const bl_body_desc* d) {
    w->bodies = (orc_body*)realloc(w->bodies, sizeof(orc_body) * (w->nbodies + 1));
    orc_body* b = &w->bodies[w->nbodies];
    memset(b, 0, sizeof(*b));
    b->n = d->n_nodes, b->nb = d->n_beams;
    b->nodes = (bl_node*)malloc(sizeof(bl_node) * (b->n ? b->n : 1));
    b->info = (bl_node_info*)malloc(sizeof(bl_node_info) * (b->n ? b->n : 1));
    b->beams = (bl_beam*)malloc(sizeof(bl_beam) * (b->nb ? b->nb : 1));
    b->force = (v3*)calloc(b->n ? b->n : 1, sizeof(v3));
    if (b->n) memcpy(b->nodes, d->nodes, sizeof(bl_node) * b->n), memcpy(b->info, d->info, sizeof(bl_node_info) * b->n);
    if (b->nb) memcpy(b->beams, d->beams, sizeof(bl_beam) * b->nb);
    b->opts = d->opts;
    b->ntri = d->n_tris, b->ncap = d->n_capsules;
    b->tris = (bl_triangle*)malloc(sizeof(bl_triangle) * (b->ntri ? b->ntri : 1));
    b->caps = (bl_capsule*)malloc(sizeof(bl_capsule) * (b->ncap ? b->ncap : 1));
    if (b->ntri) memcpy(b->tris, d->tris, sizeof(bl_triangle) * b->ntri);
    if (b->ncap) memcpy(b->caps, d->capsules, sizeof(bl_capsule) * b->ncap);
    b->nsh = d->n_shocks, b->nsl = d->n_slides;
    b->shocks = (bl_shock*)malloc(sizeof(bl_shock) * (b->nsh ? b->nsh : 1));
    b->slides = (bl_slide*)malloc(sizeof(bl_slide) * (b->nsl ? b->nsl : 1));
    b->rail = (uint32_t*)malloc(sizeof(uint32_t) * (d->n_rail ? d->n_rail : 1));
    if (b->nsh) memcpy(b->shocks, d->shocks, sizeof(bl_shock) * b->nsh);
    if (b->nsl) memcpy(b->slides, d->slides, sizeof(bl_slide) * b->nsl);
    if (d->n_rail) memcpy(b->rail, d->rail, sizeof(uint32_t) * d->n_rail);
    b->nfr = d->n_frames, b->nj = d->n_joints;
    b->frames = (bl_frame*)malloc(sizeof(bl_frame) * (b->nfr ? b->nfr : 1));
    b->torque = (v3*)calloc(b->nfr ? b->nfr : 1, sizeof(v3));
    b->joints = (bl_joint*)malloc(sizeof(bl_joint) * (b->nj ? b->nj : 1));
    if (b->nfr) memcpy(b->frames, d->frames, sizeof(bl_frame) * b->nfr);
    if (b->nj) memcpy(b->joints, d->joints, sizeof(bl_joint) * b->nj);
    b->nw = d->n_wheels;
    b->wheels = (bl_wheel*)malloc(sizeof(bl_wheel) * (b->nw ? b->nw : 1));
    b->wnodes = (uint32_t*)malloc(sizeof(uint32_t) * (d->n_wheel_nodes ? d->n_wheel_nodes : 1));
    b->ctl = (float*)calloc(2 * (b->nw ? b->nw : 1), sizeof(float));
    if (b->nw) memcpy(b->wheels, d->wheels, sizeof(bl_wheel) * b->nw);
    if (d->n_wheel_nodes) memcpy(b->wnodes, d->wheel_nodes, sizeof(uint32_t) * d->n_wheel_nodes);
    /* (World::collide_static's node_wheel, src/phys/world.cpp:873-877: a later wheel listing a node takes it) */
    b->node_wheel = (int16_t*)malloc(sizeof(int16_t) * (b->n ? b->n : 1));
    for (uint32_t i = 0; i < b->n; i++) b->node_wheel[i] = -1;
    for (uint32_t k = 0; k < b->nw; k++)
        for (uint32_t j = 0; j < b->wheels[k].n_nodes; j++) b->node_wheel[b->wnodes[b->wheels[k].node0 + j]] = (int16_t)k;
    int hulls = 0;
    for (uint32_t t = 0; t < b->ntri; t++) hulls += !b->tris[t].two_sided;
    b->hull_reach = hulls > 0 ? b->opts.hull_depth : 0.0f;
    /* SoftBody::compute_aabb at World::add_body (src/phys/world.cpp:283-289, src/phys/softbody.cpp:271-275) */
    b->mn = V(1e30f, 1e30f, 1e30f), b->mx = V(-1e30f, -1e30f, -1e30f);
    for (uint32_t i = 0; i < b->n; i++) b->mn = vmin3(b->mn, ld3(b->nodes[i].p)), b->mx = vmax3(b->mx, ld3(b->nodes[i].p));
    return w->nbodies++;
}

/* SoftBody::compute_beam_forces (src/phys/softbody.cpp:412-470) */
static void beam_forces(orc_body* B) {
    bl_node* nd = B->nodes;
    v3* f = B->force;
    const int deform = B->opts.allow_deform != 0, brk = B->opts.allow_break != 0, ductile = B->opts.ductile != 0;
    for (uint32_t j = 0; j < B->nb; j++) {
        bl_beam* b = &B->beams[j];
        if (b->flags & (BL_BF_BROKEN | BL_BF_SHOCK)) continue;
        const bl_node* na = &nd[b->a];
        const bl_node* nb = &nd[b->b];
        v3 dis = vsub(ld3(na->p), ld3(nb->p));
        float len2 = vdot(dis, dis);
        if (len2 < 1e-12f) continue;
        float inv_len = 1.0f / sqrtf(len2);
        float len = len2 * inv_len;
        float diff = len - b->L;
        v3 dv = vsub(ld3(na->v), ld3(nb->v));
        float k = b->k, d = b->d;
        if (b->type == BL_BT_ROPE) {
            if (diff < 0) { k = 0; d *= 0.1f; }
        } else if (b->type == BL_BT_SUPPORT) {
            if (diff > 0) {
                k = 0;
                d *= 0.1f;
                if (brk && diff > b->L * b->support_limit) { b->flags |= BL_BF_BROKEN; B->broken++; continue; }
            }
        }
        float slen = -k * diff - d * vdot(dv, dis) * inv_len;
        if (deform && !(b->flags & BL_BF_NO_DEFORM) && k != 0) {
            if (slen > b->max_pos && diff < 0) {
                float yield_len = b->max_pos / k;
                float def = diff + yield_len * (1.0f - b->plastic);
                float Lold = b->L;
                b->L = smax(smin(0.1f, 0.3f * b->L0), b->L + def);
                slen = slen - (slen - b->max_pos) * 0.5f;
                if (b->L < Lold) b->max_pos *= Lold / b->L;
            } else if (slen < b->max_neg && diff > 0) {
                float yield_len = b->max_neg / k;
                float def = diff + yield_len * (1.0f - b->plastic);
                float Lold = b->L;
                b->L = b->L + def;
                slen = slen - (slen - b->max_neg) * 0.5f;
                if (Lold > 0 && b->L > Lold) b->max_neg *= b->L / Lold;
                if (!ductile) b->strength -= def * k;
            }
        }
        if (brk && fabsf(slen) > b->strength && !(b->flags & BL_BF_NO_BREAK)) {
            b->flags |= BL_BF_BROKEN;
            B->broken++;
            continue;
        }
        b->stress = slen;
        v3 fv = vmul(dis, slen * inv_len);
        f[b->a] = vadd(f[b->a], fv);
        f[b->b] = vsub(f[b->b], fv);
    }
}

/* SoftBody::compute_shock_forces (src/phys/softbody.cpp:472-572): each shock's beam in shock order, after the beams */
static inline float sqrf(float x) { return x * x; }
typedef struct { const bl_shock* s; float diff, L; } shock_ctx;
/* shocks2's progressive rate, one way (the prog lambda, src/phys/softbody.cpp:503-517) */
static void shock_prog(const shock_ctx* c, int out, float* kk, float* dd) {
    const bl_shock* s = c->s;
    if (out) {
        *kk = s->spring_out;
        *dd = s->damp_out;
        float fr = s->long_bound != 0 ? smin(sqrf(c->diff / (s->long_bound * c->L)), 1.0f) : 1.0f;
        *kk += s->prog_spring_out * *kk * fr;
        *dd += s->prog_damp_out * *dd * fr;
    } else {
        *kk = s->spring_in;
        *dd = s->damp_in;
        float fr = s->short_bound != 0 ? smin(sqrf(c->diff / (s->short_bound * c->L)), 1.0f) : 1.0f;
        *kk += s->prog_spring_in * *kk * fr;
        *dd += s->prog_damp_in * *dd * fr;
    }
}

static void shock_forces(orc_body* B) {
    bl_node* nd = B->nodes;
    v3* f = B->force;
    for (uint32_t q = 0; q < B->nsh; q++) {
        const bl_shock* s = &B->shocks[q];
        bl_beam* b = &B->beams[s->beam];
        if (b->flags & BL_BF_BROKEN) continue;
        const bl_node* na = &nd[b->a];
        const bl_node* nb = &nd[b->b];
        v3 dis = vsub(ld3(na->p), ld3(nb->p));
        float len2 = vdot(dis, dis);
        if (len2 < 1e-12f) continue;
        float inv_len = 1.0f / sqrtf(len2);
        float len = len2 * inv_len;
        const float L = b->L;
        float diff = len - L;
        float v = vdot(vsub(ld3(na->v), ld3(nb->v)), dis) * inv_len;
        float k, d;
        const float lb = s->long_bound * L, sb = s->short_bound * L;
        if (s->type == 1) {
            k = s->spring;
            d = s->damp;
            float r = 0;
            if (diff > lb) r = diff - lb;
            else if (diff < -sb) r = -diff - sb;
            if (r != 0) {
                r = smin(r, 1.0f);
                k += (s->bound_spring - k) * r;
                d += (s->bound_damp - d) * r;
            }
        } else if (s->type == 2) {
            const shock_ctx c = {s, diff, L};
            shock_prog(&c, v > 0, &k, &d);
            if (s->soft_bump) {
                float Lp = 0.8f * L, lpre = s->long_bound * Lp, spre = -s->short_bound * Lp;
                if (diff > lpre) {
                    shock_prog(&c, 1, &k, &d);
                    float f2 = s->long_bound != 0 ? smin(sqrf((diff - lpre) * 5.0f / (s->long_bound * L)), 1.0f) : 1.0f;
                    k += (k + 100) * s->prog_spring_out * f2;
                    d += (d + 100) * s->prog_damp_out * f2;
                    if (v < 0) { k = s->spring_in; d = s->damp_in; }
                } else if (diff < spre) {
                    shock_prog(&c, 0, &k, &d);
                    float f2 = s->short_bound != 0 ? smin(sqrf((diff - spre) * 5.0f / (s->short_bound * L)), 1.0f) : 1.0f;
                    k += (k + 100) * s->prog_spring_out * f2;
                    d += (d + 100) * s->prog_damp_out * f2;
                    if (v > 0) { k = s->spring_out; d = s->damp_out; }
                }
                if (diff > lb || diff < -sb) {
                    k = smax(k, s->bound_spring);
                    d = smax(d, s->bound_damp);
                }
            } else if (diff > lb || diff < -sb) {
                k = s->bound_spring;
                d = s->bound_damp;
            }
        } else {
            k = s->spring_in;
            d = s->damp_in;
            if (diff > lb) {
                float r = smin(diff - lb, 1.0f);
                k += (s->bound_spring - k) * r;
                d += (s->bound_damp - d) * r;
            } else if (diff < -sb) {
                float r = smin(-diff - sb, 1.0f);
                k += (s->bound_spring - k) * r;
                d += (s->bound_damp - d) * r;
            } else if (v > 0) {
                float wv = clampf_(fabsf(v), 0.15f, 20.0f);
                k = s->spring_out;
                d = (s->damp_out * s->damp_out_slow * smin(wv, s->split_out) + s->damp_out * s->damp_out_fast * smax(0.0f, wv - s->split_out)) / wv;
            } else if (v < 0) {
                float wv = clampf_(fabsf(v), 0.15f, 20.0f);
                k = s->spring_in;
                d = (s->damp_in * s->damp_in_slow * smin(wv, s->split_in) + s->damp_in * s->damp_in_fast * smax(0.0f, wv - s->split_in)) / wv;
            }
        }
        float slen = -k * diff - d * v;
        if (B->opts.allow_break && fabsf(slen) > b->strength && !(b->flags & BL_BF_NO_BREAK)) {
            b->flags |= BL_BF_BROKEN;
            B->broken++;
            continue;
        }
        b->stress = slen;
        v3 fv = vmul(dis, slen * inv_len);
        f[b->a] = vadd(f[b->a], fv);
        f[b->b] = vsub(f[b->b], fv);
    }
}

/* SoftBody::compute_slide_forces (src/phys/softbody.cpp:574-623): each slide node pulled onto the nearest point of its
 * rail segment (the current one and its neighbours; the whole rail the first time), the reaction split over the
 * segment's ends */
static float slide_seg_dist2(const bl_node* nd, const uint32_t* rail, int i, v3 p, float* t) {
    v3 a = ld3(nd[rail[i]].p), b = ld3(nd[rail[i + 1]].p), ab = vsub(b, a);
    float l2 = vdot(ab, ab);
    *t = l2 > 1e-12f ? clampf_(vdot(vsub(p, a), ab) / l2, 0.0f, 1.0f) : 0.0f;
    v3 q = vsub(vadd(a, vmul(ab, *t)), p);
    return vdot(q, q);
}

static void slide_forces(orc_body* B) {
    bl_node* nd = B->nodes;
    v3* f = B->force;
    for (uint32_t q = 0; q < B->nsl; q++) {
        bl_slide* s = &B->slides[q];
        if (s->broken || s->nrail < 2) continue;
        const uint32_t* rail = B->rail + s->rail0;
        const v3 p = ld3(nd[s->node].p);
        const int nseg = (int)s->nrail - 1;
        float t, best_t = 0;
        if (s->seg < 0) {
            float best = 1e30f;
            for (int i = 0; i < nseg; i++) {
                float dd = slide_seg_dist2(nd, rail, i, p, &t);
                if (dd < best) { best = dd; s->seg = i; best_t = t; }
            }
        }
        float best = slide_seg_dist2(nd, rail, s->seg, p, &best_t);
        int bi = s->seg;
        if (s->seg > 0) {
            float dd = slide_seg_dist2(nd, rail, s->seg - 1, p, &t);
            if (dd < best) { best = dd; bi = s->seg - 1; best_t = t; }
        }
        if (s->seg + 1 < nseg) {
            float dd = slide_seg_dist2(nd, rail, s->seg + 1, p, &t);
            if (dd < best) { best = dd; bi = s->seg + 1; best_t = t; }
        }
        s->seg = bi;
        uint32_t na = rail[bi], nb = rail[bi + 1];
        v3 ideal = vadd(ld3(nd[na].p), vmul(vsub(ld3(nd[nb].p), ld3(nd[na].p)), best_t));
        v3 d = vsub(ideal, p);
        float dist = vlen(d);
        float stretch = dist - s->threshold;
        if (stretch <= 0 || dist < 1e-9f) continue;
        v3 F = vmul(d, s->k * stretch / dist);
        if (s->break_force > 0 && vlen(F) > s->break_force) {
            s->broken = 1;
            continue;
        }
        f[s->node] = vadd(f[s->node], F);
        f[na] = vsub(f[na], vmul(F, 1.0f - best_t));
        f[nb] = vsub(f[nb], vmul(F, best_t));
    }
}

/* ------------------------------------------------------------------ quaternions (src/core/math.h:131-222) */
typedef struct { float x, y, z, w; } qt;
static inline qt Q(float x, float y, float z, float w) { qt r = {x, y, z, w}; return r; }
static inline qt qmul(qt a, qt b) {
    return Q(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
static inline qt qscale(qt a, float s) { return Q(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline qt qadd(qt a, qt b) { return Q(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline qt qconj(qt a) { return Q(-a.x, -a.y, -a.z, a.w); }
static inline qt qnormalize(qt a) {
    float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
    return l > 1e-12f ? qscale(a, 1.0f / l) : Q(0, 0, 0, 1);
}
static inline v3 vcross_(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 qrotate(qt q, v3 p) {
    v3 u = V(q.x, q.y, q.z);
    v3 t = vmul(vcross_(u, p), 2.0f);
    return vadd(vadd(p, vmul(t, q.w)), vcross_(u, t));
}
static inline v3 vnormalize_(v3 a) { float l = vlen(a); return l > 1e-12f ? vmul(a, 1.0f / l) : V(0, 0, 0); }
static qt qaxis_angle(v3 axis, float angle) {
    float h = angle * 0.5f, sn = sinf(h);
    v3 a = vnormalize_(axis);
    return Q(a.x * sn, a.y * sn, a.z * sn, cosf(h));
}
static v3 qlog(qt qin) {
    qt q = qin.w < 0 ? qscale(qin, -1.0f) : qin;
    v3 v = V(q.x, q.y, q.z);
    float sn = vlen(v);
    if (sn < 1e-8f) return vmul(v, 2.0f);
    float angle = 2.0f * atan2f(sn, q.w);
    return vmul(v, angle / sn);
}
static inline qt qload(const float* f) { return Q(f[0], f[1], f[2], f[3]); }
static inline void qstore(float* f, qt q) { f[0] = q.x, f[1] = q.y, f[2] = q.z, f[3] = q.w; }

/* SoftBody::compute_joint_forces (src/phys/softbody.cpp:625-679): each joint in order */
static void joint_forces(orc_body* B) {
    bl_node* nd = B->nodes;
    v3* f = B->force;
    const int deform = B->opts.allow_deform != 0, brk = B->opts.allow_break != 0;
    for (uint32_t q = 0; q < B->nj; q++) {
        bl_joint* j = &B->joints[q];
        if (j->broken) continue;
        bl_frame* pf = &B->frames[j->parent_frame];
        const bl_node* pn = &nd[pf->node];
        const bl_node* cn = &nd[j->child_node];
        const qt pq = qload(pf->q);
        v3 r = qrotate(pq, ld3(j->local_offset));
        v3 err = vsub(ld3(cn->p), vadd(ld3(pn->p), r));
        if (j->break_force > 0 && deform) {
            float ylen = j->break_force / j->k_lin;
            float el = vlen(err);
            if (el > ylen) {
                float excess = el - ylen;
                v3 dir = vdivs(err, el);
                st3(j->local_offset, vadd(ld3(j->local_offset), qrotate(qconj(pq), vmul(dir, excess))));
                r = qrotate(pq, ld3(j->local_offset));
                err = vmul(dir, ylen);
                j->plastic_lin += excess;
            }
        }
        v3 vrel = vsub(ld3(cn->v), vadd(ld3(pn->v), vcross_(ld3(pf->w), r)));
        v3 F = vsub(vmul(err, -j->k_lin), vmul(vrel, j->d_lin));
        f[j->child_node] = vadd(f[j->child_node], F);
        f[pf->node] = vsub(f[pf->node], F);
        v3* pt = &B->torque[j->parent_frame];
        *pt = vadd(*pt, vcross_(r, V(-F.x, -F.y, -F.z)));
        float sres = j->break_force > 0 ? vlen(err) * j->k_lin / j->break_force : 0.0f;
        if (j->child_frame >= 0) {
            bl_frame* cf = &B->frames[j->child_frame];
            qt qe = qmul(qload(cf->q), qconj(qmul(pq, qload(j->rel_rot))));
            v3 theta = qlog(qe);
            float ang = vlen(theta);
            float yield = j->yield_angle > 0 ? j->yield_angle : 1e9f;
            if (j->break_torque > 0) yield = smin(yield, j->break_torque / j->k_ang);
            if (ang > yield && deform) {
                float excess = ang - yield;
                v3 axis = vdivs(theta, ang);
                qt bend = qaxis_angle(axis, excess);
                qstore(j->rel_rot, qnormalize(qmul(qmul(qmul(qconj(pq), bend), pq), qload(j->rel_rot))));
                theta = vmul(axis, yield);
                j->plastic_ang += excess;
            }
            v3 T = vsub(vmul(theta, -j->k_ang), vmul(vsub(ld3(cf->w), ld3(pf->w)), j->d_ang));
            v3* ct = &B->torque[j->child_frame];
            *ct = vadd(*ct, T);
            *pt = vsub(*pt, T);
            if (j->break_torque > 0) sres = smax(sres, vlen(theta) * j->k_ang / j->break_torque);
        }
        j->stress = sres;
        int snap = (j->break_force > 0 && j->plastic_lin > j->max_stretch) || (j->break_torque > 0 && j->plastic_ang > j->max_bend);
        if (brk && snap) j->broken = 1;
    }
}

static inline v3 vnormalize_or_(v3 a, v3 fb) { float l = vlen(a); return l > 1e-12f ? vmul(a, 1.0f / l) : fb; }

/* from_mat3 (src/core/math.h:171-201): the rotation of the matrix with columns c0, c1, c2 */
static qt quat_from_cols(v3 c0, v3 c1, v3 c2) {
    const float m00 = c0.x, m11 = c1.y, m22 = c2.z;
    const float tr = m00 + m11 + m22;
    qt q;
    if (tr > 0) {
        float s = sqrtf(tr + 1.0f) * 2;
        q = Q((c1.z - c2.y) / s, (c2.x - c0.z) / s, (c0.y - c1.x) / s, 0.25f * s);
    } else if (m00 > m11 && m00 > m22) {
        float s = sqrtf(1.0f + m00 - m11 - m22) * 2;
        q = Q(0.25f * s, (c1.x + c0.y) / s, (c2.x + c0.z) / s, (c1.z - c2.y) / s);
    } else if (m11 > m22) {
        float s = sqrtf(1.0f + m11 - m00 - m22) * 2;
        q = Q((c1.x + c0.y) / s, 0.25f * s, (c2.y + c1.z) / s, (c2.x - c0.z) / s);
    } else {
        float s = sqrtf(1.0f + m22 - m00 - m11) * 2;
        q = Q((c2.x + c0.z) / s, (c2.y + c1.z) / s, 0.25f * s, (c0.y - c1.x) / s);
    }
    return qnormalize(q);
}

/* SoftBody::triad_orientation (src/phys/softbody.cpp:997-1004): x towards ref_x, z normal to the plane of ref_y */
static qt triad_orientation(const orc_body* B, uint32_t node, int32_t ref_x, int32_t ref_y) {
    const v3 p = ld3(B->nodes[node].p);
    const v3 e1 = vnormalize_or_(vsub(ld3(B->nodes[ref_x].p), p), V(1, 0, 0));
    v3 e3 = vnormalize_or_(vcross_(e1, vsub(ld3(B->nodes[ref_y].p), p)), V(0, 0, 1));
    const v3 t = vcross_(e1, e3);
    if (vdot(t, t) < 1e-12f) e3 = vnormalize_or_(vcross_(e1, V(0, 1, 0)), V(0, 0, 1));
    const v3 e2 = vcross_(e3, e1);
    return qnormalize(quat_from_cols(e1, e2, e3));
}

/* SoftBody::integrate_frames (src/phys/softbody.cpp:1006-1021): a kinematic frame takes its triad's orientation,
 * a dynamic one integrates its torque */
static void integrate_frames(orc_body* B, float dt) {
    for (uint32_t k = 0; k < B->nfr; k++) {
        bl_frame* fr = &B->frames[k];
        if (fr->ref_x >= 0 && fr->ref_y >= 0 && fr->ref_x < (int32_t)B->n && fr->ref_y < (int32_t)B->n) {
            qstore(fr->q, triad_orientation(B, fr->node, fr->ref_x, fr->ref_y));
            fr->w[0] = fr->w[1] = fr->w[2] = 0;
            B->torque[k] = V(0, 0, 0);
            continue;
        }
        if (fr->inv_inertia <= 0) continue;
        v3 w = vadd(ld3(fr->w), vmul(B->torque[k], fr->inv_inertia * dt));
        w = vmul(w, 1.0f - 0.5f * dt);
        st3(fr->w, w);
        qt dq = Q(w.x, w.y, w.z, 0);
        qt q = qload(fr->q);
        q = qadd(q, qscale(qmul(dq, q), 0.5f * dt));
        qstore(fr->q, qnormalize(q));
    }
}

/* StaticWorld::collide_point (src/phys/static_world.cpp:364-474): the terrain (without road or paving) when near it,
 * the boxes in index order, the cylinders */
static int collide_point(const orc_world* w, v3 p, float r, int near_terrain, v3* normal, float* depth_out,
                         int* surface, float* max_force) {
    int hit = 0;
    float out_depth = 0;
    *max_force = 0;
    if (near_terrain && w->tnx) {
        float th;
        v3 n;
        if (terrain_sample(w, p.x, p.z, &th, &n)) {
            float d = (th - p.y) * n.y + r;
            if (d > 0) {
                out_depth = d;
                *normal = n;
                *surface = terrain_surface(w, p.x, p.z);
                hit = 1;
            }
        }
    }
    for (int k = 0; k < w->nbox; k++) {
        const bl_box* b = &w->boxes[k];
        if (p.x < b->aabb_mn[0] - r || p.x > b->aabb_mx[0] + r || p.y < b->aabb_mn[1] - r || p.y > b->aabb_mx[1] + r ||
            p.z < b->aabb_mn[2] - r || p.z > b->aabb_mx[2] + r)
            continue;
        v3 rel = vsub(p, ld3(b->center));
        v3 l = V(vdot(rel, ld3(b->rot)), vdot(rel, ld3(b->rot + 3)), vdot(rel, ld3(b->rot + 6)));
        v3 q = V(clampf_(l.x, -b->half[0], b->half[0]), clampf_(l.y, -b->half[1], b->half[1]),
                 clampf_(l.z, -b->half[2], b->half[2]));
        v3 d = vsub(l, q);
        float dist2 = vdot(d, d);
        float depth;
        v3 nl;
        if (dist2 > 1e-12f) {
            if (dist2 >= r * r) continue;
            float dist = sqrtf(dist2);
            depth = r - dist;
            nl = vdivs(d, dist);
        } else {
            v3 pen = vsub(ld3(b->half), vabs3(l));
            if (pen.x <= pen.y && pen.x <= pen.z) { depth = pen.x; nl = V(signf_(l.x), 0, 0); }
            else if (pen.y <= pen.z) { depth = pen.y; nl = V(0, signf_(l.y), 0); }
            else { depth = pen.z; nl = V(0, 0, signf_(l.z)); }
            depth += r;
        }
        if (depth > out_depth) {
            out_depth = depth;
            *normal = mat_mul(b->rot, nl);
            *surface = (int)b->surface;
            *max_force = b->max_force;
            hit = 1;
        }
    }
    for (int k = 0; k < w->ncyl; k++) {
        const bl_cylinder* c = &w->cyl[k];
        v3 rel = vsub(p, ld3(c->base));
        if (rel.y < -r || rel.y > c->height + r) continue;
        if (fabsf(rel.x) >= c->radius + r || fabsf(rel.z) >= c->radius + r) continue;
        float dxz = sqrtf(rel.x * rel.x + rel.z * rel.z);
        float side = c->radius + r - dxz;
        float top = c->height + r - rel.y;
        if (side <= 0) continue;
        float depth;
        v3 n;
        if (top < side && rel.y > c->height * 0.5f) depth = top, n = V(0, 1, 0);
        else depth = side, n = dxz > 1e-6f ? V(rel.x / dxz, 0, rel.z / dxz) : V(1, 0, 0);
        /* (the deepest so far: a cylinder leaves max_force as it was) */
        if (depth > out_depth) {
            out_depth = depth;
            *normal = n;
            *surface = (int)c->surface;
            hit = 1;
        }
    }
    *depth_out = out_depth;
    return hit;
}

/* SoftBody::compute_wheel_forces (src/phys/softbody.cpp:689-767) with first = true (the controllers' substep), the
 * node wheels (ring tyres are refused); wheel_speed and wheel_spin over the forward-propelled ones */
static void wheel_forces(orc_body* B, float dt) {
    bl_node* nd = B->nodes;
    v3* f = B->force;
    float speed_sum = 0, spin_sum = 0;
    int nprop = 0;
    for (uint32_t k = 0; k < B->nw; k++) {
        bl_wheel* w = &B->wheels[k];
        if (w->detached || w->n_nodes == 0) continue;
        w->drive_torque = w->torque;
        if (w->brake > 0) {
            float stop = -w->avg_speed * w->radius * w->mass / dt - w->last_retorque;
            w->torque += w->speed > 0 ? clampf_(stop, -w->brake, 0.0f) : clampf_(stop, 0.0f, w->brake);
        }
        const bl_node* h0 = &nd[w->axle0];
        const bl_node* h1 = &nd[w->axle1];
        const v3 axis = vnormalize_(vsub(ld3(h1->p), ld3(h0->p)));
        const int nn = (int)w->n_nodes;
        const float per = w->torque / (float)nn;
        float expected = w->speed;
        float speed = 0;
        for (int j = 0; j < nn; j++) {
            const bl_node* hub = (j & 1) ? h1 : h0;
            const uint32_t ni = B->wnodes[w->node0 + j];
            v3 rad = vsub(ld3(nd[ni].p), ld3(hub->p));
            float rl2 = vdot(rad, rad);
            if (rl2 < 1e-8f) continue;
            float inv = 1.0f / sqrtf(rl2);
            if (w->propulsed == 2) rad = V(-rad.x, -rad.y, -rad.z);
            v3 dir = vmul(vcross_(axis, rad), inv);
            f[ni] = vadd(f[ni], vmul(dir, per * inv));
            speed += vdot(vsub(ld3(nd[ni].v), ld3(hub->v)), dir);
        }
        w->speed = speed / (float)nn;
        w->avg_speed = 0.99f * w->avg_speed + 0.1f * w->speed;
        if (w->propulsed == 1) {
            speed_sum += w->speed;
            spin_sum += w->speed / smax(0.05f, w->radius);
            nprop++;
        }
        if (w->mass > 0 && w->radius > 0) {
            expected += (w->last_torque / w->radius) / w->mass * dt;
            w->last_retorque = w->mass * (w->speed - expected) / dt;
        }
        if (w->arm >= 0 && w->near_attach >= 0 && fabsf(w->torque) > 0.01f) {
            const v3 rr = vsub(ld3(nd[w->arm].p), ld3(nd[w->near_attach].p));
            v3 r = vsub(rr, vmul(axis, vdot(rr, axis)));
            const float off = vlen(vsub(rr, r)), rl = vlen(r);
            if (rl > 0.01f && 2 * off < rl) {
                r = vdivs(r, rl);
                const v3 cf = vmul(vmul(vcross_(axis, r), 0.5f * w->torque / rl), 1.0f - 2.0f * off / rl);
                f[w->arm] = vsub(f[w->arm], cf);
                f[w->near_attach] = vadd(f[w->near_attach], cf);
            }
        }
        w->last_torque = w->torque;
        w->torque = 0;
    }
    if (nprop > 0) {
        B->wheel_speed = speed_sum / (float)nprop;
        B->wheel_spin = spin_sum / (float)nprop;
    }
}

/* tyre_contact (src/phys/static_world.cpp:53-80) */
typedef struct {
    v3 normal_force, need, slide;
    float cap;
    int touching;
} tyre_hit;
static tyre_hit tyre_contact(v3 F, v3 vel, float mass, v3 normal, float dt, const ground_model* gm, float pen,
                             float friction_coef) {
    tyre_hit t;
    memset(&t, 0, sizeof(t));
    float Vn = vdot(vel, normal);
    float Fn = vdot(F, normal);
    float sgl = gm->soft_depth;
    if (sgl > 0 && pen >= 0) {
        t.normal_force = vadd(t.normal_force, vmul(vel, -gm->fluid_drag * smin(1.0f, pen / sgl + 0.2f)));
        float Fb = gm->fluid_density * smin(pen, sgl) * 9.807f * 0.02f;
        if (Vn >= 0 && Fn < 0 && Fb > -Fn) Fb = -Fn;
        t.normal_force = vadd(t.normal_force, vmul(normal, Fb));
    }
    if (pen < sgl) return t;
    float R = -Fn;
    if (Vn < 0) R -= (0.8f * Vn - smin(0.2f * (pen - sgl) / dt, 2.0f)) * mass / dt;
    if (R <= 0) return t;
    t.touching = 1;
    t.normal_force = vadd(t.normal_force, vmul(normal, R));
    v3 slipF = vsub(F, vmul(normal, Fn));
    v3 slipV = vsub(vel, vmul(normal, Vn));
    float G = R * gm->strength * friction_coef;
    t.need = vsub(V(-slipF.x, -slipF.y, -slipF.z), vmul(slipV, 0.5f * mass / dt));
    t.cap = gm->ms * G;
    float sv = vlen(slipV);
    float mu_slide = smax(gm->mc, 0.85f * gm->ms);
    float g = mu_slide + (gm->ms - mu_slide) * expf(-powf(sv / gm->vs, gm->alpha)) + smin(gm->t2 * sv, 5.0f);
    float nl = vlen(t.need);
    v3 dir = sv > 0.05f ? vmul(slipV, -1.0f / sv) : (nl > 1e-6f ? vmul(t.need, 1.0f / nl) : V(0, 0, 0));
    t.slide = vmul(dir, g * G);
    return t;
}

/* World::collide_static for one body's nodes, boxes only (src/phys/world.cpp:781-869); tyre nodes by tyre_contact,
 * each wheel's touching tyre nodes then sharing its grip (:850-892) */
static int collide_static(orc_world* w, orc_body* B, float dt) {
    int contacts = 0;
    uint32_t ntyre = 0;
    uint32_t* tyre_node = NULL;
    tyre_hit* tyre = NULL;
    const float fric_body = B->opts.ground_friction * 1.0f; /* resting is set only by the rest damping (off) */
    for (uint32_t i = 0; i < B->n; i++) {
        bl_node* x = &B->nodes[i];
        const bl_node_info* in = &B->info[i];
        if (!(in->flags & BL_NF_GROUND) || x->inv_mass <= 0) continue;
        if (x->p[1] < w->s.kill_y) {
            x->v[0] = x->v[1] = x->v[2] = 0;
            x->inv_mass = 0;
            continue;
        }
        float r = in->radius;
        const int near_terrain = w->tnx && x->p[1] - r <= B->terrain_max + 0.05f;
        if (!near_terrain && w->nbox == 0 && w->ncyl == 0) continue;
        v3 n;
        float depth, max_force;
        int surf;
        if (!collide_point(w, ld3(x->p), r, near_terrain, &n, &depth, &surf, &max_force)) continue;
        contacts++;
        const ground_model* gm = &k_gm[surf < BL_SURF_COUNT ? surf : 0];
        if (max_force > 0) {
            v3 f = prim_coll(B->force[i], ld3(x->v), x->mass, n, dt, gm, depth, in->friction * fric_body,
                             B->opts.contact_push_max, B->opts.contact_slop, -1.0f);
            float fl = vlen(f);
            if (fl > max_force) f = vmul(f, max_force / fl);
            B->force[i] = vadd(B->force[i], f);
        } else if (in->flags & BL_NF_TYRE) {
            tyre_hit t = tyre_contact(B->force[i], ld3(x->v), x->mass, n, dt, gm, depth,
                                      in->friction * fric_body * w->s.tyre_grip);
            B->force[i] = vadd(B->force[i], t.normal_force);
            if (t.touching) {
                tyre_node = (uint32_t*)realloc(tyre_node, sizeof(uint32_t) * (ntyre + 1));
                tyre = (tyre_hit*)realloc(tyre, sizeof(tyre_hit) * (ntyre + 1));
                tyre_node[ntyre] = i, tyre[ntyre] = t;
                ntyre++;
            }
        } else {
            v3 f = prim_coll(B->force[i], ld3(x->v), x->mass, n, dt, gm, depth, in->friction * fric_body,
                             B->opts.contact_push_max, B->opts.contact_slop, B->opts.bounce);
            B->force[i] = vadd(B->force[i], f);
        }
    }
    if (ntyre) {
        v3* need_sum = (v3*)calloc(B->nw ? B->nw : 1, sizeof(v3));
        float* cap_sum = (float*)calloc(B->nw ? B->nw : 1, sizeof(float));
        for (uint32_t q = 0; q < ntyre; q++) {
            const int wk = B->node_wheel[tyre_node[q]];
            if (wk >= 0) need_sum[wk] = vadd(need_sum[wk], tyre[q].need), cap_sum[wk] += tyre[q].cap;
        }
        for (uint32_t q = 0; q < ntyre; q++) {
            const int wk = B->node_wheel[tyre_node[q]];
            const int stick = wk >= 0 ? vlen(need_sum[wk]) <= cap_sum[wk] : vlen(tyre[q].need) <= tyre[q].cap;
            B->force[tyre_node[q]] = vadd(B->force[tyre_node[q]], stick ? tyre[q].need : tyre[q].slide);
        }
        free(need_sum), free(cap_sum);
    }
    free(tyre_node), free(tyre);
    return contacts;
}

/* SoftBody::integrate_nodes (src/phys/softbody.cpp:769-795) */
static void integrate(orc_world* w, orc_body* B, float dt, v3* mn_out, v3* mx_out, float* max_v2_out) {
    float max_v2 = *max_v2_out;
    v3 mn = *mn_out, mx = *mx_out;
    const float kMax = 400.0f;
    for (uint32_t i = 0; i < B->n; i++) {
        bl_node* x = &B->nodes[i];
        if (x->inv_mass > 0) {
            v3 v = vadd(ld3(x->v), vmul(B->force[i], x->inv_mass * dt));
            float v2 = vdot(v, v);
            if (v2 > kMax * kMax) {
                v = vmul(v, kMax / sqrtf(v2));
                v2 = kMax * kMax;
                w->st.vel_clamps++;
            } else if (!(v2 >= 0.0f)) {
                v = V(0, 0, 0);
                v2 = 0;
                w->st.nan_resets++;
            }
            st3(x->v, v);
            st3(x->p, vadd(ld3(x->p), vmul(v, dt)));
            max_v2 = v2 > max_v2 ? v2 : max_v2;
        }
        mn = vmin3(mn, ld3(x->p));
        mx = vmax3(mx, ld3(x->p));
    }
    *mn_out = mn, *mx_out = mx, *max_v2_out = max_v2;
}

/* The substep loop of World::simulate_island for bodies without inter-body pairs (src/phys/world.cpp:3326-3591):
 * forces (clear_forces src/phys/softbody.cpp:330-381 with no drag: f = g m; then the beams), then per node the static
 * contact and the integration. */
static int cmp_float(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* World::repair_nodes (src/phys/world.cpp:4022-4111), beams only: returns the nodes reset, -1 when more than half
 * of the body is bad. The median is the element at size/2 of the sorted finite coordinates, as nth_element finds. */
static int repair_nodes(orc_body* B) {
    const uint32_t n = B->n;
    float* xs = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    float *ys = xs + n, *zs = xs + 2 * n;
    uint32_t m = 0;
    for (uint32_t i = 0; i < n; i++) {
        const float* p = B->nodes[i].p;
        if (isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2])) xs[m] = p[0], ys[m] = p[1], zs[m] = p[2], m++;
    }
    if (m == 0) { free(xs); return -1; }
    qsort(xs, m, sizeof(float), cmp_float), qsort(ys, m, sizeof(float), cmp_float), qsort(zs, m, sizeof(float), cmp_float);
    const v3 med = V(xs[m / 2], ys[m / 2], zs[m / 2]);
    free(xs);
    const float reach = 300.0f, vmax = 0.9f * 400.0f;
    uint8_t* bad = (uint8_t*)calloc(n ? n : 1, 1);
    int nbad = 0;
    for (uint32_t i = 0; i < n; i++) {
        const bl_node* x = &B->nodes[i];
        int finite = isfinite(x->p[0]) && isfinite(x->p[1]) && isfinite(x->p[2]) && isfinite(x->v[0]) && isfinite(x->v[1]) && isfinite(x->v[2]);
        v3 v = ld3(x->v);
        if (!finite || vdot(v, v) > vmax * vmax || vlen(vsub(ld3(x->p), med)) > reach) bad[i] = 1, nbad++;
    }
    if (nbad == 0) { free(bad); return 0; }
    if (nbad * 2 > (int)n) { free(bad); return -1; }
    const int reset = nbad;
    v3* sp = (v3*)malloc(sizeof(v3) * n);
    v3* sv = (v3*)malloc(sizeof(v3) * n);
    int* cnt = (int*)malloc(sizeof(int) * n);
    for (int pass = 0; pass < 8 && nbad > 0; pass++) {
        for (uint32_t i = 0; i < n; i++) sp[i] = V(0, 0, 0), sv[i] = V(0, 0, 0), cnt[i] = 0;
        for (uint32_t j = 0; j < B->nb; j++) {
            const bl_beam* bm = &B->beams[j];
            if (bm->flags & BL_BF_BROKEN) continue;
            uint32_t x = bm->a, y = bm->b;
            if (bad[x] && !bad[y]) {
                sp[x] = vadd(sp[x], ld3(B->nodes[y].p)), sv[x] = vadd(sv[x], ld3(B->nodes[y].v)), cnt[x]++;
            } else if (bad[y] && !bad[x]) {
                sp[y] = vadd(sp[y], ld3(B->nodes[x].p)), sv[y] = vadd(sv[y], ld3(B->nodes[x].v)), cnt[y]++;
            }
        }
        for (uint32_t i = 0; i < n; i++)
            if (bad[i] && cnt[i] > 0) {
                st3(B->nodes[i].p, vdivs(sp[i], (float)cnt[i]));
                st3(B->nodes[i].v, vdivs(sv[i], (float)cnt[i]));
                bad[i] = 0;
                nbad--;
            }
    }
    for (uint32_t i = 0; i < n; i++)
        if (bad[i]) st3(B->nodes[i].p, med), st3(B->nodes[i].v, V(0, 0, 0));
    free(sp), free(sv), free(cnt), free(bad);
    for (uint32_t k = 0; k < B->nfr; k++) B->frames[k].w[0] = B->frames[k].w[1] = B->frames[k].w[2] = 0;
    /* bounds and top speed of the repaired body (src/phys/world.cpp:4106-4109) */
    B->mn = V(1e30f, 1e30f, 1e30f), B->mx = V(-1e30f, -1e30f, -1e30f);
    B->max_speed = 0;
    for (uint32_t i = 0; i < n; i++) {
        B->mn = vmin3(B->mn, ld3(B->nodes[i].p)), B->mx = vmax3(B->mx, ld3(B->nodes[i].p));
        B->max_speed = smax(B->max_speed, vlen(ld3(B->nodes[i].v)));
    }
    return reset;
}

/* The post-frame check of World::simulate_island (src/phys/world.cpp:3962-3979): a body whose bounds are not
 * finite, wider than 800 m, or whose top speed in the frame passed 360 m/s is repaired, or frozen when it cannot be. */
/* SoftBody::motion_energy (src/phys/softbody.cpp:798-803), the nodes' part */
static void motion_energy(const orc_body* B, v3 g, double* ke, double* pe) {
    *ke = *pe = 0;
    for (uint32_t i = 0; i < B->n; i++) {
        const bl_node* x = &B->nodes[i];
        if (x->inv_mass > 0) {
            const v3 v = ld3(x->v);
            *ke += 0.5 * x->mass * vdot(v, v);
            *pe -= x->mass * (double)vdot(g, ld3(x->p));
        }
    }
}

/* SoftBody::guard_energy (src/phys/softbody.cpp:805-815): a frame that gained energy beyond 2 J and 5 % of its
 * kinetic energy at the start is scaled back */
static void guard_energy(orc_body* B, v3 g) {
    double ke, pe;
    motion_energy(B, g, &ke, &pe);
    const double excess = (ke + pe) - (B->guard_ke + B->guard_pe) - (2.0 + 0.05 * B->guard_ke);
    if (!(excess > 0) || !(ke > 0)) return;
    const double left = ke - excess;
    const float s = (float)sqrt((0.0 < left ? left : 0.0) / ke);
    for (uint32_t i = 0; i < B->n; i++) {
        bl_node* x = &B->nodes[i];
        x->v[0] *= s, x->v[1] *= s, x->v[2] *= s;
    }
    B->guard_cuts++;
}

static void post_frame(orc_world* w, orc_body* B) {
    const float lim = 400.0f * 0.9f;
    v3 e = vsub(B->mx, B->mn);
    if (!isfinite(B->mn.x) || !isfinite(B->mx.y) || smax(e.x, smax(e.y, e.z)) > 800.0f || B->max_speed > lim) {
        int fixed = repair_nodes(B);
        if (fixed < 0) {
            for (uint32_t i = 0; i < B->n; i++) {
                bl_node* x = &B->nodes[i];
                if (!isfinite(x->p[0]) || !isfinite(x->p[1]) || !isfinite(x->p[2])) x->p[0] = 0, x->p[1] = -1000, x->p[2] = 0;
                x->v[0] = x->v[1] = x->v[2] = 0;
                x->inv_mass = 0;
            }
            B->mn = V(1e30f, 1e30f, 1e30f), B->mx = V(-1e30f, -1e30f, -1e30f);
            for (uint32_t i = 0; i < B->n; i++) B->mn = vmin3(B->mn, ld3(B->nodes[i].p)), B->mx = vmax3(B->mx, ld3(B->nodes[i].p));
            B->repairs += 1;
            return; /* (frozen: src/phys/world.cpp:3967-3976 skips the rest) */
        }
        B->repairs += 1;
    }
    /* (src/phys/world.cpp:3995-3996: no collision volumes or sphere contacts here, so every frame is a quiet one) */
    B->guard_quiet++;
    if (B->opts.energy_guard && B->guard_quiet > 15) guard_energy(B, ld3(w->s.gravity));
}

/* ------------------------------------------------------------------ inter-body contacts (slow bodies) */
typedef struct { uint16_t bn, bt; uint32_t node, tri; } orc_nt;
typedef struct { uint16_t bn, bc; uint32_t node, cap; } orc_nc;
typedef struct { uint16_t bc, bt; uint32_t cap, tri; } orc_ct;
typedef struct { uint32_t pair; float s; v3 bary, nrm; float pen; } orc_hit;
typedef struct { uint16_t body; uint32_t node; int x, y, z; v3 p; } orc_pt;
typedef struct { uint32_t key, idx; } orc_entry;

#define VEC_PUSH(arr, n, cap, val)                                                                                     \
    do {                                                                                                               \
        if ((n) == (cap)) {                                                                                            \
            (cap) = (cap) ? 2 * (cap) : 64;                                                                            \
            (arr) = realloc((arr), sizeof(*(arr)) * (cap));                                                            \
        }                                                                                                              \
        (arr)[(n)++] = (val);                                                                                          \
    } while (0)

typedef struct {
    int* bodies;  /* world indices, in body order */
    int nb;
    int needs_pairs;
    int interval, next_rebuild;
    float margin, inv_cell;
    orc_nt* nt; size_t nnt, cnt;
    orc_nc* nc; size_t nnc, cnc;
    orc_ct* ct; size_t nct, cct;
    orc_hit* hits; size_t nhit, chit;
} orc_island;

/* cell_hash (src/phys/world.cpp:23-25): int products wrap as uint32 */
static inline uint32_t cell_hash(int x, int y, int z) {
    return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}

static inline float max_edge2(v3 a, v3 b, v3 c) {
    v3 ba = vsub(b, a), cb = vsub(c, b), ac = vsub(a, c);
    return smax(vdot(ba, ba), smax(vdot(cb, cb), vdot(ac, ac)));
}
static inline v3 vcross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 vnormalize(v3 a) { float l = vlen(a); return l > 1e-12f ? vmul(a, 1.0f / l) : V(0, 0, 0); }
static inline float clampf01(float v) { return clampf_(v, 0, 1); }

/* closest_on_triangle (src/phys/softbody.h:158-190, Ericson 5.1.5) */
static v3 closest_on_triangle(v3 p, v3 a, v3 b, v3 c, v3* bary) {
    v3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
    float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { *bary = V(1, 0, 0); return a; }
    v3 bp = vsub(p, b);
    float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { *bary = V(0, 1, 0); return b; }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        float v = d1 / (d1 - d3);
        *bary = V(1 - v, v, 0);
        return vadd(a, vmul(ab, v));
    }
    v3 cp = vsub(p, c);
    float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { *bary = V(0, 0, 1); return c; }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        float ww = d2 / (d2 - d6);
        *bary = V(1 - ww, 0, ww);
        return vadd(a, vmul(ac, ww));
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        float ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *bary = V(0, 1 - ww, ww);
        return vadd(b, vmul(vsub(c, b), ww));
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom, ww = vc * denom;
    *bary = V(1 - v - ww, v, ww);
    return vadd(vadd(a, vmul(ab, v)), vmul(ac, ww));
}

/* closest_on_segment, closest_seg_seg (src/phys/world.cpp:27-47) */
static float closest_on_segment(v3 p, v3 a, v3 b) {
    v3 ab = vsub(b, a);
    float l2 = vdot(ab, ab);
    if (l2 < 1e-12f) return 0;
    return clampf01(vdot(vsub(p, a), ab) / l2);
}
static void closest_seg_seg(v3 p1, v3 q1, v3 p2, v3 q2, float* s, float* t) {
    v3 d1 = vsub(q1, p1), d2 = vsub(q2, p2), r = vsub(p1, p2);
    float a = vdot(d1, d1), e = vdot(d2, d2), f = vdot(d2, r);
    if (a <= 1e-12f && e <= 1e-12f) { *s = *t = 0; return; }
    if (a <= 1e-12f) { *s = 0; *t = clampf01(f / e); return; }
    float c = vdot(d1, r);
    if (e <= 1e-12f) { *t = 0; *s = clampf01(-c / a); return; }
    float b = vdot(d1, d2), denom = a * e - b * b;
    *s = denom != 0 ? clampf01((b * f - c * e) / denom) : 0;
    *t = (b * *s + f) / e;
    if (*t < 0) { *t = 0; *s = clampf01(-c / a); }
    else if (*t > 1) { *t = 1; *s = clampf01((b - c) / a); }
}

/* near_triangle (src/phys/world.cpp:1135-1142): a two-sided triangle within r of p; a hull (one-sided) triangle also
 * takes p up to `depth` behind it, over its face */
static int near_triangle(v3 p, v3 a, v3 b, v3 c, v3 tu, float r, int two_sided, float depth) {
    const float s = vdot(vsub(p, a), tu);
    v3 bary;
    if (two_sided) {
        if (!(fabsf(s) <= r)) return 0;
        v3 d = vsub(closest_on_triangle(p, a, b, c, &bary), p);
        return vdot(d, d) <= r * r;
    }
    if (s > r || s < -r - depth) return 0;
    const v3 q = vsub(p, vmul(tu, s));
    v3 d = vsub(closest_on_triangle(q, a, b, c, &bary), q);
    return vdot(d, d) <= r * r;
}

static inline int cell_of(float v, float inv_cell) { return (int)floorf(v * inv_cell); }

static int has_contacters(const orc_body* B) {
    for (uint32_t i = 0; i < B->n; i++)
        if (B->info[i].flags & BL_NF_CONTACTER) return 1;
    return 0;
}

typedef struct { v3 mn, mx; } aabb3;
static inline int aabb_overlaps(aabb3 a, aabb3 b) {
    return a.mn.x <= b.mx.x && a.mx.x >= b.mn.x && a.mn.y <= b.mx.y && a.mx.y >= b.mn.y && a.mn.z <= b.mx.z && a.mx.z >= b.mn.z;
}
static inline aabb3 aabb_expand(aabb3 a, float r) {
    a.mn = vsub(a.mn, V(r, r, r)), a.mx = vadd(a.mx, V(r, r, r));
    return a;
}

static int uf_find(int* parent, int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
}

/* World::build_islands (src/phys/world.cpp:636-762), for awake bodies of node/beam type: union-find over the
 * expanded bounds of bodies that can touch (contacters against triangles or capsules, capsules against triangles). */
static int build_islands(orc_world* w, float frame_time, orc_island** out) {
    const int nb = w->nbodies;
    int* parent = malloc(sizeof(int) * (nb ? nb : 1));
    aabb3* boxes = malloc(sizeof(aabb3) * (nb ? nb : 1));
    int* hc = malloc(sizeof(int) * (nb ? nb : 1));
    for (int i = 0; i < nb; i++) {
        orc_body* B = &w->bodies[i];
        parent[i] = i;
        aabb3 a = {B->mn, B->mx};
        float m = smin(B->max_speed, 200.0f) * frame_time * 2.0f + 0.3f;
        boxes[i] = aabb_expand(a, m + B->hull_reach);
        hc[i] = has_contacters(B);
    }
    if (w->s.inter_body_collisions)
        for (int i = 0; i < nb; i++)
            for (int j = i + 1; j < nb; j++) {
                if (!aabb_overlaps(boxes[i], boxes[j])) continue;
                orc_body *bi = &w->bodies[i], *bj = &w->bodies[j];
                if (bi->opts.collision_group != 0 && bi->opts.collision_group == bj->opts.collision_group) continue;
                int ti = bi->ntri > 0, tj = bj->ntri > 0, ci = bi->ncap > 0, cj = bj->ncap > 0;
                int can = (hc[i] && (tj || cj)) || (hc[j] && (ti || ci)) || (ci && tj) || (cj && ti);
                if (!can) continue;
                int a = uf_find(parent, i), b = uf_find(parent, j);
                if (a != b) parent[a] = b;
            }
    int* island_of = malloc(sizeof(int) * (nb ? nb : 1));
    for (int i = 0; i < nb; i++) island_of[i] = -1;
    orc_island* isl = calloc(nb ? nb : 1, sizeof(orc_island));
    int nisl = 0;
    for (int i = 0; i < nb; i++) {
        int r = uf_find(parent, i);
        if (island_of[r] < 0) island_of[r] = nisl++;
        orc_island* I = &isl[island_of[r]];
        I->bodies = realloc(I->bodies, sizeof(int) * (I->nb + 1));
        I->bodies[I->nb++] = i;
    }
    for (int k = 0; k < nisl; k++)
        isl[k].needs_pairs = w->s.inter_body_collisions && (isl[k].nb > 1 || w->bodies[isl[k].bodies[0]].opts.self_collision);
    free(parent), free(boxes), free(hc), free(island_of);
    *out = isl;
    return nisl;
}

static void free_island(orc_island* I) {
    free(I->bodies), free(I->nt), free(I->nc), free(I->ct), free(I->hits);
}

typedef struct {
    orc_world* w;
    orc_island* I;
    aabb3* boxes;
    int** partners;
    int* npart;
} bp_ctx;

static int near_partner(const bp_ctx* c, int bi, v3 mn, v3 mx) {
    orc_body* B = &c->w->bodies[c->I->bodies[bi]];
    if (B->opts.self_collision) return 1;
    for (int k = 0; k < c->npart[bi]; k++) {
        aabb3 q = c->boxes[c->partners[bi][k]];
        if (mn.x <= q.mx.x && mx.x >= q.mn.x && mn.y <= q.mx.y && mx.y >= q.mn.y && mn.z <= q.mx.z && mx.z >= q.mn.z) return 1;
    }
    return 0;
}

static int same_group(orc_world* w, orc_island* I, int a, int b) {
    int ga = w->bodies[I->bodies[a]].opts.collision_group, gb = w->bodies[I->bodies[b]].opts.collision_group;
    return ga != 0 && ga == gb;
}

/* build_buckets of rebuild_pairs (src/phys/world.cpp:1985-1995): power-of-two buckets, stable counting fill */
static uint32_t build_buckets(const orc_entry* ent, size_t n, uint32_t** bstart, uint32_t** bitems) {
    uint32_t nbk = 64;
    while (nbk < n * 2) nbk <<= 1;
    uint32_t mask = nbk - 1;
    uint32_t* st = calloc(nbk + 1, sizeof(uint32_t));
    for (size_t e = 0; e < n; e++) st[(ent[e].key & mask) + 1]++;
    for (uint32_t i = 0; i < nbk; i++) st[i + 1] += st[i];
    uint32_t* items = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* fill = malloc(sizeof(uint32_t) * nbk);
    memcpy(fill, st, sizeof(uint32_t) * nbk);
    for (size_t e = 0; e < n; e++) items[fill[ent[e].key & mask]++] = ent[e].idx;
    free(fill);
    *bstart = st, *bitems = items;
    return mask;
}

static int cmp_ct(const void* x, const void* y) {
    const orc_ct *a = x, *b = y;
    if (a->bt != b->bt) return a->bt < b->bt ? -1 : 1;
    return a->tri < b->tri ? -1 : a->tri > b->tri ? 1 : 0;
}

static void fast_pairs(orc_world* w, orc_island* I, const bp_ctx* c, const float* extra, const float* bulk,
                       float slow_speed);

/* World::rebuild_pairs (src/phys/world.cpp:1851-2308) without FEM mid points: the slow bodies' node hash, the
 * triangles' and capsules' queries of it (margins grown by a fast body's own travel, a hull triangle's depth behind
 * it), the fast bodies' own pairs after them, then capsules against triangles */
static void rebuild_pairs(orc_world* w, orc_island* I) {
    w->st.pair_rebuilds++;
    I->nnt = I->nnc = I->nct = 0;
    const int nb = I->nb;
    float rmax = 0;
    int any_caps = 0, any_tris = 0;
    for (int k = 0; k < nb; k++) {
        orc_body* B = &w->bodies[I->bodies[k]];
        rmax = smax(rmax, B->opts.collision_radius);
        any_caps |= B->ncap > 0;
        any_tris |= B->ntri > 0;
        for (uint32_t t = 0; t < B->ntri; t++) {
            bl_triangle* tr = &B->tris[t];
            if (!tr->torn && tr->rest_edge2 > 0 &&
                max_edge2(ld3(B->nodes[tr->a].p), ld3(B->nodes[tr->b].p), ld3(B->nodes[tr->c].p)) > 2.5f * 2.5f * tr->rest_edge2)
                tr->torn = 1;
        }
    }
    float slow_speed = 0;
    for (int k = 0; k < nb; k++) {
        orc_body* B = &w->bodies[I->bodies[k]];
        if (B->max_speed <= 8.0f) slow_speed = smax(slow_speed, B->max_speed);
    }
    const float speed = slow_speed + 0.5f;
    const double hz = (double)smax(1.0f, w->s.collision_hz);
    int interval = (int)ceil(1.0 / (hz * (double)w->s.dt) - 1e-6);
    if (interval < 1) interval = 1;
    I->interval = interval;
    /* a fast body's bulk speed (its nodes' mass-weighted RMS speed, +25 %) and its extra margin (:1893-1915) */
    float* bulk = calloc(nb, sizeof(float));
    float* extra = calloc(nb, sizeof(float));
    for (int k = 0; k < nb; k++) {
        orc_body* B = &w->bodies[I->bodies[k]];
        if (B->max_speed <= 8.0f) continue;
        double mv2 = 0, m = 0;
        for (uint32_t i = 0; i < B->n; i++) {
            const v3 v = ld3(B->nodes[i].v);
            mv2 += (double)B->nodes[i].mass * vdot(v, v);
            m += B->nodes[i].mass;
        }
        bulk[k] = clampf_(m > 0 ? 1.25f * (float)sqrt(mv2 / m) : B->max_speed, 8.0f, 400.0f);
    }
    const float margin = 2.0f * speed * (float)interval * w->s.dt + 0.02f;
    I->margin = margin;
    for (int k = 0; k < nb; k++)
        if (w->bodies[I->bodies[k]].max_speed > 8.0f) extra[k] = bulk[k] * (float)interval * w->s.dt;
    {
        double tri_size = 0;
        int samples = 0;
        for (int k = 0; k < nb; k++) {
            orc_body* B = &w->bodies[I->bodies[k]];
            for (uint32_t t = 0; t < B->ntri; t += 7) {
                bl_triangle* tr = &B->tris[t];
                tri_size += sqrtf(max_edge2(ld3(B->nodes[tr->a].p), ld3(B->nodes[tr->b].p), ld3(B->nodes[tr->c].p)));
                samples++;
            }
        }
        const float mean = samples ? (float)(tri_size / samples) : 1.0f;
        I->inv_cell = 1.0f / clampf_(mean + 2.0f * (rmax + margin), 0.25f, 1.0f);
    }
    const float inv_cell = I->inv_cell;
    const float reach = rmax + margin;
    bp_ctx c;
    c.w = w, c.I = I;
    c.boxes = malloc(sizeof(aabb3) * nb);
    c.partners = calloc(nb, sizeof(int*));
    c.npart = calloc(nb, sizeof(int));
    for (int k = 0; k < nb; k++) {
        orc_body* B = &w->bodies[I->bodies[k]];
        aabb3 a = {B->mn, B->mx};
        c.boxes[k] = aabb_expand(a, reach + extra[k] + B->hull_reach);
        c.partners[k] = malloc(sizeof(int) * nb);
    }
    for (int a = 0; a < nb; a++)
        for (int b = a + 1; b < nb; b++) {
            if (same_group(w, I, a, b) || !aabb_overlaps(c.boxes[a], c.boxes[b])) continue;
            c.partners[a][c.npart[a]++] = b;
            c.partners[b][c.npart[b]++] = a;
        }
    /* 1) the spatial hash of contacter nodes */
    orc_pt* pts = NULL;
    size_t npts = 0, cpts = 0;
    orc_entry* ent = NULL;
    size_t nent = 0, cent = 0;
    for (int bi = 0; bi < nb; bi++) {
        orc_body* B = &w->bodies[I->bodies[bi]];
        if ((c.npart[bi] == 0 && !B->opts.self_collision) || extra[bi] > 0) continue; /* (fast bodies: fast_pairs) */
        for (uint32_t ni = 0; ni < B->n; ni++) {
            if (!(B->info[ni].flags & BL_NF_CONTACTER)) continue;
            v3 p = ld3(B->nodes[ni].p);
            if (!near_partner(&c, bi, p, p)) continue;
            orc_pt pt = {(uint16_t)bi, ni, cell_of(p.x, inv_cell), cell_of(p.y, inv_cell), cell_of(p.z, inv_cell), p};
            orc_entry e = {cell_hash(pt.x, pt.y, pt.z), (uint32_t)npts};
            VEC_PUSH(ent, nent, cent, e);
            VEC_PUSH(pts, npts, cpts, pt);
        }
    }
    w->last_rb[0] += (int64_t)npts;
    if (npts && (any_tris || any_caps)) {
        uint32_t *bstart, *bitems;
        const uint32_t mask = build_buckets(ent, nent, &bstart, &bitems);
        uint32_t* run_end = malloc(sizeof(uint32_t) * (npts ? npts : 1));
        for (uint32_t bk = 0; bk <= mask; bk++) {
            uint32_t e = bstart[bk + 1];
            for (uint32_t k = e; k-- > bstart[bk];) {
                int same = k + 1 < e && pts[bitems[k + 1]].body == pts[bitems[k]].body;
                run_end[k] = same ? run_end[k + 1] : k + 1;
            }
        }
        /* 2) triangles, then capsules, of each body query the node hash */
        for (int bt = 0; bt < nb; bt++) {
            orc_body* O = &w->bodies[I->bodies[bt]];
            if (c.npart[bt] == 0 && !O->opts.self_collision) continue;
            const float rq0 = rmax + margin + extra[bt];
            for (uint32_t t = 0; t < O->ntri; t++) {
                const bl_triangle* tr = &O->tris[t];
                if (tr->torn) continue;
                const float depth = tr->two_sided ? 0.0f : O->opts.hull_depth, rq = rq0 + depth;
                v3 a = ld3(O->nodes[tr->a].p), bb = ld3(O->nodes[tr->b].p), cc = ld3(O->nodes[tr->c].p);
                v3 tmn = vmin3(a, vmin3(bb, cc)), tmx = vmax3(a, vmax3(bb, cc));
                if (!near_partner(&c, bt, vsub(tmn, V(rq, rq, rq)), vadd(tmx, V(rq, rq, rq)))) continue;
                w->last_rb[1]++;
                const v3 tn = vcross(vsub(bb, a), vsub(cc, a));
                const float tn2 = vdot(tn, tn);
                const v3 tu = tn2 > 1e-20f ? vmul(tn, 1.0f / sqrtf(tn2)) : V(0, 0, 0);
                int x0 = cell_of(tmn.x - rq, inv_cell), x1 = cell_of(tmx.x + rq, inv_cell);
                int y0 = cell_of(tmn.y - rq, inv_cell), y1 = cell_of(tmx.y + rq, inv_cell);
                int z0 = cell_of(tmn.z - rq, inv_cell), z1 = cell_of(tmx.z + rq, inv_cell);
                if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) > 512) continue;
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            uint32_t bk = cell_hash(x, y, z) & mask;
                            for (uint32_t k = bstart[bk]; k < bstart[bk + 1]; k++) {
                                const orc_pt* pt = &pts[bitems[k]];
                                if (pt->body == bt && !O->opts.self_collision) { k = run_end[k] - 1; continue; }
                                if (pt->x != x || pt->y != y || pt->z != z) continue;
                                if (pt->p.x < tmn.x - rq || pt->p.y < tmn.y - rq || pt->p.z < tmn.z - rq || pt->p.x > tmx.x + rq ||
                                    pt->p.y > tmx.y + rq || pt->p.z > tmx.z + rq)
                                    continue;
                                const int bn = pt->body;
                                orc_body* Bn = &w->bodies[I->bodies[bn]];
                                int self = bn == bt;
                                if (self && (!Bn->opts.self_collision || tr->a == pt->node || tr->b == pt->node || tr->c == pt->node)) continue;
                                if (!self && same_group(w, I, bn, bt)) continue;
                                v3 p = ld3(Bn->nodes[pt->node].p);
                                float r = Bn->opts.collision_radius + margin + extra[bt];
                                const float rb = r + depth;
                                if (p.x < tmn.x - rb || p.y < tmn.y - rb || p.z < tmn.z - rb || p.x > tmx.x + rb || p.y > tmx.y + rb || p.z > tmx.z + rb) continue;
                                if (!near_triangle(p, a, bb, cc, tu, r, tr->two_sided, depth)) continue;
                                orc_nt q = {(uint16_t)bn, (uint16_t)bt, pt->node, t};
                                VEC_PUSH(I->nt, I->nnt, I->cnt, q);
                            }
                        }
            }
            for (uint32_t ci = 0; ci < O->ncap; ci++) {
                const bl_capsule* cp = &O->caps[ci];
                v3 a = ld3(O->nodes[cp->a].p), bb = ld3(O->nodes[cp->b].p);
                float rq2 = cp->radius + rmax + margin + extra[bt];
                v3 mn = vsub(vmin3(a, bb), V(rq2, rq2, rq2)), mx = vadd(vmax3(a, bb), V(rq2, rq2, rq2));
                if (!near_partner(&c, bt, mn, mx)) continue;
                int x0 = cell_of(mn.x, inv_cell), x1 = cell_of(mx.x, inv_cell), y0 = cell_of(mn.y, inv_cell), y1 = cell_of(mx.y, inv_cell);
                int z0 = cell_of(mn.z, inv_cell), z1 = cell_of(mx.z, inv_cell);
                if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) > 512) continue;
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            uint32_t bk = cell_hash(x, y, z) & mask;
                            for (uint32_t k = bstart[bk]; k < bstart[bk + 1]; k++) {
                                const orc_pt* pt = &pts[bitems[k]];
                                if (pt->body == bt) { k = run_end[k] - 1; continue; }
                                if (pt->x != x || pt->y != y || pt->z != z) continue;
                                const int bn = pt->body;
                                if (same_group(w, I, bn, bt)) continue;
                                orc_body* Bn = &w->bodies[I->bodies[bn]];
                                v3 p = ld3(Bn->nodes[pt->node].p);
                                float r = cp->radius + Bn->opts.collision_radius + margin + extra[bt];
                                float sg = closest_on_segment(p, a, bb);
                                v3 d = vsub(vadd(a, vmul(vsub(bb, a), sg)), p);
                                if (vdot(d, d) > r * r) continue;
                                orc_nc q = {(uint16_t)bn, (uint16_t)bt, pt->node, ci};
                                VEC_PUSH(I->nc, I->nnc, I->cnc, q);
                            }
                        }
            }
        }
        free(bstart), free(bitems), free(run_end);
    }
    /* 2b) the fast bodies' nodes against their partners' triangles and capsules */
    {
        int any_fast = 0;
        for (int k = 0; k < nb; k++) any_fast |= extra[k] > 0;
        if (any_fast) {
            const size_t n0 = I->nnt + I->nnc;
            fast_pairs(w, I, &c, extra, bulk, speed);
            w->fast_pairs += (int64_t)(I->nnt + I->nnc - n0);
        }
    }
    /* 3) capsules against triangles: the triangles hashed into every cell of their boxes */
    if (any_caps && any_tris) {
        typedef struct { uint16_t body; uint32_t idx; } prim;
        prim* prims = NULL;
        size_t np = 0, cpn = 0;
        nent = 0;
        for (int bi = 0; bi < nb; bi++) {
            orc_body* B = &w->bodies[I->bodies[bi]];
            if (c.npart[bi] == 0 && !B->opts.self_collision) continue;
            const float r = B->opts.collision_radius + margin + extra[bi];
            for (uint32_t t = 0; t < B->ntri; t++) {
                const bl_triangle* tr = &B->tris[t];
                if (tr->torn) continue;
                v3 a = ld3(B->nodes[tr->a].p), bb = ld3(B->nodes[tr->b].p), cc = ld3(B->nodes[tr->c].p);
                v3 mn = vsub(vmin3(a, vmin3(bb, cc)), V(r, r, r)), mx = vadd(vmax3(a, vmax3(bb, cc)), V(r, r, r));
                if (!near_partner(&c, bi, mn, mx)) continue;
                int x0 = cell_of(mn.x, inv_cell), x1 = cell_of(mx.x, inv_cell), y0 = cell_of(mn.y, inv_cell), y1 = cell_of(mx.y, inv_cell);
                int z0 = cell_of(mn.z, inv_cell), z1 = cell_of(mx.z, inv_cell);
                if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) > 512) continue;
                uint32_t pid = (uint32_t)np;
                prim pr = {(uint16_t)bi, t};
                VEC_PUSH(prims, np, cpn, pr);
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            orc_entry e = {cell_hash(x, y, z), pid};
                            VEC_PUSH(ent, nent, cent, e);
                        }
            }
        }
        uint32_t *bstart, *bitems;
        const uint32_t mask = build_buckets(ent, nent, &bstart, &bitems);
        for (int bc = 0; bc < nb; bc++) {
            orc_body* B = &w->bodies[I->bodies[bc]];
            for (uint32_t ci = 0; ci < B->ncap; ci++) {
                const bl_capsule* cp = &B->caps[ci];
                v3 a = ld3(B->nodes[cp->a].p), bb = ld3(B->nodes[cp->b].p);
                float rr = cp->radius + B->opts.collision_radius + margin + extra[bc];
                v3 mn = vsub(vmin3(a, bb), V(rr, rr, rr)), mx = vadd(vmax3(a, bb), V(rr, rr, rr));
                int x0 = cell_of(mn.x, inv_cell), x1 = cell_of(mx.x, inv_cell), y0 = cell_of(mn.y, inv_cell), y1 = cell_of(mx.y, inv_cell);
                int z0 = cell_of(mn.z, inv_cell), z1 = cell_of(mx.z, inv_cell);
                if ((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) > 512) continue;
                size_t start = I->nct;
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            uint32_t bk = cell_hash(x, y, z) & mask;
                            for (uint32_t k = bstart[bk]; k < bstart[bk + 1]; k++) {
                                const prim* pr = &prims[bitems[k]];
                                if (pr->body == bc || same_group(w, I, bc, pr->body)) continue;
                                orc_body* O = &w->bodies[I->bodies[pr->body]];
                                const bl_triangle* t = &O->tris[pr->idx];
                                v3 ta = ld3(O->nodes[t->a].p), tb = ld3(O->nodes[t->b].p), tc = ld3(O->nodes[t->c].p);
                                v3 tmn = vmin3(ta, vmin3(tb, tc)), tmx = vmax3(ta, vmax3(tb, tc));
                                if (tmx.x < mn.x || tmx.y < mn.y || tmx.z < mn.z || tmn.x > mx.x || tmn.y > mx.y || tmn.z > mx.z) continue;
                                orc_ct q = {(uint16_t)bc, pr->body, ci, pr->idx};
                                VEC_PUSH(I->ct, I->nct, I->cct, q);
                            }
                        }
                /* dedupe: a capsule sees a triangle through several cells (sorted by body, triangle) */
                size_t m = I->nct - start;
                if (m > 1) {
                    qsort(I->ct + start, m, sizeof(orc_ct), cmp_ct);
                    size_t wr = start + 1;
                    for (size_t k = start + 1; k < I->nct; k++)
                        if (I->ct[k].bt != I->ct[wr - 1].bt || I->ct[k].tri != I->ct[wr - 1].tri) I->ct[wr++] = I->ct[k];
                    I->nct = wr;
                }
            }
        }
        free(bstart), free(bitems), free(prims);
    }
    free(pts), free(ent);
    for (int k = 0; k < nb; k++) free(c.partners[k]);
    free(c.partners), free(c.npart), free(c.boxes);
    free(bulk), free(extra);
    w->last_rb[2] += (int64_t)I->nnt, w->last_rb[4] += (int64_t)I->nnc, w->last_rb[5] += (int64_t)I->nct;
    w->st.contact_pairs = (int64_t)(I->nnt + I->nnc + I->nct);
}

/* World::fast_pairs (src/phys/world.cpp:2310-2522): each fast body's contacter nodes in a grid of its own (cells about
 * the reach, at most 48 an axis; a node in every cell within its own travel until the next rebuild), each partner
 * triangle (then capsule) scanning the cells of its reach box: the nodes found there that are within reach, in
 * ascending order. (The fast pairs are refreshed with the rest, at the slow schedule's interval.) */
typedef struct {
    float inv_cell;
    v3 origin;
    int dim[3];
    uint32_t *start, *items;
} fast_grid;

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }
static inline float v3c(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

static uint32_t fg_cell(const fast_grid* G, v3 p) {
    const int x = clampi((int)((p.x - G->origin.x) * G->inv_cell), 0, G->dim[0] - 1);
    const int y = clampi((int)((p.y - G->origin.y) * G->inv_cell), 0, G->dim[1] - 1);
    const int z = clampi((int)((p.z - G->origin.z) * G->inv_cell), 0, G->dim[2] - 1);
    return (uint32_t)((z * G->dim[1] + y) * G->dim[0] + x);
}

/* the cells of node ni (the box of cells between its position -+ its travel), visited in z, y, x order */
static uint32_t fg_cells(const fast_grid* G, const orc_body* F, uint32_t ni, float travel, uint32_t* out) {
    const v3 p = ld3(F->nodes[ni].p);
    const float m = vlen(ld3(F->nodes[ni].v)) * travel;
    const uint32_t lo = fg_cell(G, vsub(p, V(m, m, m))), hi = fg_cell(G, vadd(p, V(m, m, m)));
    if (lo == hi) {
        out[0] = lo;
        return 1;
    }
    const int x0 = (int)(lo % G->dim[0]), y0 = (int)(lo / G->dim[0] % G->dim[1]), z0 = (int)(lo / (G->dim[0] * G->dim[1]));
    const int x1 = (int)(hi % G->dim[0]), y1 = (int)(hi / G->dim[0] % G->dim[1]), z1 = (int)(hi / (G->dim[0] * G->dim[1]));
    uint32_t n = 0;
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) out[n++] = (uint32_t)((z * G->dim[1] + y) * G->dim[0] + x);
    return n;
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

/* the nodes listed in the cells of [mn, mx], ascending, each once */
static size_t fg_scan(const fast_grid* G, v3 mn, v3 mx, uint32_t** buf, size_t* cap) {
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        const int a = (int)floorf((v3c(mn, k) - v3c(G->origin, k)) * G->inv_cell);
        const int b = (int)((v3c(mx, k) - v3c(G->origin, k)) * G->inv_cell);
        lo[k] = a < 0 ? 0 : a;
        hi[k] = b < G->dim[k] - 1 ? b : G->dim[k] - 1;
        if (lo[k] > hi[k]) return 0;
    }
    size_t n = 0;
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++) {
            const uint32_t row = (uint32_t)((z * G->dim[1] + y) * G->dim[0]);
            const uint32_t k0 = G->start[row + lo[0]], k1 = G->start[row + hi[0] + 1];
            for (uint32_t k = k0; k < k1; k++) VEC_PUSH(*buf, n, *cap, G->items[k]);
        }
    qsort(*buf, n, sizeof(uint32_t), cmp_u32);
    size_t u = 0;
    for (size_t i = 0; i < n; i++)
        if (i == 0 || (*buf)[i] != (*buf)[i - 1]) (*buf)[u++] = (*buf)[i];
    return u;
}

static void fast_pairs(orc_world* w, orc_island* I, const bp_ctx* c, const float* extra, const float* bulk,
                       float slow_speed) {
    const int nb = I->nb;
    const float dt = w->s.dt;
    const int fi = I->interval; /* fast_interval: the slow schedule's */
#define FAST_EXTRA(b) (extra[(b)] <= 0 ? 0.0f : bulk[(b)] * (float)fi * dt)
    const float travel = 1.25f * (float)fi * dt;
    const float slow_reach = 2.0f * slow_speed * (float)fi * dt + 0.01f;
    uint32_t* found = NULL;
    size_t fcap = 0;
    for (int bf = 0; bf < nb; bf++) {
        if (FAST_EXTRA(bf) <= 0) continue;
        orc_body* F = &w->bodies[I->bodies[bf]];
        float rmax = 0;
        for (int k = 0; k < c->npart[bf]; k++)
            rmax = smax(rmax, F->opts.collision_radius + slow_reach + FAST_EXTRA(c->partners[bf][k]));
        fast_grid G;
        const v3 ext = vadd(vsub(F->mx, F->mn), V(2e-3f, 2e-3f, 2e-3f));
        float cellsz = clampf_(2.0f * rmax, 0.15f, 2.0f);
        cellsz = smax(cellsz, smax(ext.x, smax(ext.y, ext.z)) / 48.0f);
        G.inv_cell = 1.0f / cellsz;
        G.origin = vsub(F->mn, V(1e-3f, 1e-3f, 1e-3f));
        for (int k = 0; k < 3; k++) {
            int d = (int)(v3c(ext, k) * G.inv_cell) + 1;
            G.dim[k] = d > 48 ? 48 : d;
            if (G.dim[k] < 1) G.dim[k] = 1;
        }
        const uint32_t ncell = (uint32_t)(G.dim[0] * G.dim[1] * G.dim[2]);
        G.start = calloc(ncell + 1, sizeof(uint32_t));
        uint32_t count = 0;
        uint32_t* big = malloc(sizeof(uint32_t) * ncell); /* (a node's cells: at most the whole grid) */
        for (uint32_t ni = 0; ni < F->n; ni++) {
            if (!(F->info[ni].flags & BL_NF_CONTACTER)) continue;
            const uint32_t nc = fg_cells(&G, F, ni, travel, big);
            for (uint32_t k = 0; k < nc; k++) G.start[big[k] + 1]++, count++;
        }
        for (uint32_t i = 0; i < ncell; i++) G.start[i + 1] += G.start[i];
        G.items = malloc(sizeof(uint32_t) * (count ? count : 1));
        uint32_t* fill = malloc(sizeof(uint32_t) * (ncell ? ncell : 1));
        memcpy(fill, G.start, sizeof(uint32_t) * ncell);
        for (uint32_t ni = 0; ni < F->n; ni++) {
            if (!(F->info[ni].flags & BL_NF_CONTACTER)) continue;
            const uint32_t nc = fg_cells(&G, F, ni, travel, big);
            for (uint32_t k = 0; k < nc; k++) G.items[fill[big[k]]++] = ni;
        }
        free(fill), free(big);
        for (int pk = 0; pk < c->npart[bf]; pk++) {
            const int bo = c->partners[bf][pk];
            orc_body* O = &w->bodies[I->bodies[bo]];
            const int o_fast = extra[bo] > 0;
            const float r0 = F->opts.collision_radius + (o_fast ? FAST_EXTRA(bo) + 0.01f : slow_reach);
            const float r = F->opts.collision_radius + slow_reach + travel * (F->max_speed + (o_fast ? O->max_speed : 0.0f)) + FAST_EXTRA(bo);
            const aabb3 fb = aabb_expand((aabb3){F->mn, F->mx}, r);
            const float r00 = F->opts.collision_radius + (o_fast ? 0.01f : slow_reach);
            for (uint32_t t = 0; t < O->ntri; t++) {
                const bl_triangle* tr = &O->tris[t];
                if (tr->torn) continue;
                const v3 a = ld3(O->nodes[tr->a].p), bb = ld3(O->nodes[tr->b].p), cc = ld3(O->nodes[tr->c].p);
                const v3 tmn = vmin3(a, vmin3(bb, cc)), tmx = vmax3(a, vmax3(bb, cc));
                if (tmx.x < fb.mn.x || tmx.y < fb.mn.y || tmx.z < fb.mn.z || tmn.x > fb.mx.x || tmn.y > fb.mx.y || tmn.z > fb.mx.z) continue;
                float rt = r00;
                if (o_fast) {
                    const v3 va = ld3(O->nodes[tr->a].v), vb = ld3(O->nodes[tr->b].v), vc = ld3(O->nodes[tr->c].v);
                    rt = r00 + travel * sqrtf(smax(vdot(va, va), smax(vdot(vb, vb), vdot(vc, vc))));
                }
                const float depth = tr->two_sided ? 0.0f : O->opts.hull_depth;
                const v3 tn = vcross(vsub(bb, a), vsub(cc, a));
                const float tn2 = vdot(tn, tn);
                const v3 tu = tn2 > 1e-20f ? vmul(tn, 1.0f / sqrtf(tn2)) : V(0, 0, 0);
                const float e = rt + depth;
                const size_t n = fg_scan(&G, vsub(tmn, V(e, e, e)), vadd(tmx, V(e, e, e)), &found, &fcap);
                for (size_t i = 0; i < n; i++) {
                    const uint32_t ni = found[i];
                    const v3 p = ld3(F->nodes[ni].p), v = ld3(F->nodes[ni].v);
                    const float rn = rt + vlen(v) * travel, rb = rn + depth;
                    if (p.x < tmn.x - rb || p.y < tmn.y - rb || p.z < tmn.z - rb || p.x > tmx.x + rb || p.y > tmx.y + rb || p.z > tmx.z + rb) continue;
                    if (!near_triangle(p, a, bb, cc, tu, rn, tr->two_sided, depth)) continue;
                    orc_nt q = {(uint16_t)bf, (uint16_t)bo, ni, t};
                    VEC_PUSH(I->nt, I->nnt, I->cnt, q);
                }
            }
            for (uint32_t ci = 0; ci < O->ncap; ci++) {
                const bl_capsule* cp = &O->caps[ci];
                const v3 a = ld3(O->nodes[cp->a].p), bb = ld3(O->nodes[cp->b].p);
                const float rc = cp->radius + r;
                v3 cmn = vsub(vmin3(a, bb), V(rc, rc, rc)), cmx = vadd(vmax3(a, bb), V(rc, rc, rc));
                if (cmx.x < F->mn.x || cmx.y < F->mn.y || cmx.z < F->mn.z || cmn.x > F->mx.x || cmn.y > F->mx.y || cmn.z > F->mx.z) continue;
                const float rc0 = cp->radius + r0;
                cmn = vsub(vmin3(a, bb), V(rc0, rc0, rc0)), cmx = vadd(vmax3(a, bb), V(rc0, rc0, rc0));
                const size_t n = fg_scan(&G, cmn, cmx, &found, &fcap);
                for (size_t i = 0; i < n; i++) {
                    const uint32_t ni = found[i];
                    const v3 p = ld3(F->nodes[ni].p), v = ld3(F->nodes[ni].v);
                    const float rn = rc0 + vlen(v) * travel;
                    const float sg = closest_on_segment(p, a, bb);
                    const v3 d = vsub(vadd(a, vmul(vsub(bb, a), sg)), p);
                    if (!(vdot(d, d) <= rn * rn)) continue;
                    orc_nc q = {(uint16_t)bf, (uint16_t)bo, ni, ci};
                    VEC_PUSH(I->nc, I->nnc, I->cnc, q);
                }
            }
        }
        free(G.start), free(G.items);
    }
    free(found);
#undef FAST_EXTRA
}

/* World::near_pairs (src/phys/world.cpp:2525-2596): the node-triangle candidates within reach of their relative
 * travel until the next rebuild, and not receding across the triangle's plane; compacted in order */
static void near_pairs(orc_world* w, orc_island* I) {
    const size_t n = I->nnt;
    if (n == 0) return;
    const float T = (float)(I->interval > 1 ? I->interval : 1) * w->s.dt;
    size_t wr = 0;
    for (size_t i = 0; i < n; i++) {
        const orc_nt pr = I->nt[i];
        orc_body* Bn = &w->bodies[I->bodies[pr.bn]];
        orc_body* Bt = &w->bodies[I->bodies[pr.bt]];
        const bl_node* x = &Bn->nodes[pr.node];
        const bl_triangle* t = &Bt->tris[pr.tri];
        if (t->torn) continue;
        const bl_node *a = &Bt->nodes[t->a], *b = &Bt->nodes[t->b], *c3 = &Bt->nodes[t->c];
        v3 bary;
        const v3 q = closest_on_triangle(ld3(x->p), ld3(a->p), ld3(b->p), ld3(c3->p), &bary);
        const float r0 = Bn->opts.collision_radius + (t->two_sided ? 0.0f : Bt->opts.hull_depth);
        const v3 da = vsub(ld3(x->v), ld3(a->v)), db = vsub(ld3(x->v), ld3(b->v)), dc = vsub(ld3(x->v), ld3(c3->v));
        const float vrel = sqrtf(smax(vdot(da, da), smax(vdot(db, db), vdot(dc, dc))));
        const float reach = r0 + 1.25f * vrel * T + 0.02f;
        v3 xq = vsub(ld3(x->p), q);
        if (vdot(xq, xq) >= reach * reach) continue;
        const v3 nn = vcross(vsub(ld3(b->p), ld3(a->p)), vsub(ld3(c3->p), ld3(a->p)));
        const float nl = vlen(nn);
        if (nl > 1e-12f) {
            const v3 n0 = vdivs(nn, nl);
            const float s0 = vdot(n0, vsub(ld3(x->p), ld3(a->p)));
            /* (a hull's triangle only from in front) */
            if (t->two_sided || s0 > Bn->opts.collision_radius) {
                const float sg = s0 >= 0 ? 1.0f : -1.0f;
                const float close = smax(0.0f, smax(-sg * vdot(n0, da), smax(-sg * vdot(n0, db), -sg * vdot(n0, dc))));
                if (fabsf(s0) - 1.25f * close * T - 0.02f >= Bn->opts.collision_radius) continue;
            }
        }
        I->nt[wr++] = pr;
    }
    I->nnt = wr;
    w->last_rb[3] += (int64_t)wr;
    w->st.contact_pairs = (int64_t)(I->nnt + I->nnc + I->nct);
}

/* World::pair_detect (src/phys/world.cpp:2807-2962) */
static void pair_detect(orc_world* w, orc_island* I) {
    I->nhit = 0;
    for (size_t i = 0; i < I->nnt; i++) {
        const orc_nt pr = I->nt[i];
        orc_body* Bn = &w->bodies[I->bodies[pr.bn]];
        orc_body* Bt = &w->bodies[I->bodies[pr.bt]];
        const bl_node* n = &Bn->nodes[pr.node];
        const bl_triangle* t = &Bt->tris[pr.tri];
        if (t->torn) continue;
        const bl_node *a = &Bt->nodes[t->a], *b = &Bt->nodes[t->b], *cc = &Bt->nodes[t->c];
        const float r = Bt->opts.faces_only && !t->two_sided ? Bt->opts.face_skin : Bn->opts.collision_radius;
        v3 np = ld3(n->p), ap = ld3(a->p), bp = ld3(b->p), cp3 = ld3(cc->p);
        v3 bary;
        if (!t->two_sided) {
            /* a hull triangle: over its face the node is pushed out along the face's normal, from as deep as the
             * hull's depth behind it; off the face, outside, the ordinary contact with its edges */
            const v3 fn = vnormalize(vcross(vsub(bp, ap), vsub(cp3, ap)));
            const float sf = vdot(vsub(np, ap), fn);
            if (sf >= r || sf < -Bt->opts.hull_depth) continue;
            const v3 q = vsub(np, vmul(fn, sf));
            const v3 cq = vsub(closest_on_triangle(q, ap, bp, cp3, &bary), q);
            if (vdot(cq, cq) <= 1e-6f) {
                orc_hit h = {(uint32_t)i, 0.0f, bary, fn, r - sf};
                VEC_PUSH(I->hits, I->nhit, I->chit, h);
                continue;
            }
            if (sf < 0 || Bt->opts.faces_only) continue;
        } else {
            const v3 lo = vsub(vmin3(ap, vmin3(bp, cp3)), V(r, r, r)), hi = vadd(vmax3(ap, vmax3(bp, cp3)), V(r, r, r));
            if (np.x < lo.x || np.y < lo.y || np.z < lo.z || np.x > hi.x || np.y > hi.y || np.z > hi.z) continue;
        }
        v3 cp = closest_on_triangle(np, ap, bp, cp3, &bary);
        v3 d = vsub(np, cp);
        float dist2 = vdot(d, d);
        if (dist2 >= r * r) continue;
        float dist = sqrtf(dist2);
        v3 nrm;
        if (dist > 1e-5f) nrm = vdivs(d, dist);
        else {
            nrm = vnormalize(vcross(vsub(bp, ap), vsub(cp3, ap)));
            v3 vt = vadd(vadd(vmul(ld3(a->v), bary.x), vmul(ld3(b->v), bary.y)), vmul(ld3(cc->v), bary.z));
            if (vdot(vsub(ld3(n->v), vt), nrm) > 0) nrm = V(-nrm.x, -nrm.y, -nrm.z);
        }
        orc_hit h = {(uint32_t)i, 0.0f, bary, nrm, r - dist};
        VEC_PUSH(I->hits, I->nhit, I->chit, h);
    }
    for (size_t i = 0; i < I->nnc; i++) {
        const orc_nc pr = I->nc[i];
        orc_body* Bn = &w->bodies[I->bodies[pr.bn]];
        orc_body* Bc = &w->bodies[I->bodies[pr.bc]];
        const bl_capsule* cap = &Bc->caps[pr.cap];
        v3 np = ld3(Bn->nodes[pr.node].p), ap = ld3(Bc->nodes[cap->a].p), bp = ld3(Bc->nodes[cap->b].p);
        float t = closest_on_segment(np, ap, bp);
        v3 cp = vadd(ap, vmul(vsub(bp, ap), t));
        v3 d = vsub(np, cp);
        float r = Bn->opts.collision_radius + cap->radius;
        float dist2 = vdot(d, d);
        if (dist2 >= r * r) continue;
        float dist = sqrtf(dist2);
        v3 nrm = dist > 1e-5f ? vdivs(d, dist) : V(0, 1, 0);
        orc_hit h = {(uint32_t)(I->nnt + i), t, V(0, 0, 0), nrm, r - dist};
        VEC_PUSH(I->hits, I->nhit, I->chit, h);
    }
    for (size_t i = 0; i < I->nct; i++) {
        const orc_ct pr = I->ct[i];
        orc_body* Bc = &w->bodies[I->bodies[pr.bc]];
        orc_body* Bt = &w->bodies[I->bodies[pr.bt]];
        const bl_capsule* cap = &Bc->caps[pr.cap];
        const v3 p0 = ld3(Bc->nodes[cap->a].p), p1 = ld3(Bc->nodes[cap->b].p);
        const bl_triangle* tri = &Bt->tris[pr.tri];
        const v3 a = ld3(Bt->nodes[tri->a].p), b = ld3(Bt->nodes[tri->b].p), cc = ld3(Bt->nodes[tri->c].p);
        float best = 1e30f, best_s = 0;
        v3 best_bary = V(0, 0, 0), best_seg = V(0, 0, 0), best_tri = V(0, 0, 0);
#define CONSIDER(S, SP, TP, BARY)                                                                                      \
    do {                                                                                                               \
        v3 dd_ = vsub((SP), (TP));                                                                                     \
        float d2_ = vdot(dd_, dd_);                                                                                    \
        if (d2_ < best) { best = d2_; best_s = (S); best_seg = (SP); best_tri = (TP); best_bary = (BARY); }            \
    } while (0)
        v3 bary;
        v3 q = closest_on_triangle(p0, a, b, cc, &bary);
        CONSIDER(0.0f, p0, q, bary);
        q = closest_on_triangle(p1, a, b, cc, &bary);
        CONSIDER(1.0f, p1, q, bary);
        const v3 tv[3] = {a, b, cc};
        for (int e = 0; e < 3; e++) {
            float sg, u;
            closest_seg_seg(p0, p1, tv[e], tv[(e + 1) % 3], &sg, &u);
            v3 sp = vadd(p0, vmul(vsub(p1, p0), sg));
            v3 tp = vadd(tv[e], vmul(vsub(tv[(e + 1) % 3], tv[e]), u));
            float bb3[3] = {0, 0, 0};
            bb3[e] = 1 - u;
            bb3[(e + 1) % 3] = u;
            CONSIDER(sg, sp, tp, V(bb3[0], bb3[1], bb3[2]));
        }
#undef CONSIDER
        v3 tn = vcross(vsub(b, a), vsub(cc, a));
        float d0 = vdot(vsub(p0, a), tn), d1 = vdot(vsub(p1, a), tn);
        int crossing = 0;
        if ((d0 < 0) != (d1 < 0) && fabsf(d0 - d1) > 1e-12f) {
            float sg = d0 / (d0 - d1);
            v3 ip = vadd(p0, vmul(vsub(p1, p0), sg));
            q = closest_on_triangle(ip, a, b, cc, &bary);
            v3 qi = vsub(q, ip);
            if (vdot(qi, qi) < 1e-8f) {
                crossing = 1;
                best = 0, best_s = sg, best_seg = ip, best_tri = q, best_bary = bary;
            }
        }
        float r = cap->radius + Bc->opts.collision_radius;
        if (best >= r * r) continue;
        float dist = sqrtf(best);
        v3 nrm;
        float pen;
        if (!crossing && dist > 1e-5f) {
            nrm = vdivs(vsub(best_seg, best_tri), dist);
            pen = r - dist;
        } else {
            nrm = vnormalize(tn);
            v3 mid = vmul(vadd(p0, p1), 0.5f);
            if (vdot(vsub(mid, a), nrm) < 0) nrm = V(-nrm.x, -nrm.y, -nrm.z);
            float dm = smin(fabsf(d0), fabsf(d1)) / smax(1e-6f, vlen(tn));
            pen = r + smin(dm, 0.2f);
        }
        orc_hit h = {(uint32_t)(I->nnt + I->nnc + i), best_s, best_bary, nrm, pen};
        VEC_PUSH(I->hits, I->nhit, I->chit, h);
    }
}

/* contact_force<NA, NB> (src/phys/world.cpp:59-91) with both sides movable and no force snapshots */
static int contact_force(orc_world* w, bl_node** A, v3** FA, const float* wa, int na, bl_node** B, v3** FB,
                         const float* wb, int nbn, v3 n, float pen, float dt, float fric) {
    float W = 0;
    for (int i = 0; i < na; i++) W += wa[i] * wa[i] * A[i]->inv_mass;
    for (int i = 0; i < nbn; i++) W += wb[i] * wb[i] * B[i]->inv_mass;
    if (W <= 1e-12f) return 0;
    float m = 1.0f / W;
    v3 va = V(0, 0, 0), vb = V(0, 0, 0), acc = V(0, 0, 0);
    for (int i = 0; i < na; i++) {
        va = vadd(va, vmul(ld3(A[i]->v), wa[i]));
        acc = vadd(acc, vmul(*FA[i], wa[i] * A[i]->inv_mass));
    }
    for (int i = 0; i < nbn; i++) {
        vb = vadd(vb, vmul(ld3(B[i]->v), wb[i]));
        acc = vsub(acc, vmul(*FB[i], wb[i] * B[i]->inv_mass));
    }
    const v3 f_rel = vmul(acc, m);
    v3 f = prim_coll(f_rel, vsub(va, vb), m, n, dt, &k_gm[BL_SURF_CONCRETE], pen, fric, 2.0f, 0.0f, -1.0f);
    for (int i = 0; i < na; i++) *FA[i] = vadd(*FA[i], vmul(f, wa[i]));
    for (int i = 0; i < nbn; i++) *FB[i] = vsub(*FB[i], vmul(f, wb[i]));
    (void)w;
    return 1;
}

/* hull_push_out<3> (src/phys/world.cpp:93-120), both sides movable, no force snapshots: the normal force that makes
 * the node's velocity relative to the face after the step the push-out speed (a fifth of the depth a step, at most
 * 2 m/s), counting the forces on both sides so far, only when it would come out slower */
static int hull_push_out(bl_node* A, v3* FA, bl_node** B, v3** FB, const float* wb, v3 n, float pen, float dt) {
    float W = A->inv_mass;
    for (int i = 0; i < 3; i++) W += wb[i] * wb[i] * B[i]->inv_mass;
    if (W <= 1e-12f) return 0;
    v3 vb = V(0, 0, 0), acc = V(0, 0, 0);
    acc = vadd(acc, vmul(*FA, A->inv_mass));
    for (int i = 0; i < 3; i++) {
        vb = vadd(vb, vmul(ld3(B[i]->v), wb[i]));
        acc = vsub(acc, vmul(*FB[i], wb[i] * B[i]->inv_mass));
    }
    const float m = 1.0f / W, vr = vdot(vsub(ld3(A->v), vb), n), target = smin(0.2f * pen / dt, 2.0f);
    const float fn = (target - vr) * m / dt - vdot(acc, n) * m;
    if (fn <= 0) return 1;
    const v3 f = vmul(n, fn);
    *FA = vadd(*FA, f);
    for (int i = 0; i < 3; i++) *FB[i] = vsub(*FB[i], vmul(f, wb[i]));
    return 1;
}

/* World::collide_pairs (src/phys/world.cpp:2964-3079): the narrow phase, then the response in pair order */
static void collide_pairs(orc_world* w, orc_island* I) {
    w->st.narrow_tests += (int64_t)(I->nnt + I->nnc + I->nct);
    if (I->nnt + I->nnc + I->nct == 0) return;
    pair_detect(w, I);
    const float dt = w->s.dt;
    for (size_t k = 0; k < I->nhit; k++) {
        const orc_hit* h = &I->hits[k];
        if (h->pair < I->nnt) {
            const orc_nt pr = I->nt[h->pair];
            orc_body *Bn = &w->bodies[I->bodies[pr.bn]], *Bt = &w->bodies[I->bodies[pr.bt]];
            const bl_triangle* t = &Bt->tris[pr.tri];
            bl_node* A[1] = {&Bn->nodes[pr.node]};
            v3* FA[1] = {&Bn->force[pr.node]};
            float wa[1] = {1.0f};
            bl_node* B[3] = {&Bt->nodes[t->a], &Bt->nodes[t->b], &Bt->nodes[t->c]};
            v3* FB[3] = {&Bt->force[t->a], &Bt->force[t->b], &Bt->force[t->c]};
            float wb[3] = {h->bary.x, h->bary.y, h->bary.z};
            float fric = (Bn->info[pr.node].flags & BL_NF_TYRE) ? Bn->info[pr.node].friction
                                                                 : 0.8f * smin(Bn->opts.contact_friction, Bt->opts.contact_friction);
            if (!t->two_sided && h->pen > Bn->opts.collision_radius) { /* (behind a hull's face) */
                if (hull_push_out(A[0], FA[0], B, FB, wb, h->nrm, h->pen, dt)) w->st.contacts++;
                continue;
            }
            if (contact_force(w, A, FA, wa, 1, B, FB, wb, 3, h->nrm, h->pen, dt, fric)) w->st.contacts++;
        } else if (h->pair < I->nnt + I->nnc) {
            const orc_nc pr = I->nc[h->pair - I->nnt];
            orc_body *Bn = &w->bodies[I->bodies[pr.bn]], *Bc = &w->bodies[I->bodies[pr.bc]];
            const bl_capsule* cap = &Bc->caps[pr.cap];
            bl_node* A[1] = {&Bn->nodes[pr.node]};
            v3* FA[1] = {&Bn->force[pr.node]};
            float wa[1] = {1.0f};
            bl_node* B[2] = {&Bc->nodes[cap->a], &Bc->nodes[cap->b]};
            v3* FB[2] = {&Bc->force[cap->a], &Bc->force[cap->b]};
            float wb[2] = {1 - h->s, h->s};
            if (contact_force(w, A, FA, wa, 1, B, FB, wb, 2, h->nrm, h->pen, dt,
                              0.8f * smin(Bn->opts.contact_friction, Bc->opts.contact_friction)))
                w->st.contacts++;
        } else {
            const orc_ct pr = I->ct[h->pair - I->nnt - I->nnc];
            orc_body *Bc = &w->bodies[I->bodies[pr.bc]], *Bt = &w->bodies[I->bodies[pr.bt]];
            const bl_capsule* cap = &Bc->caps[pr.cap];
            const bl_triangle* tri = &Bt->tris[pr.tri];
            bl_node* A[2] = {&Bc->nodes[cap->a], &Bc->nodes[cap->b]};
            v3* FA[2] = {&Bc->force[cap->a], &Bc->force[cap->b]};
            float wa[2] = {1 - h->s, h->s};
            bl_node* B[3] = {&Bt->nodes[tri->a], &Bt->nodes[tri->b], &Bt->nodes[tri->c]};
            v3* FB[3] = {&Bt->force[tri->a], &Bt->force[tri->b], &Bt->force[tri->c]};
            float wb[3] = {h->bary.x, h->bary.y, h->bary.z};
            if (contact_force(w, A, FA, wa, 2, B, FB, wb, 3, h->nrm, h->pen, dt,
                              0.8f * smin(Bc->opts.contact_friction, Bt->opts.contact_friction)))
                w->st.contacts++;
        }
    }
}

/* World::step_substeps (src/phys/world.cpp:368-585) -> build_islands, then simulate_island per island
 * (src/phys/world.cpp:3211-4020): per substep the forces (clear_forces src/phys/softbody.cpp:381 with no drag:
 * f = g m; the beams), the inter-body contacts (pairs rebuilt at the first substep and every `interval`
 * substeps), then per node the static contact and the integration; after the frame, the post-frame check. */
int orc_step(orc_world* w, int substeps) {
    const float dt = w->s.dt;
    const v3 g = ld3(w->s.gravity);
    if (substeps <= 0) return 0;
    orc_island* isl;
    /* (World::build_islands, src/phys/world.cpp:743-756: each body's reach this frame over the terrain) */
    for (int k = 0; k < w->nbodies && w->tnx; k++) {
        orc_body* B = &w->bodies[k];
        const float m = smin(B->max_speed, 200.0f) * ((float)substeps * dt) * 2.0f + 0.5f;
        B->terrain_max = terrain_max_height(w, B->mn.x - m, B->mn.z - m, B->mx.x + m, B->mx.z + m);
    }
    const int nisl = build_islands(w, (float)substeps * dt, &isl);
    for (int k = 0; k < nisl; k++) {
        orc_island* I = &isl[k];
        for (int q = 0; q < I->nb; q++) {
            orc_body* B = &w->bodies[I->bodies[q]];
            float m2 = 0;
            for (uint32_t i = 0; i < B->n; i++) {
                v3 v = ld3(B->nodes[i].v);
                m2 = smax(m2, vdot(v, v));
            }
            B->max_speed = sqrtf(m2);
            /* (the energy guard's start, src/phys/world.cpp:3236) */
            if (B->opts.energy_guard) motion_energy(B, g, &B->guard_ke, &B->guard_pe);
        }
        I->next_rebuild = 0;
        for (int s = 0; s < substeps; s++) {
            for (int q = 0; q < I->nb; q++) {
                orc_body* B = &w->bodies[I->bodies[q]];
                for (uint32_t i = 0; i < B->n; i++) B->force[i] = vmul(g, B->nodes[i].mass);
                beam_forces(B);
                /* (World::simulate_island's base_forces, src/phys/world.cpp:3255-3273: beams, shocks, ..., slides) */
                for (uint32_t k = 0; k < B->nfr; k++) B->torque[k] = V(0, 0, 0);
                if (B->nsh) shock_forces(B);
                if (B->nj) joint_forces(B);
                if (B->nw) {
                    if (B->controlled)
                        for (uint32_t k = 0; k < B->nw; k++) B->wheels[k].torque = B->ctl[2 * k], B->wheels[k].brake = B->ctl[2 * k + 1];
                    wheel_forces(B, dt);
                }
                if (B->nsl) slide_forces(B);
            }
            if (I->needs_pairs) {
                if (s == 0 || s >= I->next_rebuild) {
                    rebuild_pairs(w, I);
                    near_pairs(w, I);
                    I->next_rebuild = s + I->interval;
                }
                collide_pairs(w, I);
            }
            for (int q = 0; q < I->nb; q++) {
                orc_body* B = &w->bodies[I->bodies[q]];
                w->st.static_contacts += collide_static(w, B, dt);
                v3 mn = V(1e30f, 1e30f, 1e30f), mx = V(-1e30f, -1e30f, -1e30f);
                float v2 = 0;
                integrate(w, B, dt, &mn, &mx, &v2);
                B->mn = mn, B->mx = mx;
                float ms = sqrtf(v2);
                if (ms > B->max_speed) B->max_speed = ms;
            }
            /* (the frames after the nodes: src/phys/world.cpp:3929) */
            for (int q = 0; q < I->nb; q++) integrate_frames(&w->bodies[I->bodies[q]], dt);
        }
        free_island(I);
    }
    free(isl);
    w->st.substeps += substeps;
    for (int k = 0; k < w->nbodies; k++) post_frame(w, &w->bodies[k]);
    return 0;
}

int orc_read_body(orc_world* w, uint32_t id, bl_node* nodes, bl_beam* beams, int64_t* broken) {
    if ((int)id >= w->nbodies) return BL_ERR_ARG;
    orc_body* B = &w->bodies[id];
    if (nodes && B->n) memcpy(nodes, B->nodes, sizeof(bl_node) * B->n);
    if (beams && B->nb) memcpy(beams, B->beams, sizeof(bl_beam) * B->nb);
    if (broken) *broken = B->broken;
    return 0;
}

int orc_write_body(orc_world* w, uint32_t id, const bl_node* nodes, const bl_beam* beams) {
    if ((int)id >= w->nbodies) return BL_ERR_ARG;
    orc_body* B = &w->bodies[id];
    if (nodes && B->n) memcpy(B->nodes, nodes, sizeof(bl_node) * B->n);
    if (beams && B->nb) memcpy(B->beams, beams, sizeof(bl_beam) * B->nb);
    return 0;
}

int orc_read_wheels(orc_world* w, uint32_t id, bl_wheel* wheels, float* speed_spin) {
    if ((int)id >= w->nbodies) return BL_ERR_ARG;
    orc_body* B = &w->bodies[id];
    if (wheels && B->nw) memcpy(wheels, B->wheels, sizeof(bl_wheel) * B->nw);
    if (speed_spin) speed_spin[0] = B->wheel_speed, speed_spin[1] = B->wheel_spin;
    return 0;
}

int orc_set_wheel_controls(orc_world* w, uint32_t id, const float* torque, const float* brake) {
    if ((int)id >= w->nbodies) return BL_ERR_ARG;
    orc_body* B = &w->bodies[id];
    B->controlled = torque != NULL && brake != NULL;
    for (uint32_t k = 0; B->controlled && k < B->nw; k++) B->ctl[2 * k] = torque[k], B->ctl[2 * k + 1] = brake[k];
    return 0;
}

int orc_read_joints(orc_world* w, uint32_t id, bl_frame* frames, bl_joint* joints) {
    if ((int)id >= w->nbodies) return BL_ERR_ARG;
    orc_body* B = &w->bodies[id];
    if (frames && B->nfr) memcpy(frames, B->frames, sizeof(bl_frame) * B->nfr);
    if (joints && B->nj) memcpy(joints, B->joints, sizeof(bl_joint) * B->nj);
    return 0;
}

int orc_read_slides(orc_world* w, uint32_t id, bl_slide* out) {
    if ((int)id >= w->nbodies) return BL_ERR_ARG;
    orc_body* B = &w->bodies[id];
    if (out && B->nsl) memcpy(out, B->slides, sizeof(bl_slide) * B->nsl);
    return 0;
}

int64_t orc_fast_pairs(const orc_world* w) { return w->fast_pairs; }

void orc_rebuild_counts(const orc_world* w, int64_t* out) {
    for (int k = 0; k < 6; k++) out[k] = w->last_rb[k];
}

int orc_stats(orc_world* w, bl_stats* out) {
    *out = w->st;
    int64_t br = 0;
    for (int k = 0; k < w->nbodies; k++) br += w->bodies[k].broken;
    out->broken_beams = br;
    return 0;
}

/* SoftBody::finalize, beam part (src/phys/softbody.cpp:76-80) */
int orc_finalize(const bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb) {
    for (uint32_t j = 0; j < nb; j++) {
        if (beams[j].a >= n || beams[j].b >= n) return BL_ERR_ARG;
        float dd = vlen(vsub(ld3(nodes[beams[j].a].p), ld3(nodes[beams[j].b].p)));
        beams[j].L = beams[j].L0 = smax(1e-3f, dd);
    }
    return 0;
}

/* SoftBody::stabilize, beam part (src/phys/softbody.cpp:85-110) */
int orc_stabilize(const bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb, float dt, float kb, float db,
                  int* changed) {
    float* ksum = (float*)calloc(n ? n : 1, sizeof(float));
    float* dsum = (float*)calloc(n ? n : 1, sizeof(float));
    float* ks = (float*)malloc(sizeof(float) * (n ? n : 1));
    float* ds = (float*)malloc(sizeof(float) * (n ? n : 1));
    for (uint32_t j = 0; j < nb; j++) {
        ksum[beams[j].a] += beams[j].k;
        ksum[beams[j].b] += beams[j].k;
        dsum[beams[j].a] += beams[j].d;
        dsum[beams[j].b] += beams[j].d;
    }
    for (uint32_t i = 0; i < n; i++) {
        ks[i] = 1.0f, ds[i] = 1.0f;
        float m = nodes[i].mass;
        if (m <= 0) continue;
        float kl = ksum[i] * dt * dt / m;
        if (kl > kb) ks[i] = kb / kl;
        float dl = dsum[i] * dt / m;
        if (dl > db) ds[i] = db / dl;
    }
    int c = 0;
    for (uint32_t j = 0; j < nb; j++) {
        float a = smin(ks[beams[j].a], ks[beams[j].b]);
        float e = smin(ds[beams[j].a], ds[beams[j].b]);
        if (a < 1.0f || e < 1.0f) c++;
        beams[j].k *= a;
        beams[j].d *= e;
    }
    free(ksum), free(dsum), free(ks), free(ds);
    if (changed) *changed = c;
    return 0;
}

/* StaticWorld::add_box (src/phys/static_world.cpp:298-306): to_mat3 (src/core/math.h:163-170) and
 * StaticBox::update_aabb (src/phys/static_world.cpp:242-246) */
int orc_box_from_quat(const float* c, const float* half, const float* q, uint32_t surface, bl_box* out) {
    memset(out, 0, sizeof(*out));
    float x = q[0], y = q[1], z = q[2], qw = q[3];
    float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = qw * x, wy = qw * y, wz = qw * z;
    float m[9] = {1 - 2 * (yy + zz), 2 * (xy + wz), 2 * (xz - wy),
                  2 * (xy - wz), 1 - 2 * (xx + zz), 2 * (yz + wx),
                  2 * (xz + wy), 2 * (yz - wx), 1 - 2 * (xx + yy)};
    memcpy(out->rot, m, sizeof(m));
    for (int k = 0; k < 3; k++) out->center[k] = c[k], out->half[k] = half[k];
    v3 e = vadd(vadd(vmul(vabs3(ld3(m)), half[0]), vmul(vabs3(ld3(m + 3)), half[1])), vmul(vabs3(ld3(m + 6)), half[2]));
    st3(out->aabb_mn, vsub(ld3(c), e));
    st3(out->aabb_mx, vadd(ld3(c), e));
    out->surface = surface;
    out->max_force = 0;
    return 0;
}
