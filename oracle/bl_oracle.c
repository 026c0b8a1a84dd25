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
} orc_body;

struct orc_world {
    bl_settings s;
    bl_box* boxes;
    int nbox;
    orc_body* bodies;
    int nbodies;
    bl_stats st;
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
    }
    free(w->bodies);
    free(w->boxes);
    free(w);
}

int orc_set_statics(orc_world* w, const bl_box* boxes, int n) {
    free(w->boxes);
    w->boxes = (bl_box*)malloc(sizeof(bl_box) * (n > 0 ? n : 1));
    if (n > 0) memcpy(w->boxes, boxes, sizeof(bl_box) * n);
    w->nbox = n;
    return 0;
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

/* StaticWorld::collide_point, the box part (src/phys/static_world.cpp:364-374,407-454); boxes in index order */
static int collide_point(const orc_world* w, v3 p, float r, v3* normal, float* depth_out, int* surface,
                         float* max_force) {
    int hit = 0;
    float out_depth = 0;
    *max_force = 0;
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
    *depth_out = out_depth;
    return hit;
}

/* World::collide_static for one body's nodes, boxes only (src/phys/world.cpp:781-869) */
static int collide_static(orc_world* w, orc_body* B, float dt) {
    int contacts = 0;
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
        if (w->nbox == 0) continue;
        float r = in->radius;
        v3 n;
        float depth, max_force;
        int surf;
        if (!collide_point(w, ld3(x->p), r, &n, &depth, &surf, &max_force)) continue;
        contacts++;
        const ground_model* gm = &k_gm[surf < BL_SURF_COUNT ? surf : 0];
        if (max_force > 0) {
            v3 f = prim_coll(B->force[i], ld3(x->v), x->mass, n, dt, gm, depth, in->friction * fric_body,
                             B->opts.contact_push_max, B->opts.contact_slop, -1.0f);
            float fl = vlen(f);
            if (fl > max_force) f = vmul(f, max_force / fl);
            B->force[i] = vadd(B->force[i], f);
        } else {
            v3 f = prim_coll(B->force[i], ld3(x->v), x->mass, n, dt, gm, depth, in->friction * fric_body,
                             B->opts.contact_push_max, B->opts.contact_slop, B->opts.bounce);
            B->force[i] = vadd(B->force[i], f);
        }
    }
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
    return reset;
}

/* The post-frame check of World::simulate_island (src/phys/world.cpp:3962-3979): a body whose bounds are not
 * finite, wider than 800 m, or whose top speed in the frame passed 360 m/s is repaired, or frozen when it cannot be. */
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
        }
        B->repairs += 1;
    }
}

int orc_step(orc_world* w, int substeps) {
    const float dt = w->s.dt;
    const v3 g = ld3(w->s.gravity);
    /* the top speed at the island's start (src/phys/world.cpp:3217-3222) */
    for (int k = 0; k < w->nbodies; k++) {
        orc_body* B = &w->bodies[k];
        float m2 = 0;
        for (uint32_t i = 0; i < B->n; i++) {
            v3 v = ld3(B->nodes[i].v);
            m2 = smax(m2, vdot(v, v));
        }
        B->max_speed = sqrtf(m2);
    }
    for (int s = 0; s < substeps; s++) {
        for (int k = 0; k < w->nbodies; k++) {
            orc_body* B = &w->bodies[k];
            for (uint32_t i = 0; i < B->n; i++) B->force[i] = vmul(g, B->nodes[i].mass);
            beam_forces(B);
            w->st.static_contacts += collide_static(w, B, dt);
            v3 mn = V(1e30f, 1e30f, 1e30f), mx = V(-1e30f, -1e30f, -1e30f);
            float v2 = 0;
            integrate(w, B, dt, &mn, &mx, &v2);
            B->mn = mn, B->mx = mx;
            float ms = sqrtf(v2);
            if (ms > B->max_speed) B->max_speed = ms;
        }
        w->st.substeps++;
    }
    if (substeps > 0)
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
