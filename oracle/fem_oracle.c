/* Test infrastructure only: the checker of the HIP shell-FEM path. Only tests/, the smoke check and bench.py's
 * CPU-baseline leg load it; the product never does.
 *
 * A restatement in C of the reference's implicit step of a frame body of shell triangles and Timoshenko members
 * (zenkovich/beamlab src/phys/frame_fem.cpp, fem_shell.cpp, core/math.h), with plasticity but no topology events (the
 * body's allow_break off: no tears, splits or letting go), no mounts, contacts or held steps:
 *   ShellMesher::grid / tri                 fem_shell.cpp:17-46   (the plate, its masses)
 *   FemFrame::finalize (triangles)          frame_fem.cpp:259-276 (rest shape, rest frames)
 *   FemFrame::node_inertia                  frame_fem.cpp:3075-3086
 *   tri_grad, dkt_b, tri_local_k            frame_fem.cpp:1482-1579 (the element stiffness, FP64)
 *   tri_build_k                             frame_fem.cpp:1625-1647 (stored as floats)
 *   tri_kinematics, eval_tris               frame_fem.cpp:1584-1622, 1661-1831 (membrane flow, bending hinges)
 *   kinematics, stiffness, eval_forces      frame_fem.cpp:1384-1476, 1896-2070 (members: axial, torsion, hinges)
 *   effective, element_matrix, member_world_k  frame_fem.cpp:1834-1850, 2147-2186
 *   the unloading re-pass and its budget    frame_fem.cpp:2745-2760, 2959-2996, 1863-1864, 2591
 *   gather_part (triangles, no contacts)    frame_fem.cpp:916-1086 (6x6 blocks, FP64)
 *   chol6, right_solve_lt, sub_abt, sub_mv, sub_mtv, forward6, backward6   frame_fem.cpp:2193-2317
 *   the factorization and substitution loops   frame_fem.cpp:2899-2943 (here over the nodes' natural order, dense)
 *   FemFrame::begin_forces's resync of xd     frame_fem.cpp:1867-1872
 *   the update (dv as force, clamps, skips)   frame_fem.cpp:2904-2908, 3000-3027
 *   SoftBody::integrate_nodes                 softbody.cpp:769-795
 *   FemFrame::sync_positions                  frame_fem.cpp:3407-3422
 * The element stiffness is pinned bit for bit against the compiled reference's FemFrame::tri_stiffness, the whole
 * substep (plastic state included) against the compiled reference's FemFrame, and by closed forms (rigid-body modes,
 * plate theory, the frames' deflections, plastic collapse loads) and by residuals.
 */
#include "fem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>



/* ------------------------------------------------------------------ float vectors and quaternions (core/math.h) */
typedef struct { float x, y, z; } f3;
typedef struct { float x, y, z, w; } qf;
static inline f3 F3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add3(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub3(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul3(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b) { return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
static inline f3 div3(f3 a, float s) { float i = 1.0f / s; return F3(a.x * i, a.y * i, a.z * i); }
static inline f3 norm3(f3 a) { float l = len3(a); return l > 1e-12f ? mul3(a, 1.0f / l) : F3(0, 0, 0); }
static inline qf Q(float x, float y, float z, float w) { qf q = {x, y, z, w}; return q; }
static inline qf qmul(qf a, qf b) {
    return Q(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
static inline qf qconj(qf q) { return Q(-q.x, -q.y, -q.z, q.w); }
static inline qf qscale(qf q, float s) { return Q(q.x * s, q.y * s, q.z * s, q.w * s); }
static inline qf qnorm(qf q) {
    float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return l > 1e-12f ? qscale(q, 1.0f / l) : Q(0, 0, 0, 1);
}
/* from_mat3 with columns c0 c1 c2 (core/math.h:171-202) */
static qf from_cols(f3 c0, f3 c1, f3 c2) {
    float m00 = c0.x, m11 = c1.y, m22 = c2.z;
    float tr = m00 + m11 + m22;
    qf q;
    if (tr > 0) {
        float s = sqrtf(tr + 1.0f) * 2;
        q.w = 0.25f * s;
        q.x = (c1.z - c2.y) / s;
        q.y = (c2.x - c0.z) / s;
        q.z = (c0.y - c1.x) / s;
    } else if (m00 > m11 && m00 > m22) {
        float s = sqrtf(1.0f + m00 - m11 - m22) * 2;
        q.w = (c1.z - c2.y) / s;
        q.x = 0.25f * s;
        q.y = (c1.x + c0.y) / s;
        q.z = (c2.x + c0.z) / s;
    } else if (m11 > m22) {
        float s = sqrtf(1.0f + m11 - m00 - m22) * 2;
        q.w = (c2.x - c0.z) / s;
        q.x = (c1.x + c0.y) / s;
        q.y = 0.25f * s;
        q.z = (c2.y + c1.z) / s;
    } else {
        float s = sqrtf(1.0f + m22 - m00 - m11) * 2;
        q.w = (c0.y - c1.x) / s;
        q.x = (c2.x + c0.z) / s;
        q.y = (c2.y + c1.z) / s;
        q.z = 0.25f * s;
    }
    return qnorm(q);
}
/* quat_log / quat_exp (core/math.h:214-226) */
static f3 qlog(qf qin) {
    qf q = qin.w < 0 ? qscale(qin, -1.0f) : qin;
    f3 v = F3(q.x, q.y, q.z);
    float s = len3(v);
    if (s < 1e-8f) return mul3(v, 2.0f);
    float angle = 2.0f * atan2f(s, q.w);
    return mul3(v, angle / s);
}
static qf qexp(f3 rv) {
    float a = len3(rv);
    if (a < 1e-8f) return qnorm(Q(rv.x * 0.5f, rv.y * 0.5f, rv.z * 0.5f, 1.0f));
    f3 axis = norm3(div3(rv, a));
    float hh = a * 0.5f, s = sinf(hh);
    return Q(axis.x * s, axis.y * s, axis.z * s, cosf(hh));
}

/* ------------------------------------------------------------------ the element (frame_fem.cpp:1482-1579) */
static void dkt_b(const double x[3], const double y[3], double xi, double eta, double B[3][9]) {
    const double x23 = x[1] - x[2], x31 = x[2] - x[0], x12 = x[0] - x[1];
    const double y23 = y[1] - y[2], y31 = y[2] - y[0], y12 = y[0] - y[1];
    double P[7], q[7], r[7], t[7];
    const double xs[3] = {x23, x31, x12}, ys[3] = {y23, y31, y12};
    for (int k = 4; k <= 6; k++) {
        const double xij = xs[k - 4], yij = ys[k - 4], l2 = xij * xij + yij * yij;
        P[k] = -6 * xij / l2, q[k] = 3 * xij * yij / l2, t[k] = -6 * yij / l2, r[k] = 3 * yij * yij / l2;
    }
    const double a = 1 - 2 * xi, bb = 1 - 2 * eta;
    const double Hx_xi[9] = {P[6] * a + (P[5] - P[6]) * eta, q[6] * a - (q[5] + q[6]) * eta,
                             -4 + 6 * (xi + eta) + r[6] * a - eta * (r[5] + r[6]),
                             -P[6] * a + eta * (P[4] + P[6]), q[6] * a - eta * (q[6] - q[4]),
                             -2 + 6 * xi + r[6] * a + eta * (r[4] - r[6]),
                             -eta * (P[5] + P[4]), eta * (q[4] - q[5]), -eta * (r[5] - r[4])};
    const double Hy_xi[9] = {t[6] * a + eta * (t[5] - t[6]), 1 + r[6] * a - eta * (r[5] + r[6]),
                             -q[6] * a + eta * (q[5] + q[6]),
                             -t[6] * a + eta * (t[4] + t[6]), -1 + r[6] * a + eta * (r[4] - r[6]),
                             -q[6] * a - eta * (q[4] - q[6]),
                             -eta * (t[4] + t[5]), eta * (r[4] - r[5]), -eta * (q[4] - q[5])};
    const double Hx_eta[9] = {-P[5] * bb - xi * (P[6] - P[5]), q[5] * bb - xi * (q[5] + q[6]),
                              -4 + 6 * (xi + eta) + r[5] * bb - xi * (r[5] + r[6]),
                              xi * (P[4] + P[6]), xi * (q[4] - q[6]), -xi * (r[6] - r[4]),
                              P[5] * bb - xi * (P[4] + P[5]), q[5] * bb + xi * (q[4] - q[5]),
                              -2 + 6 * eta + r[5] * bb + xi * (r[4] - r[5])};
    const double Hy_eta[9] = {-t[5] * bb - xi * (t[6] - t[5]), 1 + r[5] * bb - xi * (r[5] + r[6]),
                              -q[5] * bb + xi * (q[5] + q[6]),
                              xi * (t[4] + t[6]), xi * (r[4] - r[6]), -xi * (q[4] - q[6]),
                              t[5] * bb - xi * (t[4] + t[5]), -1 + r[5] * bb + xi * (r[4] - r[5]),
                              -q[5] * bb - xi * (q[4] - q[5])};
    const double A2 = x31 * y12 - x12 * y31, iA = 1.0 / A2;
    for (int j = 0; j < 9; j++) {
        B[0][j] = (y31 * Hx_xi[j] + y12 * Hx_eta[j]) * iA;
        B[1][j] = (-x31 * Hy_xi[j] - x12 * Hy_eta[j]) * iA;
        B[2][j] = (-x31 * Hx_xi[j] - x12 * Hx_eta[j] + y31 * Hy_xi[j] + y12 * Hy_eta[j]) * iA;
    }
}
static const double kGauss[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
static const int kMemDof[3] = {0, 1, 5}, kPlateDof[3] = {2, 3, 4};

static void tri_grad(const float X[3][2], double b[3], double c[3], double* A2) {
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        b[i] = (double)X[j][1] - X[k][1];
        c[i] = (double)X[k][0] - X[j][0];
    }
    *A2 = ((double)X[1][0] - X[0][0]) * ((double)X[2][1] - X[0][1]) - ((double)X[2][0] - X[0][0]) * ((double)X[1][1] - X[0][1]);
}

void orc_fem_tri_local_k(const float Xf[6], float E, float nuf, float tf, float drill, double* Kout) {
    double K[18][18];
    const float X[3][2] = {{Xf[0], Xf[1]}, {Xf[2], Xf[3]}, {Xf[4], Xf[5]}};
    for (int i = 0; i < 18; i++)
        for (int j = 0; j < 18; j++) K[i][j] = 0;
    double b[3], c[3], A2;
    tri_grad(X, b, c, &A2);
    if (!(A2 > 0)) { memcpy(Kout, K, sizeof(K)); return; }
    const double A = 0.5 * A2, nu = nuf;
    const double Dm = (double)E * tf / (1 - nu * nu);
    double Bm[3][6];
    for (int i = 0; i < 3; i++) {
        Bm[0][2 * i] = b[i] / A2, Bm[1][2 * i] = 0, Bm[2][2 * i] = c[i] / A2;
        Bm[0][2 * i + 1] = 0, Bm[1][2 * i + 1] = c[i] / A2, Bm[2][2 * i + 1] = b[i] / A2;
    }
    const double D3[3][3] = {{1, nu, 0}, {nu, 1, 0}, {0, 0, 0.5 * (1 - nu)}};
    for (int p = 0; p < 6; p++)
        for (int q = 0; q < 6; q++) {
            double v = 0;
            for (int r = 0; r < 3; r++)
                for (int t = 0; t < 3; t++) v += Bm[r][p] * D3[r][t] * Bm[t][q];
            K[6 * (p / 2) + p % 2][6 * (q / 2) + q % 2] += A * Dm * v;
        }
    const double x[3] = {X[0][0], X[1][0], X[2][0]}, y[3] = {X[0][1], X[1][1], X[2][1]};
    const float Dbf = E * tf * tf * tf / (12.0f * (1.0f - nuf * nuf)); /* ShellSection::D() in float */
    const double Db = Dbf;
    for (int gq = 0; gq < 3; gq++) {
        double B[3][9];
        dkt_b(x, y, kGauss[gq][0], kGauss[gq][1], B);
        for (int p = 0; p < 9; p++)
            for (int q = 0; q < 9; q++) {
                double v = 0;
                for (int r = 0; r < 3; r++)
                    for (int t = 0; t < 3; t++) v += B[r][p] * D3[r][t] * B[t][q];
                K[6 * (p / 3) + 2 + p % 3][6 * (q / 3) + 2 + q % 3] += A2 / 6.0 * Db * v;
            }
    }
    const double G = (double)E / (2 * (1 + nu)), kd = (double)drill * G * tf * A;
    double gu[3], gv[3];
    for (int i = 0; i < 3; i++) gu[i] = -c[i] / (2 * A2), gv[i] = b[i] / (2 * A2);
    for (int i = 0; i < 3; i++) {
        K[6 * i + 5][6 * i + 5] += kd;
        for (int j = 0; j < 3; j++) {
            K[6 * i + 5][6 * j] -= kd * gu[j], K[6 * j][6 * i + 5] -= kd * gu[j];
            K[6 * i + 5][6 * j + 1] -= kd * gv[j], K[6 * j + 1][6 * i + 5] -= kd * gv[j];
            K[6 * i][6 * j] += 3 * kd * gu[i] * gu[j], K[6 * i][6 * j + 1] += 3 * kd * gu[i] * gv[j];
            K[6 * i + 1][6 * j] += 3 * kd * gv[i] * gu[j], K[6 * i + 1][6 * j + 1] += 3 * kd * gv[i] * gv[j];
        }
    }
    memcpy(Kout, K, sizeof(K));
}

/* ------------------------------------------------------------------ the plate */
/* FemFrame::Tangent (frame_fem.h:657-674): a member's tangent at the step's start */
typedef struct {
    f3 e1, e2, e3;
    float L, ka, kt, cy[3], cz[3], ngeo;
    int on;
    uint8_t yield, unload, hold; /* yield / unload bits: 1 hinge a, 2 hinge b, 4 axial, 8 torsion; elastic_hold */
    float ma[2], mb[2], N, T;
} mem_tan;

/* a member's plastic state (FrameElement, frame_fem.h:209-223): its order in orc_fem::mps */
enum { PS_UP, PS_TP, PS_PA, PS_PB = 4, PS_BA = 6, PS_BB = 8, PS_BT = 10, PS_DAMAGE, PS_DMG_A, PS_DMG_B, PS_UTIL, PS_N, PS_N_ };

struct orc_fem {
    uint32_t nn, nt;
    double* xd;
    float *p, *inv_mass, *force; /* the body's Node p and inv_mass; the forces the update hands the integrator */
    float* impulse;              /* gather_part's held impulse J (3 per node): f (-(h - step)), 0 for a step of h */
    float* torque;               /* FemFrame::torque: moments from outside for the next step (3 per node) */
    float *v, *w, *q, *mass, *inertia, *pos0;
    int64_t failures, clamps, nan_resets;
    uint8_t* fixed;
    uint32_t* tn;
    float *X, *r0, *Kf, *tmass, *area0;
    float E, nu, rho, t, drill, damping, yield;
    /* the triangles' plastic state: the authored rest shape, the corners' plastic rest rotations (element axes, not
     * yet baked), the plastic stretch reached, the utilisation; the bending's curvature rows at the three points
     * (tri_b_, 54 floats) and whether the stored stiffness is that of the current rest shape (tri_k_ok_) */
    float *X0, *th0, *tdmg, *tutil, *Bc;
    uint8_t* tk_ok;
    /* members (FrameElement): ends, section, joints; rest length, rest frames in the end nodes' frames, the elastic
     * stiffness (ka, kt, cy[3], cz[3]) and where the co-rotated frame takes its twist from; the plastic state and the
     * tangents */
    uint32_t nm, nsec;
    uint32_t* mem;               /* 4 per member: a, b, section, end_a | end_b << 8 */
    float *mL0, *mqa, *mqb, *mst, *mtw;
    uint8_t* mbroken;
    float* sec;                  /* ORC_FRAME_SECTION_FIELDS per section */
    float* mps;                  /* PS_N_ per member */
    mem_tan* tan;
    /* the components (FemFrame::analyse, :557-566: the frame's nodes joined by members and triangles) and their node
     * counts; the re-pass budget (:1863-1864, 2591) and the statistics */
    uint32_t *comp, *comp_size, ncomp;
    float repass_credit;
    int64_t repasses, passes;
    /* held steps (FemFrame::held_*, :2342-2512): the last step's prediction of the smooth forces and moments (pred_,
     * tpred_: all but the elements' and the contacts'), and its factorization kept (the envelope, its inverse
     * diagonals, the free nodes' indices) while it succeeded */
    float *pred, *tpred;
    int factored, nfk;
    double *Lk, *dk;
    int *fck, *fik;
    size_t* roffk;
};

int orc_fem_rest_shape(const float p[9], float X[6]) {
    const f3 x0 = F3(p[0], p[1], p[2]), x1 = F3(p[3], p[4], p[5]), x2 = F3(p[6], p[7], p[8]);
    const f3 nn = cross3(sub3(x1, x0), sub3(x2, x0));
    const float a2 = len3(nn);
    if (!(a2 > 1e-8f)) return -1;
    const f3 e3 = div3(nn, a2), e1 = norm3(sub3(x1, x0)), e2 = cross3(e3, e1);
    const f3 c = div3(add3(add3(x0, x1), x2), 3.0f);
    const f3 xs[3] = {x0, x1, x2};
    for (int i = 0; i < 3; i++) X[2 * i] = dot3(sub3(xs[i], c), e1), X[2 * i + 1] = dot3(sub3(xs[i], c), e2);
    return 0;
}

/* end_flex, condense, ends2, stiffness (frame_fem.cpp:1412-1476): the member's elastic stiffness in float */
static void mem_condense(float c[3], int pin_a, int pin_b) {
    if (pin_a && pin_b) c[0] = c[1] = c[2] = 0;
    else if (pin_a) c[2] = c[0] > 0 ? c[2] - c[1] * c[1] / c[0] : 0, c[0] = c[1] = 0;
    else if (pin_b) c[0] = c[2] > 0 ? c[0] - c[1] * c[1] / c[2] : 0, c[1] = c[2] = 0;
}
typedef struct { int rel[3]; float f; } end_flex_t;
static end_flex_t mem_end_flex(uint32_t j, const float* s) {
    end_flex_t e = {{0, 0, 0}, 0.0f};
    switch (j) {
    case 1: e.rel[0] = e.rel[1] = e.rel[2] = 1; break; /* FJ_BALL */
    case 2: e.rel[2] = 1; break;                       /* FJ_HINGE_V */
    case 3: e.rel[1] = 1; break;                       /* FJ_HINGE_H */
    case 4: e.rel[0] = 1; break;                       /* FJ_SWIVEL */
    case 5: e.f = s[ORC_SEC_JOINT_K] > 0 ? 1.0f / s[ORC_SEC_JOINT_K] : 0.0f; break; /* FJ_ELASTIC */
    default: break;
    }
    return e;
}
static void mem_ends2(float c[3], int ra, int rb, float fa, float fb) {
    if (ra || rb) {
        mem_condense(c, ra, rb);
        if (ra != rb) {
            float* k = ra ? &c[2] : &c[0];
            const float fl = ra ? fb : fa;
            if (fl > 0 && *k > 0) *k = 1.0f / (1.0f / *k + fl);
        }
        return;
    }
    if (fa <= 0 && fb <= 0) return;
    const double det = (double)c[0] * c[2] - (double)c[1] * c[1];
    if (!(det > 0)) return;
    const double F00 = c[2] / det + fa, F01 = -c[1] / det, F11 = c[0] / det + fb;
    const double d2 = F00 * F11 - F01 * F01;
    c[0] = (float)(F11 / d2), c[1] = (float)(-F01 / d2), c[2] = (float)(F00 / d2);
}
static void mem_bend(const float* s, float L, float I, float As, float c[3], int axis, end_flex_t fa, end_flex_t fb) {
    const float E = s[ORC_SEC_E], G = s[ORC_SEC_G];
    const float phi = As > 0 ? 12.0f * E * I / (G * As * L * L) : 0.0f;
    const float k0 = E * I / (L * (1.0f + phi));
    c[0] = c[2] = (4.0f + phi) * k0;
    c[1] = (2.0f - phi) * k0;
    mem_ends2(c, fa.rel[axis], fb.rel[axis], fa.f, fb.f);
}
static void mem_stiffness(const float* s, float L, uint32_t end_a, uint32_t end_b, float st[8]) {
    const end_flex_t fa = mem_end_flex(end_a, s), fb = mem_end_flex(end_b, s);
    st[0] = s[ORC_SEC_AXIAL] * s[ORC_SEC_E] * s[ORC_SEC_A] / L;
    st[1] = s[ORC_SEC_G] * s[ORC_SEC_J] / L;
    if (fa.rel[0] || fb.rel[0]) st[1] = 0;
    else if (fa.f + fb.f > 0) st[1] = 1.0f / (1.0f / st[1] + fa.f + fb.f);
    mem_bend(s, L, s[ORC_SEC_IY], s[ORC_SEC_ASZ], &st[2], 1, fa, fb);
    mem_bend(s, L, s[ORC_SEC_IZ], s[ORC_SEC_ASY], &st[5], 2, fa, fb);
}

/* A frame body (FemFrame with its nodes in the body's order): nodes p, mass (as the body has them), pinned flags;
 * shell triangles of one section (E, nu, rho, t, drill, damping); members (a, b, section, end_a | end_b << 8) of
 * elastic sections (ORC_FRAME_SECTION_FIELDS floats each). FemFrame::finalize (frame_fem.cpp:240-290) and
 * node_inertia (:3075-3086). */
/* tri_build_k (:1625-1647): the stored stiffness (float: the membrane's and the plate's 9 x 9) and the bending's
 * curvature rows at the three points from the corners' rotations alone, of the current rest shape */
static void tri_build_k(orc_fem* f, uint32_t k) {
    double K[18][18];
    const float* X = &f->X[6 * k];
    orc_fem_tri_local_k(X, f->E, f->nu, f->t, f->drill, &K[0][0]);
    float* P = &f->Kf[162 * (size_t)k];
    for (int a = 0; a < 9; a++)
        for (int c = 0; c < 9; c++) {
            const int im = 6 * (a / 3) + kMemDof[a % 3], jm = 6 * (c / 3) + kMemDof[c % 3];
            const int ip = 6 * (a / 3) + kPlateDof[a % 3], jp = 6 * (c / 3) + kPlateDof[c % 3];
            P[a * 9 + c] = (float)K[im < jm ? im : jm][im < jm ? jm : im];
            P[81 + a * 9 + c] = (float)K[ip < jp ? ip : jp][ip < jp ? jp : ip];
        }
    const double x[3] = {X[0], X[2], X[4]}, y[3] = {X[1], X[3], X[5]};
    float* Bc = &f->Bc[54 * (size_t)k];
    for (int g = 0; g < 3; g++) {
        double B[3][9];
        dkt_b(x, y, kGauss[g][0], kGauss[g][1], B);
        for (int r = 0; r < 3; r++)
            for (int i = 0; i < 3; i++) Bc[g * 18 + r * 6 + 2 * i] = (float)B[r][3 * i + 1], Bc[g * 18 + r * 6 + 2 * i + 1] = (float)B[r][3 * i + 2];
    }
    f->tk_ok[k] = 1;
}

static uint32_t uf_find(uint32_t* up, uint32_t x) {
    while (up[x] != x) x = up[x] = up[up[x]];
    return x;
}

orc_fem* orc_fem_create(uint32_t nn, const float* p, const float* mass, const uint8_t* fixed, uint32_t nt,
                        const uint32_t* tris, const float shell[ORC_SHELL_FIELDS], uint32_t nm,
                        const uint32_t* members, uint32_t nsec, const float* sections) {
    orc_fem* f = (orc_fem*)calloc(1, sizeof(orc_fem));
    f->E = shell[0], f->nu = shell[1], f->rho = shell[2], f->t = shell[3], f->drill = shell[4], f->damping = shell[5];
    f->yield = shell[6];
    f->X0 = (float*)calloc(6 * nt + 1, sizeof(float));
    f->th0 = (float*)calloc(9 * nt + 1, sizeof(float));
    f->tdmg = (float*)calloc(nt + 1, sizeof(float));
    f->tutil = (float*)calloc(nt + 1, sizeof(float));
    f->Bc = (float*)calloc(54 * (size_t)nt + 1, sizeof(float));
    f->tk_ok = (uint8_t*)calloc(nt + 1, 1);
    f->mps = (float*)calloc(PS_N_ * (size_t)nm + 1, sizeof(float));
    f->tan = (mem_tan*)calloc(nm + 1, sizeof(mem_tan));
    f->comp = (uint32_t*)calloc(nn + 1, sizeof(uint32_t));
    f->comp_size = (uint32_t*)calloc(nn + 1, sizeof(uint32_t));
    f->repass_credit = 8.0f; /* kRepassBurst */
    f->pred = (float*)calloc(3 * nn + 1, sizeof(float));
    f->tpred = (float*)calloc(3 * nn + 1, sizeof(float));
    f->nn = nn, f->nt = nt, f->nm = nm, f->nsec = nsec;
    f->xd = (double*)calloc(3 * nn + 1, sizeof(double));
    f->pos0 = (float*)calloc(3 * nn + 1, sizeof(float));
    f->p = (float*)calloc(3 * nn + 1, sizeof(float));
    f->inv_mass = (float*)calloc(nn + 1, sizeof(float));
    f->force = (float*)calloc(3 * nn + 1, sizeof(float));
    f->impulse = (float*)calloc(3 * nn + 1, sizeof(float));
    f->torque = (float*)calloc(3 * nn + 1, sizeof(float));
    f->v = (float*)calloc(3 * nn + 1, sizeof(float));
    f->w = (float*)calloc(3 * nn + 1, sizeof(float));
    f->q = (float*)calloc(4 * nn + 1, sizeof(float));
    f->mass = (float*)calloc(nn + 1, sizeof(float));
    f->inertia = (float*)calloc(nn + 1, sizeof(float));
    f->fixed = (uint8_t*)calloc(nn + 1, 1);
    f->tn = (uint32_t*)calloc(3 * nt + 1, sizeof(uint32_t));
    f->X = (float*)calloc(6 * nt + 1, sizeof(float));
    f->r0 = (float*)calloc(12 * nt + 1, sizeof(float));
    f->Kf = (float*)calloc(162 * (size_t)nt + 1, sizeof(float));
    f->tmass = (float*)calloc(nt + 1, sizeof(float));
    f->area0 = (float*)calloc(nt + 1, sizeof(float));
    f->mem = (uint32_t*)calloc(4 * nm + 1, sizeof(uint32_t));
    f->mL0 = (float*)calloc(nm + 1, sizeof(float));
    f->mqa = (float*)calloc(4 * nm + 1, sizeof(float));
    f->mqb = (float*)calloc(4 * nm + 1, sizeof(float));
    f->mst = (float*)calloc(8 * nm + 1, sizeof(float));
    f->mtw = (float*)calloc(nm + 1, sizeof(float));
    f->mbroken = (uint8_t*)calloc(nm + 1, 1);
    f->sec = (float*)calloc(ORC_FRAME_SECTION_FIELDS * nsec + 1, sizeof(float));
    memcpy(f->pos0, p, 3 * sizeof(float) * nn);
    memcpy(f->p, p, 3 * sizeof(float) * nn);
    for (uint32_t i = 0; i < 3 * nn; i++) f->xd[i] = p[i];
    memcpy(f->mass, mass, sizeof(float) * nn);
    if (fixed) memcpy(f->fixed, fixed, nn);
    for (uint32_t i = 0; i < nn; i++) {
        f->q[4 * i + 3] = 1.0f;
        f->inv_mass[i] = f->fixed[i] || f->mass[i] <= 0 ? 0.0f : 1.0f / f->mass[i];
    }
    if (nt) memcpy(f->tn, tris, 3 * sizeof(uint32_t) * nt);
    if (nm) memcpy(f->mem, members, 4 * sizeof(uint32_t) * nm);
    if (nsec) memcpy(f->sec, sections, ORC_FRAME_SECTION_FIELDS * sizeof(float) * nsec);
    /* the members' rest length and frames (x along the member, y towards the world's up; a vertical member's
     * towards x), their stiffness and twist weight (twist_weight, :1372-1375) */
    for (uint32_t k = 0; k < nm; k++) {
        const uint32_t a = f->mem[4 * k], b = f->mem[4 * k + 1], ea = f->mem[4 * k + 3] & 0xff, eb = f->mem[4 * k + 3] >> 8;
        const float* s = &f->sec[ORC_FRAME_SECTION_FIELDS * f->mem[4 * k + 2]];
        const f3 xa = F3(p[3 * a], p[3 * a + 1], p[3 * a + 2]), xb = F3(p[3 * b], p[3 * b + 1], p[3 * b + 2]);
        const f3 d = sub3(xb, xa);
        const float L0 = len3(d);
        f->mL0[k] = L0;
        if (L0 < 1e-5f || a == b) {
            f->mbroken[k] = 1;
            continue;
        }
        const f3 e1 = div3(d, L0);
        const f3 ref = fabsf(e1.y) < 0.9f ? F3(0, 1, 0) : F3(1, 0, 0);
        const f3 e2 = norm3(sub3(ref, mul3(e1, dot3(e1, ref)))), e3 = cross3(e1, e2);
        const qf E0 = from_cols(e1, e2, e3);
        const qf qa = qnorm(qmul(qconj(Q(0, 0, 0, 1)), E0)), qb = qnorm(qmul(qconj(Q(0, 0, 0, 1)), E0));
        f->mqa[4 * k] = qa.x, f->mqa[4 * k + 1] = qa.y, f->mqa[4 * k + 2] = qa.z, f->mqa[4 * k + 3] = qa.w;
        f->mqb[4 * k] = qb.x, f->mqb[4 * k + 1] = qb.y, f->mqb[4 * k + 2] = qb.z, f->mqb[4 * k + 3] = qb.w;
        mem_stiffness(s, L0, ea, eb, &f->mst[8 * k]);
        const int fa = ea == 1 || ea == 4, fb = eb == 1 || eb == 4;
        f->mtw[k] = fa == fb ? 0.5f : fa ? 1.0f : 0.0f;
    }
    /* the triangles: rest shapes, rest frames, stored stiffness (tri_build_k), masses (tri_mass) */
    const float mpa = f->rho * f->t; /* ShellSection::mass_per_m2 */
    for (uint32_t k = 0; k < nt; k++) {
        float pp[9];
        for (int c = 0; c < 3; c++) memcpy(&pp[3 * c], &f->pos0[3 * f->tn[3 * k + c]], 3 * sizeof(float));
        orc_fem_rest_shape(pp, &f->X[6 * k]);
        const f3 x0 = F3(pp[0], pp[1], pp[2]), x1 = F3(pp[3], pp[4], pp[5]), x2 = F3(pp[6], pp[7], pp[8]);
        const f3 nn3 = cross3(sub3(x1, x0), sub3(x2, x0));
        const float a2 = len3(nn3);
        const f3 e3 = div3(nn3, a2), e1 = norm3(sub3(x1, x0)), e2 = cross3(e3, e1);
        f->area0[k] = 0.5f * a2;
        f->tmass[k] = mpa * f->area0[k];
        const qf E0 = from_cols(e1, e2, e3);
        for (int c = 0; c < 3; c++) {
            const float* qn = &f->q[4 * f->tn[3 * k + c]];
            const qf r = qnorm(qmul(qconj(Q(qn[0], qn[1], qn[2], qn[3])), E0));
            f->r0[12 * k + 4 * c] = r.x, f->r0[12 * k + 4 * c + 1] = r.y, f->r0[12 * k + 4 * c + 2] = r.z, f->r0[12 * k + 4 * c + 3] = r.w;
        }
        memcpy(&f->X0[6 * k], &f->X[6 * k], 6 * sizeof(float));
        tri_build_k(f, k);
    }
    /* the components in the order of their first node (:557-566, 594-599) */
    {
        uint32_t* up = (uint32_t*)malloc(sizeof(uint32_t) * (nn + 1));
        for (uint32_t i = 0; i < nn; i++) up[i] = i;
        for (uint32_t k = 0; k < nm; k++) up[uf_find(up, f->mem[4 * k])] = uf_find(up, f->mem[4 * k + 1]);
        for (uint32_t k = 0; k < nt; k++) {
            const uint32_t* n = &f->tn[3 * k];
            up[uf_find(up, n[0])] = uf_find(up, n[1]), up[uf_find(up, n[1])] = uf_find(up, n[2]);
        }
        uint32_t* id = (uint32_t*)malloc(sizeof(uint32_t) * (nn + 1));
        for (uint32_t i = 0; i < nn; i++) id[i] = UINT32_MAX;
        for (uint32_t i = 0; i < nn; i++) {
            const uint32_t r = uf_find(up, i);
            if (id[r] == UINT32_MAX) id[r] = f->ncomp++;
            f->comp[i] = id[r];
            f->comp_size[id[r]]++;
        }
        free(up), free(id);
    }
    /* node_inertia: the members' halves (m L^2 / 78 per end, element order), then a corner's share of its triangles,
     * with the node's floor */
    for (uint32_t i = 0; i < nn; i++) f->inertia[i] = 0;
    for (uint32_t k = 0; k < nm; k++) {
        if (f->mbroken[k]) continue;
        const float* s = &f->sec[ORC_FRAME_SECTION_FIELDS * f->mem[4 * k + 2]];
        const float em = s[ORC_SEC_RHO] * s[ORC_SEC_A] * f->mL0[k]; /* FrameSection::mass_per_m() L0 */
        for (int c = 0; c < 2; c++) f->inertia[f->mem[4 * k + c]] += em * f->mL0[k] * f->mL0[k] / 78.0f;
    }
    for (uint32_t k = 0; k < nt; k++)
        for (int c = 0; c < 3; c++) f->inertia[f->tn[3 * k + c]] += f->tmass[k] * f->area0[k] / 108.0f;
    for (uint32_t i = 0; i < nn; i++) {
        const float fl = 1e-6f > f->mass[i] * 1e-4f ? 1e-6f : f->mass[i] * 1e-4f;
        f->inertia[i] = f->inertia[i] > fl ? f->inertia[i] : fl;
    }
    return f;
}

orc_fem* orc_fem_plate(int nu, int nv, float size, float y0, const float sec[6]) {
    const uint32_t nn = (uint32_t)((nu + 1) * (nv + 1)), nt = (uint32_t)(2 * nu * nv);
    float* p = (float*)calloc(3 * nn, sizeof(float));
    float* mass = (float*)calloc(nn, sizeof(float));
    uint8_t* fixed = (uint8_t*)calloc(nn, 1);
    uint32_t* tn = (uint32_t*)calloc(3 * nt, sizeof(uint32_t));
    /* ShellMesher::grid: origin + du i + dv j (float vector arithmetic) */
    const f3 origin = F3(-0.5f * size, y0, -0.5f * size), du = F3(size / (float)nu, 0, 0), dv = F3(0, 0, size / (float)nv);
    for (int j = 0; j <= nv; j++)
        for (int i = 0; i <= nu; i++) {
            const uint32_t id = (uint32_t)(j * (nu + 1) + i);
            const f3 q = add3(add3(origin, mul3(du, (float)i)), mul3(dv, (float)j));
            p[3 * id] = q.x, p[3 * id + 1] = q.y, p[3 * id + 2] = q.z;
            fixed[id] = (i == 0 || j == 0 || i == nu || j == nv);
        }
    uint32_t t = 0;
    const float mpa = sec[2] * sec[3]; /* ShellSection::mass_per_m2 */
    for (int j = 0; j < nv; j++)
        for (int i = 0; i < nu; i++) {
            const uint32_t n00 = (uint32_t)(j * (nu + 1) + i), n10 = n00 + 1, n01 = n00 + (uint32_t)(nu + 1), n11 = n01 + 1;
            uint32_t tri2[2][3];
            if ((i + j) & 1) {
                uint32_t a[2][3] = {{n00, n10, n11}, {n00, n11, n01}};
                memcpy(tri2, a, sizeof(a));
            } else {
                uint32_t a[2][3] = {{n00, n10, n01}, {n10, n11, n01}};
                memcpy(tri2, a, sizeof(a));
            }
            for (int k = 0; k < 2; k++, t++) {
                memcpy(&tn[3 * t], tri2[k], sizeof(tri2[k]));
                const float* pa = &p[3 * tri2[k][0]];
                const float* pb = &p[3 * tri2[k][1]];
                const float* pc = &p[3 * tri2[k][2]];
                const f3 A = F3(pa[0], pa[1], pa[2]), B = F3(pb[0], pb[1], pb[2]), C = F3(pc[0], pc[1], pc[2]);
                const float area = 0.5f * len3(cross3(sub3(B, A), sub3(C, A)));
                const float m = mpa * area; /* ShellMesher::tri: a third on each corner */
                for (int c = 0; c < 3; c++) mass[tri2[k][c]] += m / 3.0f;
            }
        }
    const float shell[ORC_SHELL_FIELDS] = {sec[0], sec[1], sec[2], sec[3], sec[4], sec[5], 0.0f};
    orc_fem* f = orc_fem_create(nn, p, mass, fixed, nt, tn, shell, 0, NULL, 0, NULL);
    free(p), free(mass), free(fixed), free(tn);
    return f;
}

void orc_fem_destroy(orc_fem* f) {
    if (!f) return;
    free(f->torque), free(f->mem), free(f->mL0), free(f->mqa), free(f->mqb), free(f->mst), free(f->mtw), free(f->mbroken);
    free(f->sec), free(f->X0), free(f->th0), free(f->tdmg), free(f->tutil), free(f->Bc), free(f->tk_ok), free(f->mps);
    free(f->tan), free(f->comp), free(f->comp_size), free(f->pred), free(f->tpred);
    free(f->Lk), free(f->dk), free(f->fck), free(f->fik), free(f->roffk);
    free(f->xd), free(f->pos0), free(f->p), free(f->inv_mass), free(f->force), free(f->impulse), free(f->v), free(f->w), free(f->q), free(f->mass), free(f->inertia), free(f->fixed);
    free(f->tn), free(f->X), free(f->r0), free(f->Kf), free(f->tmass), free(f->area0);
    free(f);
}
uint32_t orc_fem_nodes(const orc_fem* f) { return f->nn; }
uint32_t orc_fem_tris(const orc_fem* f) { return f->nt; }
void orc_fem_topology(const orc_fem* f, uint32_t* tri_nodes, double* xd, uint8_t* fixed, float* mass, float* inertia) {
    if (tri_nodes) memcpy(tri_nodes, f->tn, 3 * sizeof(uint32_t) * f->nt);
    if (xd) memcpy(xd, f->xd, 3 * sizeof(double) * f->nn);
    if (fixed) memcpy(fixed, f->fixed, f->nn);
    if (mass) memcpy(mass, f->mass, sizeof(float) * f->nn);
    if (inertia) memcpy(inertia, f->inertia, sizeof(float) * f->nn);
}
void orc_fem_get_state(const orc_fem* f, double* xd, float* v, float* w, float* q) {
    if (xd) memcpy(xd, f->xd, 3 * sizeof(double) * f->nn);
    if (v) memcpy(v, f->v, 3 * sizeof(float) * f->nn);
    if (w) memcpy(w, f->w, 3 * sizeof(float) * f->nn);
    if (q) memcpy(q, f->q, 4 * sizeof(float) * f->nn);
}
void orc_fem_set_state(orc_fem* f, const double* xd, const float* v, const float* w, const float* q) {
    if (xd) {
        memcpy(f->xd, xd, 3 * sizeof(double) * f->nn);
        for (uint32_t i = 0; i < 3 * f->nn; i++) f->p[i] = (float)xd[i]; /* the body's float positions follow */
    }
    if (v) memcpy(f->v, v, 3 * sizeof(float) * f->nn);
    if (w) memcpy(f->w, w, 3 * sizeof(float) * f->nn);
    if (q) memcpy(f->q, q, 4 * sizeof(float) * f->nn);
}

/* the nodes pinned (1) or free (0): ShellMesher::finish's inverse masses again (fem_shell.cpp:48-53) */
void orc_fem_set_fixed(orc_fem* f, const uint8_t* fixed) {
    for (uint32_t i = 0; i < f->nn; i++) {
        f->fixed[i] = fixed[i] ? 1 : 0;
        f->inv_mass[i] = f->fixed[i] || f->mass[i] <= 0 ? 0.0f : 1.0f / f->mass[i];
    }
}

/* the body's node records {p, inv_mass, v, mass} (as bl_node) and the frame's counters */
void orc_fem_get_nodes(const orc_fem* f, float* nodes, float* force, int64_t counters[3]) {
    if (nodes)
        for (uint32_t i = 0; i < f->nn; i++) {
            float* n = &nodes[8 * (size_t)i];
            n[0] = f->p[3 * i], n[1] = f->p[3 * i + 1], n[2] = f->p[3 * i + 2], n[3] = f->inv_mass[i];
            n[4] = f->v[3 * i], n[5] = f->v[3 * i + 1], n[6] = f->v[3 * i + 2], n[7] = f->mass[i];
        }
    if (force) memcpy(force, f->force, 3 * sizeof(float) * f->nn);
    if (counters) counters[0] = f->failures, counters[1] = f->clamps, counters[2] = f->nan_resets;
}

/* ------------------------------------------------------------------ the element pass (eval_tris, elastic) */
typedef struct {
    int on;
    f3 e1, e2, e3;
    f3 fo[3], to[3]; /* corner forces and moments (world) */
    float g[3][3];   /* membrane tension's geometric stiffness between corners */
    double kv[18];   /* K v in world axes */
} tri_out;

static int tri_kin(const orc_fem* f, uint32_t k, double d[18], f3* E1, f3* E2, f3* E3) {
    const uint32_t* n = &f->tn[3 * k];
    const float* X = &f->X[6 * k];
    double x[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) x[i][j] = f->xd[n[i] * 3 + j];
    const double a[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
    const double c[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
    double nv[3] = {a[1] * c[2] - a[2] * c[1], a[2] * c[0] - a[0] * c[2], a[0] * c[1] - a[1] * c[0]};
    const double nl = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]), al = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(nl > 1e-12) || !(al > 1e-9) || !isfinite(nl)) return 0;
    for (int j = 0; j < 3; j++) nv[j] /= nl;
    const double t1[3] = {a[0] / al, a[1] / al, a[2] / al};
    const double t2[3] = {nv[1] * t1[2] - nv[2] * t1[1], nv[2] * t1[0] - nv[0] * t1[2], nv[0] * t1[1] - nv[1] * t1[0]};
    double cen[3];
    for (int j = 0; j < 3; j++) cen[j] = (x[0][j] + x[1][j] + x[2][j]) / 3.0;
    double p[3][2];
    for (int i = 0; i < 3; i++) {
        const double r[3] = {x[i][0] - cen[0], x[i][1] - cen[1], x[i][2] - cen[2]};
        p[i][0] = r[0] * t1[0] + r[1] * t1[1] + r[2] * t1[2];
        p[i][1] = r[0] * t2[0] + r[1] * t2[1] + r[2] * t2[2];
    }
    double sn = 0, cs = 0;
    for (int i = 0; i < 3; i++) sn += X[2 * i] * p[i][1] - X[2 * i + 1] * p[i][0], cs += X[2 * i] * p[i][0] + X[2 * i + 1] * p[i][1];
    const double phi = atan2(sn, cs), cp = cos(phi), sp = sin(phi);
    double e1[3], e2[3];
    for (int j = 0; j < 3; j++) e1[j] = cp * t1[j] + sp * t2[j], e2[j] = -sp * t1[j] + cp * t2[j];
    *E1 = F3((float)e1[0], (float)e1[1], (float)e1[2]);
    *E2 = F3((float)e2[0], (float)e2[1], (float)e2[2]);
    *E3 = F3((float)nv[0], (float)nv[1], (float)nv[2]);
    const qf Ec = qconj(from_cols(*E1, *E2, *E3));
    for (int i = 0; i < 3; i++) {
        d[6 * i] = cp * p[i][0] + sp * p[i][1] - X[2 * i];
        d[6 * i + 1] = -sp * p[i][0] + cp * p[i][1] - X[2 * i + 1];
        d[6 * i + 2] = 0;
        const float* qn = &f->q[4 * n[i]];
        const float* r0 = &f->r0[12 * k + 4 * i];
        const float* t0 = &f->th0[9 * k + 3 * i];
        const f3 th = sub3(qlog(qmul(Ec, qnorm(qmul(Q(qn[0], qn[1], qn[2], qn[3]), Q(r0[0], r0[1], r0[2], r0[3]))))),
                           F3(t0[0], t0[1], t0[2]));
        d[6 * i + 3] = th.x, d[6 * i + 4] = th.y, d[6 * i + 5] = th.z;
    }
    return 1;
}

/* std::max / std::min as the reference calls them (the first argument unless the second compares greater / less) */
#define SMAX(a, b) ((a) < (b) ? (b) : (a))
#define SMIN(a, b) ((b) < (a) ? (b) : (a))

/* eval_tris's membrane strain and stress (Pa) of the corners' in-plane displacements */
static void tri_membrane(const double b3[3], const double c3[3], double A2, const double d[18], double nu, double Dm,
                         double eps[3], double sig[3]) {
    eps[0] = eps[1] = eps[2] = 0;
    for (int i = 0; i < 3; i++) {
        eps[0] += b3[i] * d[6 * i] / A2;
        eps[1] += c3[i] * d[6 * i + 1] / A2;
        eps[2] += (c3[i] * d[6 * i] + b3[i] * d[6 * i + 1]) / A2;
    }
    sig[0] = Dm * (eps[0] + nu * eps[1]), sig[1] = Dm * (nu * eps[0] + eps[1]), sig[2] = Dm * 0.5 * (1 - nu) * eps[2];
}

static double von_mises(const double m[3]) {
    const double v = m[0] * m[0] - m[0] * m[1] + m[1] * m[1] + 3 * m[2] * m[2];
    return sqrt(SMAX(0.0, v));
}

/* eval_tris's plasticity (:1690-1742, the body's allow_break off): the membrane flows radially back to the yield
 * stress, the rest shape following the stretch (the plastic stretch against the authored shape: the damage; the
 * stiffness rebuilt); the bending's moments at the three points against the plate's plastic moment, past it the
 * corners' plastic rest rotations following theirs; plastic rotations past 0.02 rad baked into the rest frames.
 * Returns the utilisation. */
static float tri_plastic(orc_fem* f, uint32_t k, double d[18], double b3[3], double c3[3], double* A2, double eps[3],
                         double sig[3]) {
    float* X = &f->X[6 * k];
    const double nu = f->nu, Dm = (double)f->E / (1 - nu * nu);
    const double seq = von_mises(sig);
    float util = (float)(seq / f->yield);
    if (seq > f->yield) {
        const double r = 1.0 - f->yield / seq;
        for (int i = 0; i < 3; i++) {
            X[2 * i] += (float)(r * d[6 * i]), X[2 * i + 1] += (float)(r * d[6 * i + 1]);
            d[6 * i] *= 1 - r, d[6 * i + 1] *= 1 - r;
        }
        const float mx = (X[0] + X[2] + X[4]) / 3.0f, my = (X[1] + X[3] + X[5]) / 3.0f;
        for (int i = 0; i < 3; i++) X[2 * i] -= mx, X[2 * i + 1] -= my;
        /* the plastic stretch: F = [dX] [dX0]^-1, its largest principal stretch */
        const float* X0 = &f->X0[6 * k];
        const double a0x = X0[2] - X0[0], a0y = X0[3] - X0[1], b0x = X0[4] - X0[0], b0y = X0[5] - X0[1];
        const double ax = X[2] - X[0], ay = X[3] - X[1], bx = X[4] - X[0], by = X[5] - X[1];
        const double det = a0x * b0y - b0x * a0y;
        if (fabs(det) > 1e-12) {
            const double i00 = b0y / det, i01 = -b0x / det, i10 = -a0y / det, i11 = a0x / det;
            const double F00 = ax * i00 + bx * i10, F01 = ax * i01 + bx * i11, F10 = ay * i00 + by * i10, F11 = ay * i01 + by * i11;
            const double C00 = F00 * F00 + F10 * F10, C01 = F00 * F01 + F10 * F11, C11 = F01 * F01 + F11 * F11;
            const double lmax = 0.5 * (C00 + C11) + sqrt(0.25 * (C00 - C11) * (C00 - C11) + C01 * C01);
            const float st = (float)(sqrt(lmax) - 1.0);
            f->tdmg[k] = SMAX(f->tdmg[k], st);
        }
        f->tk_ok[k] = 0;
        const float Xa[3][2] = {{X[0], X[1]}, {X[2], X[3]}, {X[4], X[5]}};
        tri_grad(Xa, b3, c3, A2);
        tri_membrane(b3, c3, *A2, d, nu, Dm, eps, sig);
    }
    if (!f->tk_ok[k]) tri_build_k(f, k);
    const float* Bc = &f->Bc[54 * (size_t)k];
    const double D = f->E * f->t * f->t * f->t / (12.0f * (1.0f - f->nu * f->nu)); /* ShellSection::D(), in float */
    const double Mp = 0.25 * f->yield * (double)f->t * f->t;
    double meq = 0;
    for (int g = 0; g < 3; g++) {
        double kap[3] = {0, 0, 0};
        for (int r = 0; r < 3; r++)
            for (int i = 0; i < 3; i++)
                kap[r] += Bc[g * 18 + r * 6 + 2 * i] * d[6 * i + 3] + Bc[g * 18 + r * 6 + 2 * i + 1] * d[6 * i + 4];
        const double m[3] = {D * (kap[0] + nu * kap[1]), D * (nu * kap[0] + kap[1]), D * 0.5 * (1 - nu) * kap[2]};
        const double vm = von_mises(m);
        meq = SMAX(meq, vm);
    }
    const float ub = (float)(meq / Mp);
    util = SMAX(util, ub);
    float* th0 = &f->th0[9 * k];
    if (meq > Mp) {
        const double r = 1.0 - Mp / meq;
        for (int i = 0; i < 3; i++) {
            th0[3 * i] += (float)(r * d[6 * i + 3]), th0[3 * i + 1] += (float)(r * d[6 * i + 4]);
            d[6 * i + 3] *= 1 - r, d[6 * i + 4] *= 1 - r;
        }
        f->tdmg[k] = SMAX(f->tdmg[k], 1e-5f);
    }
    for (int i = 0; i < 3; i++) {
        float* t0 = &th0[3 * i];
        if (t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2] > 0.02f * 0.02f) {
            float* r0 = &f->r0[12 * k + 4 * i];
            const qf q = qnorm(qmul(Q(r0[0], r0[1], r0[2], r0[3]), qexp(F3(-t0[0], -t0[1], -t0[2]))));
            r0[0] = q.x, r0[1] = q.y, r0[2] = q.z, r0[3] = q.w;
            t0[0] = t0[1] = t0[2] = 0;
        }
    }
    return util;
}

static void eval_tri(orc_fem* f, uint32_t k, tri_out* o) {
    o->on = 0;
    double d[18];
    f3 e1, e2, e3;
    if (!tri_kin(f, k, d, &e1, &e2, &e3)) return;
    const float Xa[3][2] = {{f->X[6 * k], f->X[6 * k + 1]}, {f->X[6 * k + 2], f->X[6 * k + 3]}, {f->X[6 * k + 4], f->X[6 * k + 5]}};
    double b3[3], c3[3], A2;
    tri_grad(Xa, b3, c3, &A2);
    if (!(A2 > 0)) return;
    const double nu = f->nu, Dm = (double)f->E / (1 - nu * nu);
    double eps[3], sig[3];
    tri_membrane(b3, c3, A2, d, nu, Dm, eps, sig);
    float util = 0;
    if (f->yield > 0) util = tri_plastic(f, k, d, b3, c3, &A2, eps, sig);
    f->tutil[k] = util;
    if (!f->tk_ok[k]) tri_build_k(f, k);
    const float* Km = &f->Kf[162 * (size_t)k];
    const float* Kp = Km + 81;
    double vl[18];
    {
        const double E[3][3] = {{e1.x, e2.x, e3.x}, {e1.y, e2.y, e3.y}, {e1.z, e2.z, e3.z}};
        for (int i = 0; i < 3; i++) {
            const uint32_t nd = f->tn[3 * k + i];
            const f3 v = F3(f->v[3 * nd], f->v[3 * nd + 1], f->v[3 * nd + 2]), om = F3(f->w[3 * nd], f->w[3 * nd + 1], f->w[3 * nd + 2]);
            for (int a2 = 0; a2 < 3; a2++) {
                vl[6 * i + a2] = E[0][a2] * v.x + E[1][a2] * v.y + E[2][a2] * v.z;
                vl[6 * i + 3 + a2] = E[0][a2] * om.x + E[1][a2] * om.y + E[2][a2] * om.z;
            }
        }
    }
    double fo[18], fl[18];
    {
        double dm[9], dp[9], vm[9], vp[9];
        for (int a = 0; a < 9; a++) {
            const int ci = 6 * (a / 3);
            dm[a] = d[ci + kMemDof[a % 3]], dp[a] = d[ci + kPlateDof[a % 3]];
            vm[a] = vl[ci + kMemDof[a % 3]], vp[a] = vl[ci + kPlateDof[a % 3]];
        }
        double fm[9] = {0}, fp[9] = {0}, um[9] = {0}, up[9] = {0};
        for (int c = 0; c < 9; c++)
            for (int a = 0; a < 9; a++) {
                const double km = (double)Km[c * 9 + a], kp = (double)Kp[c * 9 + a];
                fm[a] += km * dm[c], um[a] += km * vm[c];
                fp[a] += kp * dp[c], up[a] += kp * vp[c];
            }
        for (int a = 0; a < 9; a++) {
            const int ci = 6 * (a / 3);
            fo[ci + kMemDof[a % 3]] = fm[a], fl[ci + kMemDof[a % 3]] = um[a];
            fo[ci + kPlateDof[a % 3]] = fp[a], fl[ci + kPlateDof[a % 3]] = up[a];
        }
    }
    o->on = 1;
    o->e1 = e1, o->e2 = e2, o->e3 = e3;
    for (int i = 0; i < 3; i++) {
        const f3 fw = add3(add3(mul3(e1, (float)fo[6 * i]), mul3(e2, (float)fo[6 * i + 1])), mul3(e3, (float)fo[6 * i + 2]));
        const f3 tw = add3(add3(mul3(e1, (float)fo[6 * i + 3]), mul3(e2, (float)fo[6 * i + 4])), mul3(e3, (float)fo[6 * i + 5]));
        o->fo[i] = F3(-fw.x, -fw.y, -fw.z);
        o->to[i] = F3(-tw.x, -tw.y, -tw.z);
    }
    {
        const double sx = sig[0] * f->t, sy = sig[1] * f->t, sxy = sig[2] * f->t;
        const double tr = 0.5 * (sx + sy), df = sqrt(0.25 * (sx - sy) * (sx - sy) + sxy * sxy);
        const double l1 = fmax(0.0, tr + df), l2 = fmax(0.0, tr - df);
        double vx = sxy, vy = l1 - sx;
        if (fabs(vx) + fabs(vy) < 1e-12 * (fabs(sx) + fabs(sy) + 1)) vx = sx >= sy ? 1 : 0, vy = sx >= sy ? 0 : 1;
        const double vlen = sqrt(vx * vx + vy * vy);
        vx /= vlen, vy /= vlen;
        const double S[2][2] = {{l1 * vx * vx + l2 * vy * vy, (l1 - l2) * vx * vy}, {(l1 - l2) * vx * vy, l1 * vy * vy + l2 * vx * vx}};
        const double A = 0.5 * A2;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double gi[2] = {b3[i] / A2, c3[i] / A2}, gj[2] = {b3[j] / A2, c3[j] / A2};
                o->g[i][j] = (float)(A * (gi[0] * (S[0][0] * gj[0] + S[0][1] * gj[1]) + gi[1] * (S[1][0] * gj[0] + S[1][1] * gj[1])));
            }
    }
    {
        const double E[3][3] = {{e1.x, e2.x, e3.x}, {e1.y, e2.y, e3.y}, {e1.z, e2.z, e3.z}};
        for (int i = 0; i < 3; i++)
            for (int p = 0; p < 3; p++) {
                o->kv[6 * i + p] = E[p][0] * fl[6 * i] + E[p][1] * fl[6 * i + 1] + E[p][2] * fl[6 * i + 2];
                o->kv[6 * i + 3 + p] = E[p][0] * fl[6 * i + 3] + E[p][1] * fl[6 * i + 4] + E[p][2] * fl[6 * i + 5];
            }
    }
}

/* ------------------------------------------------------------------ members (frame_fem.cpp:1384-1405, 1896-2070) */
static inline qf qadd(qf a, qf b) { return Q(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline float qdot(qf a, qf b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
/* quat::rotate (core/math.h:150-154) */
static f3 qrot(qf q, f3 p) {
    const f3 u = F3(q.x, q.y, q.z);
    const f3 t = mul3(cross3(u, p), 2.0f);
    return add3(add3(p, mul3(t, q.w)), cross3(u, t));
}
/* quat::axis_angle, any_perpendicular, quat_from_to, slerp (core/math.h:82-85, 136-140, 203-212, 227-233) */
static qf axis_angle(f3 axis, float angle) {
    const float hh = angle * 0.5f, s = sinf(hh);
    const f3 a = norm3(axis);
    return Q(a.x * s, a.y * s, a.z * s, cosf(hh));
}
static qf quat_from_to(f3 a, f3 b) {
    const float d = dot3(a, b);
    if (d < -0.999999f) {
        const f3 t = fabsf(a.x) < 0.7f ? F3(1, 0, 0) : F3(0, 1, 0);
        return axis_angle(norm3(cross3(a, t)), 3.14159265358979323846f);
    }
    const f3 c = cross3(a, b);
    return qnorm(Q(c.x, c.y, c.z, 1.0f + d));
}
static qf qslerp(qf a, qf b, float t) {
    float d = qdot(a, b);
    if (d < 0) b = qscale(b, -1.0f), d = -d;
    if (d > 0.9995f) return qnorm(qadd(a, qscale(qadd(b, qscale(a, -1.0f)), t)));
    const float th = acosf(d), s = sinf(th);
    return qadd(qscale(a, sinf((1 - t) * th) / s), qscale(b, sinf(t * th) / s));
}

typedef struct {
    f3 e1, e2, e3;
    float L;
    double dL;
    f3 ra, rb;
} mem_kin_t;

/* kinematics (:1384-1405): the chord, the co-rotated frame (the ends' frames, their mean turned onto the chord), the
 * end rotations against it */
static int mem_kin(const orc_fem* f, uint32_t k, mem_kin_t* K) {
    const uint32_t a = f->mem[4 * k], b = f->mem[4 * k + 1];
    const double* xa = &f->xd[3 * a];
    const double* xb = &f->xd[3 * b];
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(L > 1e-6) || !isfinite(L)) return 0;
    K->L = (float)L;
    K->dL = L - (double)f->mL0[k];
    K->e1 = F3((float)(dx / L), (float)(dy / L), (float)(dz / L));
    const float* qa = &f->mqa[4 * k];
    const float* qb = &f->mqb[4 * k];
    const float* na = &f->q[4 * a];
    const float* nb = &f->q[4 * b];
    const qf Ea = qnorm(qmul(Q(na[0], na[1], na[2], na[3]), Q(qa[0], qa[1], qa[2], qa[3])));
    const qf Eb = qnorm(qmul(Q(nb[0], nb[1], nb[2], nb[3]), Q(qb[0], qb[1], qb[2], qb[3])));
    const float t = f->mtw[k];
    const qf Em = t <= 0.0f ? Ea : t >= 1.0f ? Eb : qslerp(Ea, Eb, t);
    const qf E = qnorm(qmul(quat_from_to(norm3(qrot(Em, F3(1, 0, 0))), K->e1), Em));
    const qf Ec = qconj(E);
    K->ra = qlog(qmul(Ec, Ea));
    K->rb = qlog(qmul(Ec, Eb));
    K->e2 = qrot(E, F3(0, 1, 0));
    K->e3 = qrot(E, F3(0, 0, 1));
    return 1;
}

typedef struct {
    int on;
    f3 fa, ta, tb;     /* the force on end a (end b takes -fa), the moments on the ends */
    f3 e1;             /* the tangent's chord */
    float ngeo;        /* the tension's geometric stiffness N / L (tension only) */
    double W[144];     /* the 12 x 12 matrix's blocks aa, bb, ba, ab (mem_kw_) */
    double kv[12];     /* K u at the ends' velocities and spins (mem_kv_) */
} mem_out;

/* element_matrix (:2157-2186) */
static void mem_matrix(f3 te1, f3 te2, f3 te3, float tL, float ka_f, float kt_f, const float* cy, const float* cz,
                       double K[12][12]) {
    const double iL = 1.0 / tL;
    const double e1[3] = {te1.x, te1.y, te1.z}, e2[3] = {te2.x, te2.y, te2.z}, e3[3] = {te3.x, te3.y, te3.z};
    const double g[3] = {e3[0] * iL, e3[1] * iL, e3[2] * iL}, h[3] = {e2[0] * iL, e2[1] * iL, e2[2] * iL};
    const double cy0 = cy[0], cy1 = cy[1], cy2 = cy[2], cz0 = cz[0], cz1 = cz[1], cz2 = cz[2];
    const double Sy = cy0 + 2 * cy1 + cy2, Sz = cz0 + 2 * cz1 + cz2, ay = cy0 + cy1, by = cy1 + cy2, az = cz0 + cz1, bz = cz1 + cz2;
    const double ka = ka_f, kt = kt_f;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double uu = ka * e1[i] * e1[j] + Sy * g[i] * g[j] + Sz * h[i] * h[j];
            const double up = -ay * g[i] * e2[j] + az * h[i] * e3[j];
            const double ub = -by * g[i] * e2[j] + bz * h[i] * e3[j];
            const double tt = kt * e1[i] * e1[j];
            const double pa = tt + cy0 * e2[i] * e2[j] + cz0 * e3[i] * e3[j];
            const double pb = -tt + cy1 * e2[i] * e2[j] + cz1 * e3[i] * e3[j];
            const double qb = tt + cy2 * e2[i] * e2[j] + cz2 * e3[i] * e3[j];
            K[i][j] = uu, K[i][3 + j] = up, K[i][6 + j] = -uu, K[i][9 + j] = ub;
            K[3 + i][3 + j] = pa, K[3 + i][9 + j] = pb;
            K[6 + i][6 + j] = uu, K[6 + i][9 + j] = -ub;
            K[9 + i][9 + j] = qb;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) K[3 + i][6 + j] = -K[j][3 + i];
    for (int r = 0; r < 12; r++)
        for (int c = r + 1; c < 12; c++) K[c][r] = K[r][c];
}

/* FemFrame::effective (:2147-2155): the tangent's stiffness where the member yields and the step does not unload it -
 * a flowing hinge turns like a pin, a flowing bar or twist keeps a hundredth */
static void mem_effective(const mem_tan* t, float* ka, float* kt, float cy[3], float cz[3]) {
    const uint8_t y = (uint8_t)(t->yield & ~t->unload);
    *ka = t->ka, *kt = t->kt;
    for (int i = 0; i < 3; i++) cy[i] = t->cy[i], cz[i] = t->cz[i];
    if (y & 3) mem_condense(cy, y & 1, (y & 2) != 0), mem_condense(cz, y & 1, (y & 2) != 0);
    if (y & 4) *ka *= 0.01f;
    if (y & 8) *kt *= 0.01f;
}

/* element_matrix at a member's effective tangent (member_world_k, :1834-1850): the blocks aa, bb, ba, ab */
static void mem_world_k(const mem_tan* t, double W[144], double M[12][12]) {
    float ka, kt, cy[3], cz[3];
    mem_effective(t, &ka, &kt, cy, cz);
    mem_matrix(t->e1, t->e2, t->e3, t->L, ka, kt, cy, cz, M);
    for (int r = 0; r < 6; r++)
        for (int q = 0; q < 6; q++)
            W[r * 6 + q] = M[r][q], W[36 + r * 6 + q] = M[6 + r][6 + q], W[72 + r * 6 + q] = M[6 + r][q],
            W[108 + r * 6 + q] = M[r][6 + q];
}

/* the end moments (y and z components) at the end rotations less the hinges' plastic ones (:1946-1950) */
static void mem_moments(const mem_kin_t* K, const float* st, const float* ps, float Ma[2], float Mb[2]) {
    const float dax = K->ra.y - ps[PS_PA], day = K->ra.z - ps[PS_PA + 1];
    const float dbx = K->rb.y - ps[PS_PB], dby = K->rb.z - ps[PS_PB + 1];
    Ma[0] = st[2] * dax + st[3] * dbx, Ma[1] = st[5] * day + st[6] * dby;
    Mb[0] = st[3] * dax + st[4] * dbx, Mb[1] = st[6] * day + st[7] * dby;
}

static inline float len2f(float x, float y) { return sqrtf(x * x + y * y); }

/* a hinge's radial return (:1960-1972): the yield moment hardening with the hinge's largest rotation */
static int mem_yield_end(const float* s, const float M[2], float cyy, float czz, float* p, const float* base, float* dmg,
                         float* flow) {
    const float Mp = s[ORC_SEC_MP] * (1.0f + s[ORC_SEC_HARDENING] * SMIN(1.0f, *dmg));
    const float l = len2f(M[0], M[1]);
    if (l <= Mp) return 0;
    const float r = 1.0f - Mp / l;
    const float dp0 = cyy > 0 ? M[0] * r / cyy : 0.0f, dp1 = czz > 0 ? M[1] * r / czz : 0.0f;
    p[0] += dp0, p[1] += dp1;
    *flow += len2f(dp0, dp1);
    const float lp = len2f(p[0] + base[0], p[1] + base[1]);
    *dmg = SMAX(*dmg, lp);
    return 1;
}

/* eval_forces for a member (:1905-2069, the body's allow_break off; no joint damping): the axial force (yielding in
 * tension at Np, in compression at the lower of Np and the Euler load), the torque (at Tp), the end moments (the
 * hinges' radial return, twice), the ends' forces and moments; the plastic rotations past 0.02 rad baked into the
 * rest frames; the tangent (its yield bits where it flows this step, elastic where held or the budget is spent) and
 * member_world_k */
static void eval_member(orc_fem* f, uint32_t k, mem_out* o, int elastic_step) {
    o->on = 0;
    mem_tan* t = &f->tan[k];
    t->on = 0;
    if (f->mbroken[k]) return;
    mem_kin_t K;
    if (!mem_kin(f, k, &K)) return;
    const float* s = &f->sec[ORC_FRAME_SECTION_FIELDS * f->mem[4 * k + 2]];
    const float* st = &f->mst[8 * k];
    const float ka = st[0], kt = st[1];
    float* ps = &f->mps[PS_N_ * (size_t)k];
    const float L0 = f->mL0[k];
    const int plastic = s[ORC_SEC_NP] > 0;
    const float delta = (float)K.dL;
    float N = ka * (delta - ps[PS_UP]);
    float Nt = s[ORC_SEC_NP], Nc = s[ORC_SEC_NP];
    float flow_ax = 0, flow_t = 0;
    if (plastic) {
        const float Le = 0.7f * L0, kPi = 3.14159265358979323846f;
        const float eul = kPi * kPi * s[ORC_SEC_E] * SMIN(s[ORC_SEC_IY], s[ORC_SEC_IZ]) / (Le * Le);
        Nc = SMIN(s[ORC_SEC_NP], eul);
        if (N > Nt || N < -Nc) {
            const float lim = N > 0 ? Nt : -Nc;
            const float dup = (N - lim) / ka;
            ps[PS_UP] += dup;
            flow_ax = fabsf(dup) / L0;
            const float dm = ps[PS_UP] / L0;
            ps[PS_DAMAGE] = SMAX(ps[PS_DAMAGE], dm);
            N = lim;
        }
    }
    const float tw = K.rb.x - K.ra.x;
    float T = kt * (tw - ps[PS_TP]);
    if (plastic && s[ORC_SEC_TP] > 0 && kt > 0 && fabsf(T) > s[ORC_SEC_TP]) {
        const float lim = T > 0 ? s[ORC_SEC_TP] : -s[ORC_SEC_TP];
        const float dtp = (T - lim) / kt;
        ps[PS_TP] += dtp;
        flow_t = fabsf(dtp);
        const float dm = fabsf(ps[PS_TP] + ps[PS_BT]) * s[ORC_SEC_HALF] / L0;
        ps[PS_DAMAGE] = SMAX(ps[PS_DAMAGE], dm);
        T = lim;
    }
    float Ma[2], Mb[2];
    mem_moments(&K, st, ps, Ma, Mb);
    float flow_a = 0, flow_b = 0;
    if (plastic && s[ORC_SEC_MP] > 0)
        for (int it = 0; it < 2; it++) {
            if (mem_yield_end(s, Ma, st[2], st[5], &ps[PS_PA], &ps[PS_BA], &ps[PS_DMG_A], &flow_a)) mem_moments(&K, st, ps, Ma, Mb);
            if (mem_yield_end(s, Mb, st[4], st[7], &ps[PS_PB], &ps[PS_BB], &ps[PS_DMG_B], &flow_b)) mem_moments(&K, st, ps, Ma, Mb);
        }
    float util = 0;
    if (s[ORC_SEC_NP] > 0) {
        const float u = N > 0 ? N / Nt : -N / Nc;
        util = SMAX(util, u);
    }
    if (s[ORC_SEC_MP] > 0) {
        const float ua = len2f(Ma[0], Ma[1]) / (1.0f + s[ORC_SEC_HARDENING] * SMIN(1.0f, ps[PS_DMG_A]));
        const float ub = len2f(Mb[0], Mb[1]) / (1.0f + s[ORC_SEC_HARDENING] * SMIN(1.0f, ps[PS_DMG_B]));
        const float u = SMAX(ua, ub) / s[ORC_SEC_MP];
        util = SMAX(util, u);
    }
    if (s[ORC_SEC_TP] > 0) {
        const float u = fabsf(T) / s[ORC_SEC_TP];
        util = SMAX(util, u);
    }
    ps[PS_UTIL] = util, ps[PS_N] = N;
    const float Sy = Ma[0] + Mb[0], Sz = Ma[1] + Mb[1];
    o->fa = sub3(add3(mul3(K.e1, N), mul3(K.e3, Sy / K.L)), mul3(K.e2, Sz / K.L));
    o->ta = sub3(sub3(mul3(K.e1, T), mul3(K.e2, Ma[0])), mul3(K.e3, Ma[1]));
    o->tb = sub3(sub3(mul3(K.e1, -T), mul3(K.e2, Mb[0])), mul3(K.e3, Mb[1]));
    o->on = 1;
    /* the plastic rotations into the rest frames (:2037-2045) */
    const float kBake = 0.02f;
    float* qa = &f->mqa[4 * k];
    float* qb = &f->mqb[4 * k];
    if (ps[PS_PA] * ps[PS_PA] + ps[PS_PA + 1] * ps[PS_PA + 1] > kBake * kBake) {
        const qf q = qnorm(qmul(Q(qa[0], qa[1], qa[2], qa[3]), qexp(F3(0, -ps[PS_PA], -ps[PS_PA + 1]))));
        qa[0] = q.x, qa[1] = q.y, qa[2] = q.z, qa[3] = q.w;
        ps[PS_BA] += ps[PS_PA], ps[PS_BA + 1] += ps[PS_PA + 1], ps[PS_PA] = ps[PS_PA + 1] = 0;
    }
    if (ps[PS_PB] * ps[PS_PB] + ps[PS_PB + 1] * ps[PS_PB + 1] > kBake * kBake) {
        const qf q = qnorm(qmul(Q(qb[0], qb[1], qb[2], qb[3]), qexp(F3(0, -ps[PS_PB], -ps[PS_PB + 1]))));
        qb[0] = q.x, qb[1] = q.y, qb[2] = q.z, qb[3] = q.w;
        ps[PS_BB] += ps[PS_PB], ps[PS_BB + 1] += ps[PS_PB + 1], ps[PS_PB] = ps[PS_PB + 1] = 0;
    }
    if (fabsf(ps[PS_TP]) > kBake) {
        const qf q1 = qnorm(qmul(Q(qa[0], qa[1], qa[2], qa[3]), qexp(F3(0.5f * ps[PS_TP], 0, 0))));
        const qf q2 = qnorm(qmul(Q(qb[0], qb[1], qb[2], qb[3]), qexp(F3(-0.5f * ps[PS_TP], 0, 0))));
        qa[0] = q1.x, qa[1] = q1.y, qa[2] = q1.z, qa[3] = q1.w;
        qb[0] = q2.x, qb[1] = q2.y, qb[2] = q2.z, qb[3] = q2.w;
        ps[PS_BT] += ps[PS_TP], ps[PS_TP] = 0;
    }
    /* the tangent (:2051-2068) */
    t->on = 1;
    t->e1 = K.e1, t->e2 = K.e2, t->e3 = K.e3, t->L = K.L, t->ka = ka, t->kt = kt;
    for (int i = 0; i < 3; i++) t->cy[i] = st[2 + i], t->cz[i] = st[5 + i];
    t->ngeo = N > 0 ? N / K.L : 0.0f;
    const float kFlow = 1e-4f;
    t->yield = (uint8_t)((flow_a > kFlow ? 1 : 0) | (flow_b > kFlow ? 2 : 0) | (flow_ax > 0.1f * kFlow ? 4 : 0) | (flow_t > kFlow ? 8 : 0));
    t->unload = 0;
    if (t->hold > 0) t->hold--, t->unload = t->yield;
    if (elastic_step) t->unload = t->yield;
    t->ma[0] = Ma[0], t->ma[1] = Ma[1], t->mb[0] = Mb[0], t->mb[1] = Mb[1], t->N = N, t->T = T;
    o->e1 = K.e1;
    o->ngeo = t->ngeo;
    double M[12][12];
    mem_world_k(t, o->W, M);
    const uint32_t a = f->mem[4 * k], b = f->mem[4 * k + 1];
    const double u[12] = {f->v[3 * a], f->v[3 * a + 1], f->v[3 * a + 2], f->w[3 * a], f->w[3 * a + 1], f->w[3 * a + 2],
                          f->v[3 * b], f->v[3 * b + 1], f->v[3 * b + 2], f->w[3 * b], f->w[3 * b + 1], f->w[3 * b + 2]};
    for (int r = 0; r < 12; r++) {
        double s2 = 0;
        for (int q = 0; q < 12; q++) s2 += M[r][q] * u[q];
        o->kv[r] = s2;
    }
}

/* gather_part's member_block (:926-941): block blk (0 aa, 1 bb, 2 ba, 3 ab) times c, the tension's geometric stiffness
 * with sign gs */
static void mem_block(const orc_fem* f, uint32_t k, const mem_out* o, int blk, double gs, double h2, double h, double* D) {
    const float* s = &f->sec[ORC_FRAME_SECTION_FIELDS * f->mem[4 * k + 2]];
    const double c = h2 + (double)s[ORC_SEC_DAMPING] * h;
    const double* B = &o->W[blk * 36];
    for (int x = 0; x < 36; x++) D[x] += c * B[x];
    const double gg = h2 * o->ngeo;
    if (gg != 0) {
        double P[9];
        const double ee[3] = {o->e1.x, o->e1.y, o->e1.z};
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) P[r * 3 + q] = gg * ((r == q ? 1.0 : 0.0) - ee[r] * ee[q]);
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) D[r * 6 + q] += gs * P[r * 3 + q];
    }
}

void orc_fem_add_torques(orc_fem* f, const float* torques) {
    for (uint32_t i = 0; i < 3 * f->nn; i++) f->torque[i] += torques[i];
}

/* gather_part's tri_block (frame_fem.cpp:948-985): corners (i, j) in world axes, times c, tension on the diagonal */
static void tri_block(const orc_fem* f, uint32_t k, const tri_out* o, int i, int j, double c, double h2, double* D) {
    const double E[3][3] = {{o->e1.x, o->e2.x, o->e3.x}, {o->e1.y, o->e2.y, o->e3.y}, {o->e1.z, o->e2.z, o->e3.z}};
    const float* Km = &f->Kf[162 * (size_t)k];
    const float* Kp = Km + 81;
#define M_(s2, t2) ((double)Km[(3 * i + (s2)) * 9 + 3 * j + (t2)])
#define P_(s2, t2) ((double)Kp[(3 * i + (s2)) * 9 + 3 * j + (t2)])
    for (int ra = 0; ra < 2; ra++)
        for (int cb = 0; cb < 2; cb++) {
            double T[3][3];
            for (int p = 0; p < 3; p++) {
                const double E0 = E[p][0], E1 = E[p][1], E2 = E[p][2];
                if (ra == 0 && cb == 0) {
                    T[p][0] = E0 * M_(0, 0) + E1 * M_(1, 0);
                    T[p][1] = E0 * M_(0, 1) + E1 * M_(1, 1);
                    T[p][2] = E2 * P_(0, 0);
                } else if (ra == 0) {
                    T[p][0] = E2 * P_(0, 1);
                    T[p][1] = E2 * P_(0, 2);
                    T[p][2] = E0 * M_(0, 2) + E1 * M_(1, 2);
                } else if (cb == 0) {
                    T[p][0] = E2 * M_(2, 0);
                    T[p][1] = E2 * M_(2, 1);
                    T[p][2] = E0 * P_(1, 0) + E1 * P_(2, 0);
                } else {
                    T[p][0] = E0 * P_(1, 1) + E1 * P_(2, 1);
                    T[p][1] = E0 * P_(1, 2) + E1 * P_(2, 2);
                    T[p][2] = E2 * M_(2, 2);
                }
            }
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) D[(3 * ra + p) * 6 + 3 * cb + q] += c * (T[p][0] * E[q][0] + T[p][1] * E[q][1] + T[p][2] * E[q][2]);
        }
#undef M_
#undef P_
    const double gg = h2 * o->g[i][j];
    if (gg != 0)
        for (int p = 0; p < 3; p++) D[p * 7] += gg;
}

/* ------------------------------------------------------------------ the reference's 6x6 block kernels */
static int chol6(double* A, double* dinv) {
    for (int j = 0; j < 6; j++) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; k++) s -= A[j * 6 + k] * A[j * 6 + k];
        if (!(s > 0)) return 0;
        const double d = sqrt(s), id = 1.0 / d;
        A[j * 6 + j] = d;
        dinv[j] = id;
        for (int i = j + 1; i < 6; i++) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; k++) t -= A[i * 6 + k] * A[j * 6 + k];
            A[i * 6 + j] = t * id;
        }
        for (int i = 0; i < j; i++) A[i * 6 + j] = 0;
    }
    return 1;
}
static void right_solve_lt(const double* L, const double* dinv, double* X) {
    for (int r = 0; r < 6; r++) {
        double* x = X + r * 6;
        for (int j = 0; j < 6; j++) {
            double s = x[j];
            for (int m = 0; m < j; m++) s -= L[j * 6 + m] * x[m];
            x[j] = s * dinv[j];
        }
    }
}
static void sub_abt(double* T, const double* P, const double* Qm) {
    double Qt[36];
    for (int c = 0; c < 6; c++)
        for (int m = 0; m < 6; m++) Qt[m * 6 + c] = Qm[c * 6 + m];
    for (int r = 0; r < 6; r++) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int m = 0; m < 6; m++) {
            const double p = P[r * 6 + m];
            for (int c = 0; c < 6; c++) acc[c] += p * Qt[m * 6 + c];
        }
        for (int c = 0; c < 6; c++) T[r * 6 + c] -= acc[c];
    }
}
static void sub_mv(const double* L, const double* y, double* r) {
    for (int i = 0; i < 6; i++) {
        double s = 0;
        for (int j = 0; j < 6; j++) s += L[i * 6 + j] * y[j];
        r[i] -= s;
    }
}
static void sub_mtv(const double* L, const double* xr, double* x) {
    for (int j = 0; j < 6; j++) {
        double s = 0;
        for (int i = 0; i < 6; i++) s += L[i * 6 + j] * xr[i];
        x[j] -= s;
    }
}
static void forward6(const double* L, const double* dinv, double* y) {
    for (int j = 0; j < 6; j++) {
        double s = y[j];
        for (int m = 0; m < j; m++) s -= L[j * 6 + m] * y[m];
        y[j] = s * dinv[j];
    }
}
static void backward6(const double* L, const double* dinv, double* x) {
    for (int j = 5; j >= 0; j--) {
        double s = x[j];
        for (int m = j + 1; m < 6; m++) s -= L[m * 6 + j] * x[m];
        x[j] = s * dinv[j];
    }
}

/* dense (6n)^2 row-major <-> blocks */
static void get_block(const double* A, int n, int r, int c, double* B) {
    for (int p = 0; p < 6; p++)
        for (int q = 0; q < 6; q++) B[p * 6 + q] = A[(size_t)(6 * r + p) * 6 * n + 6 * c + q];
}
static void put_block(double* A, int n, int r, int c, const double* B) {
    for (int p = 0; p < 6; p++)
        for (int q = 0; q < 6; q++) A[(size_t)(6 * r + p) * 6 * n + 6 * c + q] = B[p * 6 + q];
}

/* The block lower triangle in envelope storage: block row r holds its blocks of columns fc[r] .. r (36 doubles each,
 * from roff[r]); the right-looking factorization keeps its fill inside that envelope. The reference's loops
 * (frame_fem.cpp:2899-2919, then L L^T x = b, :2933-2942) over the blocks of the natural order; a block outside the
 * envelope is zero, and its updates (T -= 0) would leave every value as it is, so they are skipped. */
#define ENV(r, c) (&L[(roff[r] + (size_t)((c) - fc[r])) * 36])
static int env_factor(int n, const int* fc, const size_t* roff, double* L, double* dinv) {
    for (int k = 0; k < n; k++) {
        double* Lkk = ENV(k, k);
        if (!chol6(Lkk, &dinv[6 * k])) return 0;
        for (int r = k + 1; r < n; r++)
            if (fc[r] <= k) right_solve_lt(Lkk, &dinv[6 * k], ENV(r, k));
        for (int r = k + 1; r < n; r++) {
            if (fc[r] > k) continue;
            for (int c = k + 1; c <= r; c++)
                if (fc[c] <= k) sub_abt(ENV(r, c), ENV(r, k), ENV(c, k));
        }
    }
    return 1;
}

static void env_subst(int n, const int* fc, const size_t* roff, const double* L, const double* dinv, double* b) {
    for (int k = 0; k < n; k++) {
        forward6(ENV(k, k), &dinv[6 * k], &b[6 * k]);
        for (int r = k + 1; r < n; r++)
            if (fc[r] <= k) sub_mv(ENV(r, k), &b[6 * k], &b[6 * r]);
    }
    for (int k = n - 1; k >= 0; k--) {
        for (int r = k + 1; r < n; r++)
            if (fc[r] <= k) sub_mtv(ENV(r, k), &b[6 * r], &b[6 * k]);
        backward6(ENV(k, k), &dinv[6 * k], &b[6 * k]);
    }
}

/* factor in place (dinv: 6 per block row) and solve; 0, or -1 when a pivot block is not positive definite */
static int env_solve(int n, const int* fc, const size_t* roff, double* L, double* dinv, double* b) {
    if (!env_factor(n, fc, roff, L, dinv)) return -1;
    env_subst(n, fc, roff, L, dinv, b);
    return 0;
}

void orc_fem_plastic(const orc_fem* f, float* tris, float* members, int64_t counters[2]) {
    if (tris)
        for (uint32_t k = 0; k < f->nt; k++) {
            float* o = &tris[ORC_TRI_STATE * (size_t)k];
            memcpy(o, &f->X[6 * k], 6 * sizeof(float));
            memcpy(o + 6, &f->th0[9 * k], 9 * sizeof(float));
            memcpy(o + 15, &f->r0[12 * k], 12 * sizeof(float));
            o[27] = f->tdmg[k], o[28] = f->tutil[k];
        }
    if (members)
        for (uint32_t k = 0; k < f->nm; k++) {
            float* o = &members[ORC_MEM_STATE * (size_t)k];
            memcpy(o, &f->mps[PS_N_ * (size_t)k], PS_N_ * sizeof(float));
            memcpy(o + 16, &f->mqa[4 * k], 4 * sizeof(float));
            memcpy(o + 20, &f->mqb[4 * k], 4 * sizeof(float));
            o[24] = f->tan[k].yield, o[25] = f->tan[k].unload, o[26] = f->tan[k].hold;
        }
    if (counters) counters[0] = f->repasses, counters[1] = f->passes;
}

int orc_fem_block_solve(int n, double* A, double* b) {
    /* the envelope of A's nonzero lower blocks; the factor written back over A's lower blocks */
    int* fc = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
    size_t* roff = (size_t*)malloc(sizeof(size_t) * (size_t)(n + 1));
    size_t tot = 0;
    for (int r = 0; r < n; r++) {
        int first = r;
        for (int c = 0; c < r && first == r; c++) {
            double B[36];
            get_block(A, n, r, c, B);
            for (int q = 0; q < 36; q++)
                if (B[q] != 0) { first = c; break; }
        }
        fc[r] = first, roff[r] = tot, tot += (size_t)(r - first + 1);
    }
    roff[n] = tot;
    double* L = (double*)malloc(sizeof(double) * 36 * (tot ? tot : 1));
    double* dinv = (double*)malloc(sizeof(double) * 6 * (size_t)(n ? n : 1));
    for (int r = 0; r < n; r++)
        for (int c = fc[r]; c <= r; c++) get_block(A, n, r, c, ENV(r, c));
    const int rc = env_solve(n, fc, roff, L, dinv, b);
    free(dinv);
    if (rc == 0)
        for (int r = 0; r < n; r++)
            for (int c = fc[r]; c <= r; c++) put_block(A, n, r, c, ENV(r, c));
    free(fc), free(roff), free(L);
    return rc;
}

/* solve_impl's add_member (:2684-2730) for a pass again: a member's effective-tangent matrix times sign into the
 * assembled system (its ends' diagonal blocks, the block of the pair, the right side's damping term); the geometric
 * stiffness is left as it was */
static void add_member(const orc_fem* f, uint32_t k, const mem_tan* t, double sign, const int* fi, const int* fc,
                       const size_t* roff, double* L, double* R, double h, double h2, double hd) {
    double K[12][12], W[144];
    mem_world_k(t, W, K);
    const float* s = &f->sec[ORC_FRAME_SECTION_FIELDS * f->mem[4 * k + 2]];
    const double beta = (double)s[ORC_SEC_DAMPING] * h;
    const double c = sign * (h2 + beta), cr = sign * (hd + beta);
    const uint32_t a = f->mem[4 * k], b = f->mem[4 * k + 1];
    const int ka = fi[a], kb = fi[b], fa = ka < 0, fb = kb < 0;
    if (cr != 0) {
        const double u[12] = {f->v[3 * a], f->v[3 * a + 1], f->v[3 * a + 2], f->w[3 * a], f->w[3 * a + 1], f->w[3 * a + 2],
                              f->v[3 * b], f->v[3 * b + 1], f->v[3 * b + 2], f->w[3 * b], f->w[3 * b + 1], f->w[3 * b + 2]};
        for (int r = 0; r < 12; r++) {
            double s2 = 0;
            for (int q = 0; q < 12; q++) s2 += K[r][q] * u[q];
            if (r < 6) {
                if (!fa) R[6 * (size_t)ka + r] -= cr * s2;
            } else if (!fb) {
                R[6 * (size_t)kb + r - 6] -= cr * s2;
            }
        }
    }
#define ADD_BLOCK(D, ro, co)                                                                                          \
    for (int r = 0; r < 6; r++)                                                                                       \
        for (int q = 0; q < 6; q++) (D)[r * 6 + q] += c * K[(ro) + r][(co) + q];
    if (!fa) { ADD_BLOCK(ENV(ka, ka), 0, 0) }
    if (!fb) { ADD_BLOCK(ENV(kb, kb), 6, 6) }
    if (!fa && !fb && ka != kb) {
        if (kb > ka) { ADD_BLOCK(ENV(kb, ka), 6, 0) }
        else { ADD_BLOCK(ENV(ka, kb), 0, 6) }
    }
#undef ADD_BLOCK
}

/* solve_impl's passes (:2741-2760, 2899-2996): factor and solve; if a member yields, the rates of its deformations at
 * the solution's velocities - a hinge turning against its moment, a bar against its force, a twist against its
 * torque - unload it: its elastic tangent (held 16 steps), the system corrected and solved again, at most three
 * passes. A pass again of a component of 60 nodes or more is counted and spends the budget. Returns env_solve's. */
static int solve_passes(orc_fem* f, const mem_out* mo, const int* fi, int nf, const int* fc, const size_t* roff,
                        size_t tot, double* L, double* dinv, double* r, double h, double h2, double hd) {
    const uint32_t nm = f->nm;
    int any_yield = 0;
    for (uint32_t k = 0; k < nm; k++) any_yield |= mo[k].on && (f->tan[k].yield & ~f->tan[k].unload);
    double *La = NULL, *ra = NULL;
    uint32_t* chg = NULL;
    uint8_t *was = NULL, *again = NULL;
    if (any_yield) {
        La = (double*)malloc(sizeof(double) * 36 * (tot ? tot : 1));
        ra = (double*)malloc(sizeof(double) * 6 * (size_t)(nf ? nf : 1));
        memcpy(La, L, sizeof(double) * 36 * tot), memcpy(ra, r, sizeof(double) * 6 * (size_t)nf);
        chg = (uint32_t*)malloc(sizeof(uint32_t) * (nm + 1));
        was = (uint8_t*)malloc(nm + 1);
        again = (uint8_t*)malloc(f->ncomp + 1);
    }
    int rc = 0, nchg = 0;
    for (int pass = 0; pass < 3; pass++) {
        f->passes++;
        if (pass > 0) {
            for (int c = 0; c < nchg; c++) {
                mem_tan t0 = f->tan[chg[c]];
                t0.unload = was[c];
                add_member(f, chg[c], &t0, -1.0, fi, fc, roff, La, ra, h, h2, hd);
                add_member(f, chg[c], &f->tan[chg[c]], 1.0, fi, fc, roff, La, ra, h, h2, hd);
            }
            nchg = 0;
            memcpy(L, La, sizeof(double) * 36 * tot), memcpy(r, ra, sizeof(double) * 6 * (size_t)nf);
        }
        rc = env_solve(nf, fc, roff, L, dinv, r);
        if (rc || !any_yield || pass == 2) break;
        memset(again, 0, f->ncomp + 1);
        for (uint32_t k = 0; k < nm; k++) {
            mem_tan* t = &f->tan[k];
            const uint8_t y = (uint8_t)(t->yield & ~t->unload);
            if (!t->on || !y || f->mbroken[k]) continue;
            f3 vv[2], ww[2];
            for (int e = 0; e < 2; e++) {
                const uint32_t n = f->mem[4 * k + e];
                if (fi[n] < 0) {
                    vv[e] = ww[e] = F3(0, 0, 0);
                    continue;
                }
                const double* x = &r[6 * (size_t)fi[n]];
                vv[e] = add3(F3(f->v[3 * n], f->v[3 * n + 1], f->v[3 * n + 2]), F3((float)x[0], (float)x[1], (float)x[2]));
                ww[e] = add3(F3(f->w[3 * n], f->w[3 * n + 1], f->w[3 * n + 2]), F3((float)x[3], (float)x[4], (float)x[5]));
            }
            const float iL = 1.0f / t->L;
            const f3 dv = sub3(vv[1], vv[0]);
            const float rax = dot3(t->e3, dv) * iL + dot3(t->e2, ww[0]), ray = -dot3(t->e2, dv) * iL + dot3(t->e3, ww[0]);
            const float rbx = dot3(t->e3, dv) * iL + dot3(t->e2, ww[1]), rby = -dot3(t->e2, dv) * iL + dot3(t->e3, ww[1]);
            uint8_t u = 0;
            if ((y & 1) && rax * t->ma[0] + ray * t->ma[1] < 0) u |= 1;
            if ((y & 2) && rbx * t->mb[0] + rby * t->mb[1] < 0) u |= 2;
            if ((y & 4) && dot3(t->e1, dv) * t->N < 0) u |= 4;
            if ((y & 8) && dot3(t->e1, sub3(ww[1], ww[0])) * t->T < 0) u |= 8;
            if (u) {
                chg[nchg] = k, was[nchg++] = t->unload;
                t->unload |= u, t->hold = 16;
                again[f->comp[f->mem[4 * k]]] = 1;
            }
        }
        for (uint32_t c = 0; c < f->ncomp; c++)
            if (again[c] && f->comp_size[c] >= 60) f->repasses++, f->repass_credit -= 1.0f;
        if (!nchg) break;
    }
    free(La), free(ra), free(chg), free(was), free(again);
    return rc;
}

/* SoftBody::integrate_nodes (softbody.cpp:769-795) and FemFrame::sync_positions (frame_fem.cpp:3407-3422) over a
 * substep of `step`: v += F (m^-1 step), clamped at 400 m/s, a non-finite v reset (counted); p += v step; x += step v in
 * double, the orientation turned by w step, the float position taken from the double one where within the float's
 * rounding, else the other way. F is kept as the body's forces. */
static void integrate_sync(orc_fem* f, const f3* F, float h) {
    const uint32_t nn = f->nn;
    for (uint32_t i = 0; i < nn; i++) {
        f->force[3 * i] = F[i].x, f->force[3 * i + 1] = F[i].y, f->force[3 * i + 2] = F[i].z;
        if (!(f->inv_mass[i] > 0)) continue;
        f3 v = add3(F3(f->v[3 * i], f->v[3 * i + 1], f->v[3 * i + 2]), mul3(F[i], f->inv_mass[i] * h));
        const float v2 = dot3(v, v);
        if (v2 > 400.0f * 400.0f) {
            v = mul3(v, 400.0f / sqrtf(v2));
        } else if (!(v2 >= 0.0f)) {
            v = F3(0, 0, 0);
            f->nan_resets++;
        }
        f->v[3 * i] = v.x, f->v[3 * i + 1] = v.y, f->v[3 * i + 2] = v.z;
        const f3 pn = add3(F3(f->p[3 * i], f->p[3 * i + 1], f->p[3 * i + 2]), mul3(v, h));
        f->p[3 * i] = pn.x, f->p[3 * i + 1] = pn.y, f->p[3 * i + 2] = pn.z;
    }
    for (uint32_t i = 0; i < nn; i++) {
        double* x = &f->xd[3 * i];
        float* pp = &f->p[3 * i];
        x[0] += (double)h * f->v[3 * i], x[1] += (double)h * f->v[3 * i + 1], x[2] += (double)h * f->v[3 * i + 2];
        const f3 wv = F3(f->w[3 * i], f->w[3 * i + 1], f->w[3 * i + 2]);
        if (dot3(wv, wv) > 0) {
            const qf qn = qnorm(qmul(qexp(mul3(wv, h)), Q(f->q[4 * i], f->q[4 * i + 1], f->q[4 * i + 2], f->q[4 * i + 3])));
            f->q[4 * i] = qn.x, f->q[4 * i + 1] = qn.y, f->q[4 * i + 2] = qn.z, f->q[4 * i + 3] = qn.w;
        }
        const double tol = 1e-4 * (1.0 + fabsf(pp[0]) + fabsf(pp[1]) + fabsf(pp[2]));
        if (f->inv_mass[i] > 0 && fabs(x[0] - pp[0]) <= tol && fabs(x[1] - pp[1]) <= tol && fabs(x[2] - pp[2]) <= tol)
            pp[0] = (float)x[0], pp[1] = (float)x[1], pp[2] = (float)x[2];
        else
            x[0] = pp[0], x[1] = pp[1], x[2] = pp[2];
    }
}

/* ------------------------------------------------------------------ the step */
int orc_fem_step(orc_fem* f, const float* loads, float h, float theta, float dissipation, const float g[3], double* Aout,
                 double* rhs_out, double* x_out, uint32_t* free_map, uint32_t* nfree_out) {
    return orc_fem_step_span(f, loads, h, h, theta, dissipation, g, Aout, rhs_out, x_out, free_map, nfree_out);
}

int orc_fem_step_span(orc_fem* f, const float* loads, float h, float step, float theta, float dissipation,
                      const float g[3], double* Aout, double* rhs_out, double* x_out, uint32_t* free_map,
                      uint32_t* nfree_out) {
    const uint32_t nn = f->nn, nt = f->nt;
    /* theta and the dissipation as FemFrame::solve_begin takes them (frame_fem.cpp:2524-2525) */
    dissipation = dissipation < 0.0f ? 0.0f : (dissipation > 1.0f ? 1.0f : dissipation);
    {
        const float lo = 0.25f + 0.5f * dissipation;
        theta = theta < lo ? lo : (theta > 1.0f ? 1.0f : theta);
    }
    /* FemFrame::begin_forces (frame_fem.cpp:1867-1872): a node not where its double position rounds to was moved by
     * something else; its double position is taken from it */
    for (uint32_t i = 0; i < nn; i++) {
        double* x = &f->xd[3 * i];
        const float* pp = &f->p[3 * i];
        if ((float)x[0] != pp[0] || (float)x[1] != pp[1] || (float)x[2] != pp[2]) x[0] = pp[0], x[1] = pp[1], x[2] = pp[2];
    }
    /* the re-pass budget (begin_forces, :1863-1864): a pass again costs one credit, one more every fourth step, at
     * most 8; with less than one left the yielding members start from the elastic tangent */
    {
        const float cr = f->repass_credit + 0.25f;
        f->repass_credit = SMIN(8.0f, cr);
    }
    const int elastic_step = f->repass_credit < 1.0f;
    const uint32_t nm = f->nm;
    mem_out* mo = (mem_out*)calloc(nm + 1, sizeof(mem_out));
    for (uint32_t k = 0; k < nm; k++) eval_member(f, k, &mo[k], elastic_step);
    tri_out* o = (tri_out*)calloc(nt + 1, sizeof(tri_out));
    for (uint32_t k = 0; k < nt; k++) eval_tri(f, k, &o[k]);
    /* the node forces: gravity m g, the loads, then the members' ends in member order, then the triangles' corners
     * in triangle order (gather_lists, :862-898); the moments: those from outside, then the same */
    f3* F = (f3*)malloc(sizeof(f3) * (nn + 1));
    f3* T = (f3*)calloc(nn + 1, sizeof(f3));
    /* (the elements' share alone beside them: member_f, member_t_, gather_forces :2080-2100) */
    f3* Fm = (f3*)calloc(nn + 1, sizeof(f3));
    f3* Tm = (f3*)calloc(nn + 1, sizeof(f3));
    for (uint32_t i = 0; i < nn; i++) {
        F[i] = mul3(F3(g[0], g[1], g[2]), f->mass[i]);
        if (loads) F[i] = add3(F[i], F3(loads[3 * i], loads[3 * i + 1], loads[3 * i + 2]));
        T[i] = F3(f->torque[3 * i], f->torque[3 * i + 1], f->torque[3 * i + 2]);
    }
    for (uint32_t k = 0; k < nm; k++)
        if (mo[k].on) {
            const uint32_t a = f->mem[4 * k], b = f->mem[4 * k + 1];
            F[a] = add3(F[a], mo[k].fa), Fm[a] = add3(Fm[a], mo[k].fa), T[a] = add3(T[a], mo[k].ta), Tm[a] = add3(Tm[a], mo[k].ta);
            F[b] = sub3(F[b], mo[k].fa), Fm[b] = sub3(Fm[b], mo[k].fa), T[b] = add3(T[b], mo[k].tb), Tm[b] = add3(Tm[b], mo[k].tb);
        }
    for (uint32_t k = 0; k < nt; k++)
        if (o[k].on)
            for (int c = 0; c < 3; c++) {
                const uint32_t nd = f->tn[3 * k + c];
                F[nd] = add3(F[nd], o[k].fo[c]), Fm[nd] = add3(Fm[nd], o[k].fo[c]);
                T[nd] = add3(T[nd], o[k].to[c]), Tm[nd] = add3(Tm[nd], o[k].to[c]);
            }
    /* the free nodes in natural order */
    int* fi = (int*)malloc(sizeof(int) * nn);
    int nf = 0;
    for (uint32_t i = 0; i < nn; i++) fi[i] = f->fixed[i] ? -1 : nf++;
    if (free_map)
        for (uint32_t i = 0; i < nn; i++)
            if (fi[i] >= 0) free_map[fi[i]] = i;
    if (nfree_out) *nfree_out = (uint32_t)nf;
    const size_t N6 = 6 * (size_t)nf;
    /* the envelope: a free node's row from the lowest free node it shares a triangle with */
    int* fc = (int*)malloc(sizeof(int) * (size_t)(nf ? nf : 1));
    size_t* roff = (size_t*)malloc(sizeof(size_t) * (size_t)(nf + 1));
    for (int q = 0; q < nf; q++) fc[q] = q;
    for (uint32_t k = 0; k < nm; k++) {
        const int a = fi[f->mem[4 * k]], b = fi[f->mem[4 * k + 1]];
        if (a >= 0 && b >= 0 && !f->mbroken[k]) {
            if (b < fc[a]) fc[a] = b;
            if (a < fc[b]) fc[b] = a;
        }
    }
    for (uint32_t k = 0; k < nt; k++)
        for (int ci = 0; ci < 3; ci++)
            for (int cj = 0; cj < 3; cj++) {
                const int a = fi[f->tn[3 * k + ci]], b = fi[f->tn[3 * k + cj]];
                if (a >= 0 && b >= 0 && b < fc[a]) fc[a] = b;
            }
    size_t tot = 0;
    for (int q = 0; q < nf; q++) roff[q] = tot, tot += (size_t)(q - fc[q] + 1);
    roff[nf] = tot;
    double* L = (double*)calloc(36 * (tot ? tot : 1), sizeof(double));
    double* r = (double*)calloc(N6 ? N6 : 1, sizeof(double));
    const double h2 = (double)theta * h * h, hd = (double)dissipation * h * h;
    const double c = h2 + (double)f->damping * h, cr = hd + (double)f->damping * h;
    /* node -> (triangle, corner) in triangle order */
    uint32_t* ncnt = (uint32_t*)calloc(nn + 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < 3 * nt; k++) ncnt[f->tn[k] + 1]++;
    for (uint32_t i = 0; i < nn; i++) ncnt[i + 1] += ncnt[i];
    uint32_t* nc = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (nt ? nt : 1));
    {
        uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * (nn ? nn : 1));
        memcpy(fill, ncnt, sizeof(uint32_t) * nn);
        for (uint32_t k = 0; k < 3 * nt; k++) nc[fill[f->tn[k]]++] = k;
        free(fill);
    }
    /* node -> (member, end) in member order */
    uint32_t* mcnt = (uint32_t*)calloc(nn + 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < 2 * nm; k++) mcnt[f->mem[4 * (k / 2) + k % 2] + 1]++;
    for (uint32_t i = 0; i < nn; i++) mcnt[i + 1] += mcnt[i];
    uint32_t* mnc = (uint32_t*)malloc(sizeof(uint32_t) * (2 * nm + 1));
    {
        uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * (nn + 1));
        memcpy(fill, mcnt, sizeof(uint32_t) * nn);
        for (uint32_t k = 0; k < 2 * nm; k++) mnc[fill[f->mem[4 * (k / 2) + k % 2]]++] = k;
        free(fill);
    }
    for (uint32_t i = 0; i < nn; i++) {
        if (fi[i] < 0) continue;
        double* D = ENV(fi[i], fi[i]);
        double rd[6] = {0};
        D[0] = D[7] = D[14] = f->mass[i];
        D[21] = D[28] = D[35] = f->inertia[i];
        /* its members' blocks first, then its triangles' (gather_part, :1004-1024) */
        for (uint32_t e = mcnt[i]; e < mcnt[i + 1]; e++) {
            const uint32_t k = mnc[e] / 2, end = mnc[e] % 2;
            if (!mo[k].on) continue;
            mem_block(f, k, &mo[k], (int)end, 1.0, h2, h, D);
            const double crm = hd + (double)f->sec[ORC_FRAME_SECTION_FIELDS * f->mem[4 * k + 2] + ORC_SEC_DAMPING] * h;
            if (crm != 0)
                for (int j = 0; j < 6; j++) rd[j] -= crm * mo[k].kv[6 * end + j];
        }
        for (uint32_t e = ncnt[i]; e < ncnt[i + 1]; e++) {
            const uint32_t k = nc[e] / 3, ci = nc[e] % 3;
            if (!o[k].on) continue;
            tri_block(f, k, &o[k], (int)ci, (int)ci, c, h2, D);
            if (cr != 0)
                for (int j = 0; j < 6; j++) rd[j] -= cr * o[k].kv[6 * ci + j];
        }
        double* rr = &r[6 * (size_t)fi[i]];
        /* h f - (h - step) f_c + J with no contact forces; then the held impulse (f - f_m - f_c) (-(h - step)) (+-0 for a
         * step of one substep, or NaN after a non-finite force) and the prediction of the substeps the step spans,
         * f - f_m - f_c and the moments' m - t_m (frame_fem.cpp:1026-1042) */
        const double hs = (double)h - (double)step;
        float* J = &f->impulse[3 * i];
        rr[0] = h * (double)F[i].x - hs * 0.0f + J[0], rr[1] = h * (double)F[i].y - hs * 0.0f + J[1];
        rr[2] = h * (double)F[i].z - hs * 0.0f + J[2];
        const f3 pr = sub3(sub3(F[i], Fm[i]), F3(0, 0, 0)), tp = sub3(T[i], Tm[i]);
        const f3 Jn = mul3(pr, -(float)hs);
        J[0] = Jn.x, J[1] = Jn.y, J[2] = Jn.z;
        f->pred[3 * i] = pr.x, f->pred[3 * i + 1] = pr.y, f->pred[3 * i + 2] = pr.z;
        f->tpred[3 * i] = tp.x, f->tpred[3 * i + 1] = tp.y, f->tpred[3 * i + 2] = tp.z;
        rr[3] = h * (double)T[i].x, rr[4] = h * (double)T[i].y, rr[5] = h * (double)T[i].z;
        for (int j = 0; j < 6; j++) rr[j] += rd[j];
    }
    free(ncnt), free(nc), free(mcnt), free(mnc);
    /* the lower off-diagonal blocks: the members' (the row node's block: ba when b comes later, else ab), then the
     * triangles' corners (i, j), both free, the row node later */
    for (uint32_t k = 0; k < nm; k++) {
        if (!mo[k].on) continue;
        const int a = fi[f->mem[4 * k]], b = fi[f->mem[4 * k + 1]];
        if (a < 0 || b < 0 || a == b) continue;
        if (b > a) mem_block(f, k, &mo[k], 2, -1.0, h2, h, ENV(b, a));
        else mem_block(f, k, &mo[k], 3, -1.0, h2, h, ENV(a, b));
    }
    for (uint32_t k = 0; k < nt; k++) {
        if (!o[k].on) continue;
        for (int ci = 0; ci < 3; ci++)
            for (int cj = 0; cj < 3; cj++) {
                if (ci == cj) continue;
                const int a = fi[f->tn[3 * k + ci]], b = fi[f->tn[3 * k + cj]];
                if (a < 0 || b < 0 || a < b) continue;
                tri_block(f, k, &o[k], ci, cj, c, h2, ENV(a, b));
            }
    }
    if (Aout) { /* dense, the lower blocks and their mirror */
        memset(Aout, 0, sizeof(double) * N6 * N6);
        for (int q = 0; q < nf; q++)
            for (int cc = fc[q]; cc <= q; cc++) {
                const double* B = ENV(q, cc);
                put_block(Aout, nf, q, cc, B);
                if (cc != q)
                    for (int p2 = 0; p2 < 6; p2++)
                        for (int q2 = 0; q2 < 6; q2++) Aout[(size_t)(6 * cc + q2) * N6 + 6 * q + p2] = B[p2 * 6 + q2];
            }
    }
    if (rhs_out) memcpy(rhs_out, r, sizeof(double) * N6);
    double* dinv = (double*)malloc(sizeof(double) * 6 * (size_t)(nf ? nf : 1));
    int rc = solve_passes(f, mo, fi, nf, fc, roff, tot, L, dinv, r, h, h2, hd);
    /* the factorization kept for the held steps while it succeeded (comp_factored_) */
    free(f->Lk), free(f->dk), free(f->fck), free(f->fik), free(f->roffk);
    f->Lk = L, f->dk = dinv, f->fck = fc, f->fik = fi, f->roffk = roff, f->nfk = nf, f->factored = rc == 0;
    if (rc == 0) {
        if (x_out) memcpy(x_out, r, sizeof(double) * N6);
        /* solve_impl's update (frame_fem.cpp:3000-3027): a clamped node's spin zeroed; a free node's new velocity
         * handed to the integrator as the force dv m / step, unless the solution is not finite (counted; the node's
         * forces then act explicitly); |dv| <= 40 m/s and |w| <= 3000 rad/s (counted) */
        for (uint32_t i = 0; i < nn; i++) {
            if (fi[i] < 0) {
                f->w[3 * i] = f->w[3 * i + 1] = f->w[3 * i + 2] = 0;
                continue;
            }
            const double* x = &r[6 * (size_t)fi[i]];
            f3 dv = F3((float)x[0], (float)x[1], (float)x[2]);
            const f3 dw = F3((float)x[3], (float)x[4], (float)x[5]);
            if (!isfinite(dv.x + dv.y + dv.z + dw.x + dw.y + dw.z)) {
                f->failures++;
                continue;
            }
            const float kMaxDv = 40.0f, kMaxW = 3000.0f;
            float l = len3(dv);
            if (l > kMaxDv) dv = mul3(dv, kMaxDv / l), f->clamps++;
            F[i] = mul3(dv, f->mass[i] / step);
            f3 wv = add3(F3(f->w[3 * i], f->w[3 * i + 1], f->w[3 * i + 2]), dw);
            l = len3(wv);
            if (l > kMaxW) wv = mul3(wv, kMaxW / l), f->clamps++;
            f->w[3 * i] = wv.x, f->w[3 * i + 1] = wv.y, f->w[3 * i + 2] = wv.z;
        }
    } else {
        /* a pivot block not positive definite: the held impulses cleared, the forces act explicitly this once
         * (:2904-2908) */
        f->failures++;
        for (uint32_t i = 0; i < nn; i++)
            if (fi[i] >= 0) f->impulse[3 * i] = f->impulse[3 * i + 1] = f->impulse[3 * i + 2] = 0;
    }
    integrate_sync(f, F, step);
    memset(f->torque, 0, 3 * sizeof(float) * nn); /* (FemFrame::solve_end: the moments from outside used up) */
    free(mo), free(o), free(F), free(T), free(Fm), free(Tm), free(r);
    return rc;
}

/* A held substep (FemFrame::held_begin / held_component / held_end, :2342-2512) of a frame whose implicit step spans
 * several substeps: the forces (gravity, the loads) and moments from outside beyond the last step's prediction, times
 * the substep, through that step's factorization; the prediction's share of the held impulse paid; the solution's
 * velocity changes handed to the integrator as forces (a non-finite one: no force; |dv| <= 40 m/s, |w| <= 3000 rad/s,
 * counted). Without a factorization (the last step failed, or none yet) FemFrame::hold (:2321-2340): the forces held
 * over as impulse. Then integrate_nodes and sync_positions. */
int orc_fem_held(orc_fem* f, const float* loads, float step, const float g[3]) {
    const uint32_t nn = f->nn;
    f3* F = (f3*)malloc(sizeof(f3) * (nn + 1));
    for (uint32_t i = 0; i < nn; i++) {
        F[i] = mul3(F3(g[0], g[1], g[2]), f->mass[i]);
        if (loads) F[i] = add3(F[i], F3(loads[3 * i], loads[3 * i + 1], loads[3 * i + 2]));
    }
    if (!f->factored) {
        for (uint32_t i = 0; i < nn; i++) {
            if (f->inv_mass[i] <= 0) continue;
            const f3 fc = F3(0, 0, 0), J = mul3(sub3(F[i], fc), step);
            f->impulse[3 * i] += J.x, f->impulse[3 * i + 1] += J.y, f->impulse[3 * i + 2] += J.z;
            const f3 v = add3(F3(f->v[3 * i], f->v[3 * i + 1], f->v[3 * i + 2]), mul3(fc, step * f->inv_mass[i]));
            f->v[3 * i] = v.x, f->v[3 * i + 1] = v.y, f->v[3 * i + 2] = v.z;
            F[i] = F3(0, 0, 0);
        }
    } else {
        const int nf = f->nfk;
        const int* fi = f->fik;
        double* r = (double*)calloc(6 * (size_t)(nf ? nf : 1), sizeof(double));
        for (uint32_t i = 0; i < nn; i++) {
            if (fi[i] < 0) continue;
            double* rr = &r[6 * (size_t)fi[i]];
            if (f->inv_mass[i] <= 0) continue;
            const f3 pr = F3(f->pred[3 * i], f->pred[3 * i + 1], f->pred[3 * i + 2]);
            const f3 tp = F3(f->tpred[3 * i], f->tpred[3 * i + 1], f->tpred[3 * i + 2]);
            const f3 d = mul3(sub3(F[i], pr), step);
            const f3 t = mul3(sub3(F3(f->torque[3 * i], f->torque[3 * i + 1], f->torque[3 * i + 2]), tp), step);
            rr[0] = d.x, rr[1] = d.y, rr[2] = d.z, rr[3] = t.x, rr[4] = t.y, rr[5] = t.z;
            const f3 J = mul3(pr, step);
            f->impulse[3 * i] += J.x, f->impulse[3 * i + 1] += J.y, f->impulse[3 * i + 2] += J.z;
        }
        env_subst(nf, f->fck, f->roffk, f->Lk, f->dk, r);
        for (uint32_t i = 0; i < nn; i++) {
            if (fi[i] < 0) continue;
            const double* x = &r[6 * (size_t)fi[i]];
            f3 dv = F3((float)x[0], (float)x[1], (float)x[2]);
            const f3 dw = F3((float)x[3], (float)x[4], (float)x[5]);
            if (!isfinite(dv.x + dv.y + dv.z + dw.x + dw.y + dw.z)) {
                F[i] = F3(0, 0, 0);
                continue;
            }
            const float kMaxDv = 40.0f, kMaxW = 3000.0f;
            float l = len3(dv);
            if (l > kMaxDv) dv = mul3(dv, kMaxDv / l), f->clamps++;
            F[i] = mul3(dv, f->mass[i] / step);
            f3 wv = add3(F3(f->w[3 * i], f->w[3 * i + 1], f->w[3 * i + 2]), dw);
            l = len3(wv);
            if (l > kMaxW) wv = mul3(wv, kMaxW / l), f->clamps++;
            f->w[3 * i] = wv.x, f->w[3 * i + 1] = wv.y, f->w[3 * i + 2] = wv.z;
        }
        free(r);
    }
    integrate_sync(f, F, step);
    memset(f->torque, 0, 3 * sizeof(float) * nn); /* (held_end / hold: the moments from outside used up) */
    free(F);
    return 0;
}
