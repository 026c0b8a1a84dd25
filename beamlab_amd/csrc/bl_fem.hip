// The implicit co-rotational shell step on gfx950: element pass, 6x6-block assembly, supernodal (multifrontal)
// Cholesky over a nested-dissection order with the dense fronts' trailing updates on FP64 matrix cores, the
// substitutions, the update of the state.
//
// Reference semantics (src/phys/frame_fem.cpp): FemFrame::finalize for triangles (:259-276), node_inertia
// (:3075-3086), tri_grad / dkt_b / tri_local_k / tri_build_k (:1482-1579, 1625-1647; FP64, bit-identical),
// tri_kinematics and eval_tris for elastic sections (:1584-1622, 1661-1831), gather_part's triangle blocks and right
// side (:916-1086), the update of velocities, rotations and positions (:3003-3027, 3407-3416). The factorization
// solves the same system as the reference's chol6 / right_solve_lt / sub_abt over its own order (:2193-2317,
// 2899-2943): the results agree to the rounding of FP64 (the residual is measured every step).
//
// Layout in HBM:
//   nodes       xd double[3n] positions, v / w float[3n], q float4 (x y z w), mass, inertia, fixed
//   triangles   corners uint4, rest shape X float[6], rest frames r0 float4[3], stiffness float[162] (membrane 9x9,
//               plate 9x9, as the reference stores it), the element pass's outputs (axes, corner forces and
//               moments, tension stiffness, K v)
//   system      per free node its diagonal block, per pair of free nodes sharing a triangle the lower block
//               (FP64 36 each), the right side
//   fronts      one dense column-major m x m FP64 matrix per supernode (its columns, then the rows below them),
//               holding L11 / L21 after the factorization and the update matrix for its parent in the trailing part
#include "bl_world.h"
#include "bl_fem_api.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#pragma clang fp contract(off)

using blg::f3;

namespace blf {

using blg::mk;

constexpr int kMemDof[3] = {0, 1, 5}, kPlateDof[3] = {2, 3, 4};
__constant__ int c_mem_dof[3] = {0, 1, 5};
__constant__ int c_plate_dof[3] = {2, 3, 4};

// ------------------------------------------------------------------------------------------ float quaternions
// (src/core/math.h:131-226, the same operations in the same order)
struct qf {
    float x, y, z, w;
};
BL_HD qf Q(float x, float y, float z, float w) { qf q; q.x = x, q.y = y, q.z = z, q.w = w; return q; }
BL_HD qf qmul(qf a, qf b) {
    return Q(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
BL_HD qf qconj(qf q) { return Q(-q.x, -q.y, -q.z, q.w); }
BL_HD qf qscale(qf q, float s) { return Q(q.x * s, q.y * s, q.z * s, q.w * s); }
BL_HD qf qnorm(qf q) {
    const float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return l > 1e-12f ? qscale(q, 1.0f / l) : Q(0, 0, 0, 1);
}
BL_HD qf from_cols(f3 c0, f3 c1, f3 c2) {
    const float m00 = c0.x, m11 = c1.y, m22 = c2.z;
    const float tr = m00 + m11 + m22;
    qf q;
    if (tr > 0) {
        const float s = sqrtf(tr + 1.0f) * 2;
        q.w = 0.25f * s, q.x = (c1.z - c2.y) / s, q.y = (c2.x - c0.z) / s, q.z = (c0.y - c1.x) / s;
    } else if (m00 > m11 && m00 > m22) {
        const float s = sqrtf(1.0f + m00 - m11 - m22) * 2;
        q.w = (c1.z - c2.y) / s, q.x = 0.25f * s, q.y = (c1.x + c0.y) / s, q.z = (c2.x + c0.z) / s;
    } else if (m11 > m22) {
        const float s = sqrtf(1.0f + m11 - m00 - m22) * 2;
        q.w = (c2.x - c0.z) / s, q.x = (c1.x + c0.y) / s, q.y = 0.25f * s, q.z = (c2.y + c1.z) / s;
    } else {
        const float s = sqrtf(1.0f + m22 - m00 - m11) * 2;
        q.w = (c0.y - c1.x) / s, q.x = (c2.x + c0.z) / s, q.y = (c2.y + c1.z) / s, q.z = 0.25f * s;
    }
    return qnorm(q);
}
__device__ f3 qlog(qf qin) {
    const qf q = qin.w < 0 ? qscale(qin, -1.0f) : qin;
    const f3 v = mk(q.x, q.y, q.z);
    const float s = blg::length(v);
    if (s < 1e-8f) return blg::mul(v, 2.0f);
    const float angle = 2.0f * blg::bl_atan2f(s, q.w);
    return blg::mul(v, angle / s);
}
__device__ qf qexp(f3 rv) {
    const float a = blg::length(rv);
    if (a < 1e-8f) return qnorm(Q(rv.x * 0.5f, rv.y * 0.5f, rv.z * 0.5f, 1.0f));
    const f3 d = blg::divs(rv, a);
    const float l = blg::length(d);
    const f3 axis = l > 1e-12f ? blg::mul(d, 1.0f / l) : mk(0, 0, 0);
    const float hh = a * 0.5f, s = blg::bl_sinf(hh);
    return Q(axis.x * s, axis.y * s, axis.z * s, blg::bl_cosf(hh));
}

// ------------------------------------------------------------------------------------------ the element, FP64
__device__ void dkt_b(const double x[3], const double y[3], double xi, double eta, double B[3][9]) {
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

__device__ __forceinline__ void tri_grad(const float* X, double b[3], double c[3], double& A2) {
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        b[i] = (double)X[2 * j + 1] - X[2 * k + 1];
        c[i] = (double)X[2 * k] - X[2 * j];
    }
    A2 = ((double)X[2] - X[0]) * ((double)X[5] - X[1]) - ((double)X[4] - X[0]) * ((double)X[3] - X[1]);
}

struct Section {
    float E, nu, rho, t, drill, damping, yield;
};

// tri_local_k (src/phys/frame_fem.cpp:1524-1579): K (18 x 18, row-major) of the rest shape X
__device__ void tri_local_k(const float* X, const Section& s, double* K) {
    for (int i = 0; i < 324; i++) K[i] = 0;
    double b[3], c[3], A2;
    tri_grad(X, b, c, A2);
    if (!(A2 > 0)) return;
    const double A = 0.5 * A2, nu = s.nu;
    const double Dm = (double)s.E * s.t / (1 - nu * nu);
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
            K[(6 * (p / 2) + p % 2) * 18 + 6 * (q / 2) + q % 2] += A * Dm * v;
        }
    const double x[3] = {X[0], X[2], X[4]}, y[3] = {X[1], X[3], X[5]};
    const float Dbf = s.E * s.t * s.t * s.t / (12.0f * (1.0f - s.nu * s.nu)); // (ShellSection::D(), in float)
    const double Db = Dbf;
    const double gauss[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
    for (int g = 0; g < 3; g++) {
        double B[3][9];
        dkt_b(x, y, gauss[g][0], gauss[g][1], B);
        for (int p = 0; p < 9; p++)
            for (int q = 0; q < 9; q++) {
                double v = 0;
                for (int r = 0; r < 3; r++)
                    for (int t = 0; t < 3; t++) v += B[r][p] * D3[r][t] * B[t][q];
                K[(6 * (p / 3) + 2 + p % 3) * 18 + 6 * (q / 3) + 2 + q % 3] += A2 / 6.0 * Db * v;
            }
    }
    const double G = (double)s.E / (2 * (1 + nu)), kd = (double)s.drill * G * s.t * A;
    double gu[3], gv[3];
    for (int i = 0; i < 3; i++) gu[i] = -c[i] / (2 * A2), gv[i] = b[i] / (2 * A2);
    for (int i = 0; i < 3; i++) {
        K[(6 * i + 5) * 18 + 6 * i + 5] += kd;
        for (int j = 0; j < 3; j++) {
            K[(6 * i + 5) * 18 + 6 * j] -= kd * gu[j], K[(6 * j) * 18 + 6 * i + 5] -= kd * gu[j];
            K[(6 * i + 5) * 18 + 6 * j + 1] -= kd * gv[j], K[(6 * j + 1) * 18 + 6 * i + 5] -= kd * gv[j];
            K[(6 * i) * 18 + 6 * j] += 3 * kd * gu[i] * gu[j], K[(6 * i) * 18 + 6 * j + 1] += 3 * kd * gu[i] * gv[j];
            K[(6 * i + 1) * 18 + 6 * j] += 3 * kd * gv[i] * gu[j], K[(6 * i + 1) * 18 + 6 * j + 1] += 3 * kd * gv[i] * gv[j];
        }
    }
}

struct FemArgs {
    uint32_t nn, nt, nf;
    Section sec;
    double* xd;
    float *v, *w;
    float4* q;
    const float *mass, *inertia;
    float* pf;           // the body's float positions (Node::p), 3 per node
    const float* invm;   // Node::inv_mass: 0 for a pinned node
    float* fnode;        // the node forces the step gathered (3 per node); the update's forces for the integrator
    float* imp;          // gather_part's held impulse J (3 per node)
    unsigned long long* cnt; // solve failures (not finite), clamps, NaN velocity resets
    float* torq;         // FemFrame::torque: moments from outside for this step (3 per node; the update clears them)
    // members: (a, b, -, broken), rest length, rest frames in the end nodes' frames, elastic stiffness (ka, kt, cy[3],
    // cz[3]), twist weight, damping; the element pass's outputs: the ends' force and moments (fa, ta, tb), the chord and
    // the tension's geometric stiffness (e1, N / L), the 12 x 12 matrix's blocks (aa, bb, ba, ab) and K u
    uint32_t nm;
    const uint4* mem;
    const float* mL0;
    float4* mqa;
    float4* mqb;
    const float* mst;
    const float* mtw;
    const float* mdamp;
    uint8_t* mon;
    float* mfo;
    float4* me1;
    double* mW;
    double* mkv;
    // members' plasticity: the sections (the 17 floats of bl_frame_section; mem.z the member's), the plastic state
    // (16 per member: up, tp, pa, pb, ba, bb, bt, damage, dmg_a, dmg_b, util, N), the tangent's kinematic part (17
    // per member: e1, e2, e3, L, ngeo, ma, mb, N, T) and its bits (yield | unload << 8 | elastic_hold << 16), whether
    // this step starts every yielding member from its elastic tangent (the passes-again budget spent)
    const float* msec;
    float* mps;
    float* mtan;
    uint32_t* mbits;
    int elastic_step;
    uint32_t* nyield;   // [0] members yielding at the step's start, [1] members a pass unloads
    const uint8_t* fixed;
    const uint4* tri;   // corners, 0
    float* X;           // 6 per triangle: the rest shape (the plastic one)
    float4* r0;         // 3 per triangle
    float* Kf;          // 162 per triangle
    double* Kd;         // 324 per triangle (kept for export)
    // triangles' plasticity: the authored rest shape (6), the corners' plastic rest rotations (9), the plastic
    // stretch and utilisation, the bending's curvature rows at the three points (54)
    const float* X0;
    float* th0;
    float* tdmg;
    float* tutil;
    float* Bc;
    // element pass outputs
    float* te;          // 9 per triangle: e1 e2 e3
    float* tfo;         // 9: the corners' forces
    float* tto;         // 9: the corners' moments
    float* tg;          // 9: tension stiffness g[i][j]
    double* tkv;        // 18: K v, world axes
    uint8_t* ton;
    // node -> (triangle, corner) incidences in triangle order
    const uint32_t* nc_off;
    const uint32_t* nc;
    // free nodes: order index and its front, position in the front
    const int* fidx;     // per node: its free index, -1 pinned
    const uint32_t* ford; // per free index: its position in the order
    // the system
    double* Dsys;        // 36 per free index
    double* rhs;         // 6 per free index
    double* Osys;        // 36 per pair
    const uint4* pairs;  // (row node, column node, first contribution, end)
    const uint4* pcon;   // (triangle, corner of the row node, corner of the column node, 0)
    uint32_t npairs;
    const float* loads;  // 3 per node
    float g[3];
    float h, theta, dissipation;
    // the substep the step starts with (h: the step spans h / step substeps; the update and the integration are the
    // substep's), the prediction of the smooth forces and moments for the held substeps (3 per node each)
    float step;
    float* pred;
    float* tpred;
    // a pass again: the held impulses the first pass used (3 per node), the pass, the members' components and the
    // components a pass unloaded
    float* jsave;
    int pass;
    const uint32_t* mcomp;
    uint32_t* again;
};

// tri_build_k's stored parts (src/phys/frame_fem.cpp:1628-1646) from the FP64 stiffness K of the rest shape: the float
// membrane and plate 9 x 9 and, for a section that yields, the bending's curvature rows at the three points
__device__ void tri_store_k(const FemArgs& a, uint32_t t, const double* K) {
    float* P = a.Kf + (size_t)t * 162;
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 9; j++) {
            const int im = 6 * (i / 3) + kMemDof[i % 3], jm = 6 * (j / 3) + kMemDof[j % 3];
            const int ip = 6 * (i / 3) + kPlateDof[i % 3], jp = 6 * (j / 3) + kPlateDof[j % 3];
            P[i * 9 + j] = (float)K[(im < jm ? im : jm) * 18 + (im < jm ? jm : im)];
            P[81 + i * 9 + j] = (float)K[(ip < jp ? ip : jp) * 18 + (ip < jp ? jp : ip)];
        }
    if (!(a.sec.yield > 0)) return;
    const float* X = a.X + 6 * (size_t)t;
    const double x[3] = {X[0], X[2], X[4]}, y[3] = {X[1], X[3], X[5]};
    const double gauss[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
    float* Bc = a.Bc + 54 * (size_t)t;
    for (int g = 0; g < 3; g++) {
        double B[3][9];
        dkt_b(x, y, gauss[g][0], gauss[g][1], B);
        for (int r = 0; r < 3; r++)
            for (int i = 0; i < 3; i++) Bc[g * 18 + r * 6 + 2 * i] = (float)B[r][3 * i + 1], Bc[g * 18 + r * 6 + 2 * i + 1] = (float)B[r][3 * i + 2];
    }
}

// FemFrame::tri_build_k on the device: the element stiffness (FP64) and its float copy
__global__ void k_tri_k(FemArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nt) return;
    double* K = a.Kd + (size_t)t * 324;
    tri_local_k(a.X + 6 * (size_t)t, a.sec, K);
    tri_store_k(a, t, K);
}

// ------------------------------------------------------------------------------------------ members
// (src/core/math.h:82-85, 136-154, 203-212, 227-233: quat::rotate, axis_angle, quat_from_to, slerp; the same operations
// in the same order, sinf / cosf / acosf glibc's)
__device__ f3 qrot(qf q, f3 p) {
    const f3 u = mk(q.x, q.y, q.z);
    const f3 c1 = mk(u.y * p.z - u.z * p.y, u.z * p.x - u.x * p.z, u.x * p.y - u.y * p.x);
    const f3 t = blg::mul(c1, 2.0f);
    const f3 c2 = mk(u.y * t.z - u.z * t.y, u.z * t.x - u.x * t.z, u.x * t.y - u.y * t.x);
    return blg::add(blg::add(p, blg::mul(t, q.w)), c2);
}
__device__ f3 nrm3(f3 a) {
    const float l = blg::length(a);
    return l > 1e-12f ? blg::mul(a, 1.0f / l) : mk(0, 0, 0);
}
__device__ qf qadd(qf a, qf b) { return Q(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ qf quat_from_to(f3 a, f3 b) {
    const float d = blg::dot(a, b);
    if (d < -0.999999f) {
        const f3 t = fabsf(a.x) < 0.7f ? mk(1, 0, 0) : mk(0, 1, 0);
        const f3 ax = nrm3(nrm3(mk(a.y * t.z - a.z * t.y, a.z * t.x - a.x * t.z, a.x * t.y - a.y * t.x)));
        const float hh = 3.14159265358979323846f * 0.5f, sn = blg::bl_sinf(hh);
        return Q(ax.x * sn, ax.y * sn, ax.z * sn, blg::bl_cosf(hh));
    }
    const f3 c = mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    return qnorm(Q(c.x, c.y, c.z, 1.0f + d));
}
__device__ qf qslerp(qf a, qf b, float t) {
    float d = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    if (d < 0) b = qscale(b, -1.0f), d = -d;
    if (d > 0.9995f) return qnorm(qadd(a, qscale(qadd(b, qscale(a, -1.0f)), t)));
    const float th = blg::bl_acosf(d), sn = blg::bl_sinf(th);
    return qadd(qscale(a, blg::bl_sinf((1 - t) * th) / sn), qscale(b, blg::bl_sinf(t * th) / sn));
}

// K[r][c] of element_matrix (src/phys/frame_fem.cpp:2157-2186) at its place among the stored blocks: aa, bb, ba, ab
__device__ __forceinline__ uint32_t kidx(int r, int c) {
    return r < 6 ? (c < 6 ? r * 6 + c : 108 + r * 6 + (c - 6)) : (c < 6 ? 72 + (r - 6) * 6 + c : 36 + (r - 6) * 6 + (c - 6));
}

// condense (src/phys/frame_fem.cpp:1414-1418): a pinned end's bending stiffness condensed out
__device__ void mem_condense(float* c, bool pa, bool pb) {
    if (pa && pb) c[0] = c[1] = c[2] = 0;
    else if (pa) c[2] = c[0] > 0 ? c[2] - c[1] * c[1] / c[0] : 0, c[0] = c[1] = 0;
    else if (pb) c[0] = c[2] > 0 ? c[0] - c[1] * c[1] / c[2] : 0, c[1] = c[2] = 0;
}

// member_world_k (:1834-1850): element_matrix (:2157-2186) at the member's effective tangent (effective, :2147-2155:
// a flowing hinge turns like a pin, a flowing bar or twist keeps a hundredth of its stiffness unless the step unloads
// it), written to its blocks aa, bb, ba, ab as the reference's loop sets K, the lower triangle from the upper one;
// and K u at the ends' velocities and spins
__device__ void member_world_k(const FemArgs& a, uint32_t m) {
    const float* st = a.mst + 8 * (size_t)m;
    const float* tn = a.mtan + 17 * (size_t)m;
    const uint32_t bits = a.mbits[m];
    const uint32_t y = (bits & 0xffu) & ~((bits >> 8) & 0xffu);
    float ka = st[0], kt = st[1];
    float cy[3] = {st[2], st[3], st[4]}, cz[3] = {st[5], st[6], st[7]};
    if (y & 3) mem_condense(cy, (y & 1) != 0, (y & 2) != 0), mem_condense(cz, (y & 1) != 0, (y & 2) != 0);
    if (y & 4) ka *= 0.01f;
    if (y & 8) kt *= 0.01f;
    double* W = a.mW + 144 * (size_t)m;
    const double iL = 1.0 / tn[9];
    const double E1[3] = {tn[0], tn[1], tn[2]}, E2[3] = {tn[3], tn[4], tn[5]}, E3[3] = {tn[6], tn[7], tn[8]};
    const double g[3] = {E3[0] * iL, E3[1] * iL, E3[2] * iL}, h[3] = {E2[0] * iL, E2[1] * iL, E2[2] * iL};
    const double c0 = cy[0], c1 = cy[1], c2 = cy[2], d0 = cz[0], d1 = cz[1], d2 = cz[2];
    const double Syy = c0 + 2 * c1 + c2, Szz = d0 + 2 * d1 + d2, ay = c0 + c1, by = c1 + c2, az = d0 + d1, bz = d1 + d2;
    const double kad = ka, ktd = kt;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double uu = kad * E1[i] * E1[j] + Syy * g[i] * g[j] + Szz * h[i] * h[j];
            const double up = -ay * g[i] * E2[j] + az * h[i] * E3[j];
            const double ub = -by * g[i] * E2[j] + bz * h[i] * E3[j];
            const double tt = ktd * E1[i] * E1[j];
            const double pa = tt + c0 * E2[i] * E2[j] + d0 * E3[i] * E3[j];
            const double pb = -tt + c1 * E2[i] * E2[j] + d1 * E3[i] * E3[j];
            const double qb2 = tt + c2 * E2[i] * E2[j] + d2 * E3[i] * E3[j];
            W[kidx(i, j)] = uu, W[kidx(i, 3 + j)] = up, W[kidx(i, 6 + j)] = -uu, W[kidx(i, 9 + j)] = ub;
            W[kidx(3 + i, 3 + j)] = pa, W[kidx(3 + i, 9 + j)] = pb;
            W[kidx(6 + i, 6 + j)] = uu, W[kidx(6 + i, 9 + j)] = -ub;
            W[kidx(9 + i, 9 + j)] = qb2;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) W[kidx(3 + i, 6 + j)] = -W[kidx(j, 3 + i)];
    for (int r = 0; r < 12; r++)
        for (int c = r + 1; c < 12; c++) W[kidx(c, r)] = W[kidx(r, c)];
    const uint4 me = a.mem[m];
    const uint32_t na = me.x, nb = me.y;
    const double u[12] = {a.v[3 * na], a.v[3 * na + 1], a.v[3 * na + 2], a.w[3 * na], a.w[3 * na + 1], a.w[3 * na + 2],
                          a.v[3 * nb], a.v[3 * nb + 1], a.v[3 * nb + 2], a.w[3 * nb], a.w[3 * nb + 1], a.w[3 * nb + 2]};
    double* kv = a.mkv + 12 * (size_t)m;
    for (int r = 0; r < 12; r++) {
        double s2 = 0;
        for (int c = 0; c < 12; c++) s2 += W[kidx(r, c)] * u[c];
        kv[r] = s2;
    }
}

// (std::max / std::min as the reference calls them: the first argument unless the second compares greater / less)
__device__ __forceinline__ float smax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float smin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float len2f(float x, float y) { return sqrtf(x * x + y * y); }

// the end moments at the end rotations less the hinges' plastic rotations (:1946-1950)
__device__ void mem_moments(f3 ra, f3 rb, const float* st, const float* ps, float* Ma, float* Mb) {
    const float dax = ra.y - ps[2], day = ra.z - ps[3], dbx = rb.y - ps[4], dby = rb.z - ps[5];
    Ma[0] = st[2] * dax + st[3] * dbx, Ma[1] = st[5] * day + st[6] * dby;
    Mb[0] = st[3] * dax + st[4] * dbx, Mb[1] = st[6] * day + st[7] * dby;
}

// a hinge's radial return (:1960-1972): the yield moment hardening with the hinge's largest rotation
__device__ bool mem_yield_end(const float* sc, const float* M, float cyy, float czz, float* p, const float* base, float* dmg,
                              float* flow) {
    const float Mp = sc[13] * (1.0f + sc[15] * smin(1.0f, *dmg));
    const float l = len2f(M[0], M[1]);
    if (l <= Mp) return false;
    const float r = 1.0f - Mp / l;
    const float dp0 = cyy > 0 ? M[0] * r / cyy : 0.0f, dp1 = czz > 0 ? M[1] * r / czz : 0.0f;
    p[0] += dp0, p[1] += dp1;
    *flow += len2f(dp0, dp1);
    *dmg = smax(*dmg, len2f(p[0] + base[0], p[1] + base[1]));
    return true;
}

// eval_forces (src/phys/frame_fem.cpp:1896-2069, the body's allow_break off, no joint damping), one thread per member:
// kinematics (:1384-1405) - the chord, the co-rotated frame (the ends' frames, their mean turned onto the chord), the
// end rotations against it -; the axial force (yielding in tension at Np, in compression at the lower of Np and the
// Euler load), the torque (at Tp), the end moments (each hinge's radial return, twice); the ends' force and moments;
// the plastic rotations past 0.02 rad baked into the rest frames; the tangent (its yield bits where it flows this
// step; elastic while held or with the budget spent) and member_world_k. The plastic state is the thread's own.
__global__ void k_members(FemArgs a) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.nm) return;
    a.mon[m] = 0;
    const uint4 me = a.mem[m];
    if (me.w) return; // (broken: no length)
    const uint32_t na = me.x, nb = me.y;
    const double* xa = a.xd + 3 * (size_t)na;
    const double* xb = a.xd + 3 * (size_t)nb;
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(L > 1e-6) || !__builtin_isfinite(L)) return;
    const float Lf = (float)L;
    const float L0 = a.mL0[m];
    const double dL = L - (double)L0;
    const f3 e1 = mk((float)(dx / L), (float)(dy / L), (float)(dz / L));
    const float4 qna = a.q[na], qnb = a.q[nb];
    float4 qa = a.mqa[m], qb = a.mqb[m];
    const qf Ea = qnorm(qmul(Q(qna.x, qna.y, qna.z, qna.w), Q(qa.x, qa.y, qa.z, qa.w)));
    const qf Eb = qnorm(qmul(Q(qnb.x, qnb.y, qnb.z, qnb.w), Q(qb.x, qb.y, qb.z, qb.w)));
    const float t = a.mtw[m];
    const qf Em = t <= 0.0f ? Ea : t >= 1.0f ? Eb : qslerp(Ea, Eb, t);
    const qf E = qnorm(qmul(quat_from_to(nrm3(qrot(Em, mk(1, 0, 0))), e1), Em));
    const qf Ec = qconj(E);
    const f3 ra = qlog(qmul(Ec, Ea)), rb = qlog(qmul(Ec, Eb));
    const f3 e2 = qrot(E, mk(0, 1, 0)), e3 = qrot(E, mk(0, 0, 1));
    const float* st = a.mst + 8 * (size_t)m;
    const float* sc = a.msec + 17 * (size_t)me.z;
    const float ka = st[0], kt = st[1];
    float ps[16];
    for (int k = 0; k < 16; k++) ps[k] = a.mps[16 * (size_t)m + k];
    const bool plastic = sc[12] > 0;
    const float delta = (float)dL;
    float N = ka * (delta - ps[0]);
    float Nt = sc[12], Nc = sc[12];
    float flow_ax = 0, flow_t = 0;
    if (plastic) {
        const float Le = 0.7f * L0, kPi = 3.14159265358979323846f;
        const float eul = kPi * kPi * sc[0] * smin(sc[4], sc[5]) / (Le * Le);
        Nc = smin(sc[12], eul);
        if (N > Nt || N < -Nc) {
            const float lim = N > 0 ? Nt : -Nc;
            const float dup = (N - lim) / ka;
            ps[0] += dup;
            flow_ax = fabsf(dup) / L0;
            ps[11] = smax(ps[11], ps[0] / L0);
            N = lim;
        }
    }
    const float tw = rb.x - ra.x;
    float T = kt * (tw - ps[1]);
    if (plastic && sc[14] > 0 && kt > 0 && fabsf(T) > sc[14]) {
        const float lim = T > 0 ? sc[14] : -sc[14];
        const float dtp = (T - lim) / kt;
        ps[1] += dtp;
        flow_t = fabsf(dtp);
        ps[11] = smax(ps[11], fabsf(ps[1] + ps[10]) * sc[16] / L0);
        T = lim;
    }
    float Ma[2], Mb[2];
    mem_moments(ra, rb, st, ps, Ma, Mb);
    float flow_a = 0, flow_b = 0;
    if (plastic && sc[13] > 0)
        for (int it = 0; it < 2; it++) {
            if (mem_yield_end(sc, Ma, st[2], st[5], &ps[2], &ps[6], &ps[12], &flow_a)) mem_moments(ra, rb, st, ps, Ma, Mb);
            if (mem_yield_end(sc, Mb, st[4], st[7], &ps[4], &ps[8], &ps[13], &flow_b)) mem_moments(ra, rb, st, ps, Ma, Mb);
        }
    float util = 0;
    if (sc[12] > 0) util = smax(util, N > 0 ? N / Nt : -N / Nc);
    if (sc[13] > 0) {
        const float ua = len2f(Ma[0], Ma[1]) / (1.0f + sc[15] * smin(1.0f, ps[12]));
        const float ub = len2f(Mb[0], Mb[1]) / (1.0f + sc[15] * smin(1.0f, ps[13]));
        util = smax(util, smax(ua, ub) / sc[13]);
    }
    if (sc[14] > 0) util = smax(util, fabsf(T) / sc[14]);
    ps[14] = util, ps[15] = N;
    const float Sy = Ma[0] + Mb[0], Sz = Ma[1] + Mb[1];
    const f3 fa = blg::sub(blg::add(blg::mul(e1, N), blg::mul(e3, Sy / Lf)), blg::mul(e2, Sz / Lf));
    const f3 ta = blg::sub(blg::sub(blg::mul(e1, T), blg::mul(e2, Ma[0])), blg::mul(e3, Ma[1]));
    const f3 tb = blg::sub(blg::sub(blg::mul(e1, -T), blg::mul(e2, Mb[0])), blg::mul(e3, Mb[1]));
    float* fo = a.mfo + 9 * (size_t)m;
    fo[0] = fa.x, fo[1] = fa.y, fo[2] = fa.z, fo[3] = ta.x, fo[4] = ta.y, fo[5] = ta.z, fo[6] = tb.x, fo[7] = tb.y, fo[8] = tb.z;
    // the plastic rotations into the rest frames (:2037-2045)
    const float kBake = 0.02f;
    if (ps[2] * ps[2] + ps[3] * ps[3] > kBake * kBake) {
        const qf q2 = qnorm(qmul(Q(qa.x, qa.y, qa.z, qa.w), qexp(mk(0, -ps[2], -ps[3]))));
        qa = make_float4(q2.x, q2.y, q2.z, q2.w);
        ps[6] += ps[2], ps[7] += ps[3], ps[2] = ps[3] = 0;
    }
    if (ps[4] * ps[4] + ps[5] * ps[5] > kBake * kBake) {
        const qf q2 = qnorm(qmul(Q(qb.x, qb.y, qb.z, qb.w), qexp(mk(0, -ps[4], -ps[5]))));
        qb = make_float4(q2.x, q2.y, q2.z, q2.w);
        ps[8] += ps[4], ps[9] += ps[5], ps[4] = ps[5] = 0;
    }
    if (fabsf(ps[1]) > kBake) {
        const qf q1 = qnorm(qmul(Q(qa.x, qa.y, qa.z, qa.w), qexp(mk(0.5f * ps[1], 0, 0))));
        const qf q2 = qnorm(qmul(Q(qb.x, qb.y, qb.z, qb.w), qexp(mk(-0.5f * ps[1], 0, 0))));
        qa = make_float4(q1.x, q1.y, q1.z, q1.w), qb = make_float4(q2.x, q2.y, q2.z, q2.w);
        ps[10] += ps[1], ps[1] = 0;
    }
    if (plastic) {
        a.mqa[m] = qa, a.mqb[m] = qb;
        for (int k = 0; k < 16; k++) a.mps[16 * (size_t)m + k] = ps[k];
    } else {
        a.mps[16 * (size_t)m + 14] = util, a.mps[16 * (size_t)m + 15] = N;
    }
    // the tangent (:2051-2068)
    const float ngeo = N > 0 ? N / Lf : 0.0f;
    a.me1[m] = make_float4(e1.x, e1.y, e1.z, ngeo);
    float* tn = a.mtan + 17 * (size_t)m;
    tn[0] = e1.x, tn[1] = e1.y, tn[2] = e1.z, tn[3] = e2.x, tn[4] = e2.y, tn[5] = e2.z, tn[6] = e3.x, tn[7] = e3.y, tn[8] = e3.z;
    tn[9] = Lf, tn[10] = ngeo, tn[11] = Ma[0], tn[12] = Ma[1], tn[13] = Mb[0], tn[14] = Mb[1], tn[15] = N, tn[16] = T;
    const float kFlow = 1e-4f;
    const uint32_t yb = (flow_a > kFlow ? 1u : 0u) | (flow_b > kFlow ? 2u : 0u) | (flow_ax > 0.1f * kFlow ? 4u : 0u) | (flow_t > kFlow ? 8u : 0u);
    uint32_t hold = (a.mbits[m] >> 16) & 0xffu, unload = 0;
    if (hold > 0) hold--, unload = yb;
    if (a.elastic_step) unload = yb;
    a.mbits[m] = yb | unload << 8 | hold << 16;
    if (yb & ~unload) atomicAdd(&a.nyield[0], 1u);
    member_world_k(a, m);
    a.mon[m] = 1;
}

// solve_impl's check after a pass (src/phys/frame_fem.cpp:2965-2988): a yielding member's deformation rates at the
// pass's solution (the start's velocities plus the solution's change) - a hinge turning against its moment, the bar
// stretching against its force, the twist against its torque - unload it: its elastic tangent from now (held 16
// steps), its matrix and K u again, its component marked for a pass again
__global__ void k_mem_repass(FemArgs a, const double* x) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.nm || !a.mon[m]) return;
    const uint32_t bits = a.mbits[m];
    const uint32_t y = (bits & 0xffu) & ~((bits >> 8) & 0xffu);
    if (!y) return;
    const uint4 me = a.mem[m];
    f3 vv[2], ww[2];
    for (int e = 0; e < 2; e++) {
        const uint32_t n = e ? me.y : me.x;
        const int fi = a.fidx[n];
        if (fi < 0) {
            vv[e] = ww[e] = mk(0, 0, 0);
            continue;
        }
        const double* xs = x + 6 * (size_t)fi;
        vv[e] = blg::add(mk(a.v[3 * n], a.v[3 * n + 1], a.v[3 * n + 2]), mk((float)xs[0], (float)xs[1], (float)xs[2]));
        ww[e] = blg::add(mk(a.w[3 * n], a.w[3 * n + 1], a.w[3 * n + 2]), mk((float)xs[3], (float)xs[4], (float)xs[5]));
    }
    const float* tn = a.mtan + 17 * (size_t)m;
    const f3 e1 = mk(tn[0], tn[1], tn[2]), e2 = mk(tn[3], tn[4], tn[5]), e3 = mk(tn[6], tn[7], tn[8]);
    const float iL = 1.0f / tn[9];
    const f3 dv = blg::sub(vv[1], vv[0]);
    const float rax = blg::dot(e3, dv) * iL + blg::dot(e2, ww[0]), ray = -blg::dot(e2, dv) * iL + blg::dot(e3, ww[0]);
    const float rbx = blg::dot(e3, dv) * iL + blg::dot(e2, ww[1]), rby = -blg::dot(e2, dv) * iL + blg::dot(e3, ww[1]);
    uint32_t u = 0;
    if ((y & 1) && rax * tn[11] + ray * tn[12] < 0) u |= 1;
    if ((y & 2) && rbx * tn[13] + rby * tn[14] < 0) u |= 2;
    if ((y & 4) && blg::dot(e1, dv) * tn[15] < 0) u |= 4;
    if ((y & 8) && blg::dot(e1, blg::sub(ww[1], ww[0])) * tn[16] < 0) u |= 8;
    if (!u) return;
    a.mbits[m] = (bits & 0xffu) | ((((bits >> 8) & 0xffu) | u) << 8) | 16u << 16;
    member_world_k(a, m);
    a.again[a.mcomp[m]] = 1u;
    atomicAdd(&a.nyield[1], 1u);
}

// gather_part's member_block (src/phys/frame_fem.cpp:926-941): block blk (0 aa, 1 bb, 2 ba, 3 ab) times c, the tension's
// geometric stiffness with sign gs
__device__ void member_block(const FemArgs& a, uint32_t m, int blk, double gs, double h2, double* D) {
    const double c = h2 + (double)a.mdamp[m] * a.h;
    const double* B = a.mW + 144 * (size_t)m + 36 * blk;
    for (int x = 0; x < 36; x++) D[x] += c * B[x];
    const float4 e = a.me1[m];
    const double gg = h2 * e.w;
    if (gg != 0) {
        const double ee[3] = {e.x, e.y, e.z};
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) D[r * 6 + q] += gs * (gg * ((r == q ? 1.0 : 0.0) - ee[r] * ee[q]));
    }
}

__device__ double von_mises(double m0, double m1, double m2) {
    const double v = m0 * m0 - m0 * m1 + m1 * m1 + 3 * m2 * m2;
    return sqrt(0.0 < v ? v : 0.0);
}

__device__ void tri_membrane(const double* b3, const double* c3, double A2, const double* d, double nu, double Dm,
                             double* eps, double* sig) {
    eps[0] = eps[1] = eps[2] = 0;
    for (int i = 0; i < 3; i++) {
        eps[0] += b3[i] * d[6 * i] / A2;
        eps[1] += c3[i] * d[6 * i + 1] / A2;
        eps[2] += (c3[i] * d[6 * i] + b3[i] * d[6 * i + 1]) / A2;
    }
    sig[0] = Dm * (eps[0] + nu * eps[1]), sig[1] = Dm * (nu * eps[0] + eps[1]), sig[2] = Dm * 0.5 * (1 - nu) * eps[2];
}

// eval_tris's plasticity (src/phys/frame_fem.cpp:1690-1742, the body's allow_break off): the membrane flows radially
// back to the yield stress, the rest shape following the stretch (the plastic stretch against the authored shape: the
// damage; the stiffness rebuilt, tri_build_k, into this triangle's own FP64 and float copies); the bending's moments
// at the three points against the plate's plastic moment, past it the corners' plastic rest rotations following;
// plastic rotations past 0.02 rad baked into the rest frames. Returns the utilisation.
__device__ float tri_plastic(const FemArgs& a, uint32_t t, double* d, double* b3, double* c3, double* A2, double* eps,
                             double* sig) {
    float* X = a.X + 6 * (size_t)t;
    const double nu = a.sec.nu, Dm = (double)a.sec.E / (1 - nu * nu);
    const double seq = von_mises(sig[0], sig[1], sig[2]);
    float util = (float)(seq / a.sec.yield);
    if (seq > a.sec.yield) {
        const double r = 1.0 - a.sec.yield / seq;
        for (int i = 0; i < 3; i++) {
            X[2 * i] += (float)(r * d[6 * i]), X[2 * i + 1] += (float)(r * d[6 * i + 1]);
            d[6 * i] *= 1 - r, d[6 * i + 1] *= 1 - r;
        }
        const float mx = (X[0] + X[2] + X[4]) / 3.0f, my = (X[1] + X[3] + X[5]) / 3.0f;
        for (int i = 0; i < 3; i++) X[2 * i] -= mx, X[2 * i + 1] -= my;
        const float* X0 = a.X0 + 6 * (size_t)t;
        const double a0x = X0[2] - X0[0], a0y = X0[3] - X0[1], b0x = X0[4] - X0[0], b0y = X0[5] - X0[1];
        const double ax = X[2] - X[0], ay = X[3] - X[1], bx = X[4] - X[0], by = X[5] - X[1];
        const double det = a0x * b0y - b0x * a0y;
        if (fabs(det) > 1e-12) {
            const double i00 = b0y / det, i01 = -b0x / det, i10 = -a0y / det, i11 = a0x / det;
            const double F00 = ax * i00 + bx * i10, F01 = ax * i01 + bx * i11, F10 = ay * i00 + by * i10, F11 = ay * i01 + by * i11;
            const double C00 = F00 * F00 + F10 * F10, C01 = F00 * F01 + F10 * F11, C11 = F01 * F01 + F11 * F11;
            const double lmax = 0.5 * (C00 + C11) + sqrt(0.25 * (C00 - C11) * (C00 - C11) + C01 * C01);
            a.tdmg[t] = smax(a.tdmg[t], (float)(sqrt(lmax) - 1.0));
        }
        double* K = a.Kd + (size_t)t * 324;
        tri_local_k(X, a.sec, K);
        tri_store_k(a, t, K);
        tri_grad(X, b3, c3, *A2);
        tri_membrane(b3, c3, *A2, d, nu, Dm, eps, sig);
    }
    const float* Bc = a.Bc + 54 * (size_t)t;
    const double D = a.sec.E * a.sec.t * a.sec.t * a.sec.t / (12.0f * (1.0f - a.sec.nu * a.sec.nu)); // (float)
    const double Mp = 0.25 * a.sec.yield * (double)a.sec.t * a.sec.t;
    double meq = 0;
    for (int g = 0; g < 3; g++) {
        double kap[3] = {0, 0, 0};
        for (int r = 0; r < 3; r++)
            for (int i = 0; i < 3; i++)
                kap[r] += Bc[g * 18 + r * 6 + 2 * i] * d[6 * i + 3] + Bc[g * 18 + r * 6 + 2 * i + 1] * d[6 * i + 4];
        const double vm = von_mises(D * (kap[0] + nu * kap[1]), D * (nu * kap[0] + kap[1]), D * 0.5 * (1 - nu) * kap[2]);
        meq = meq < vm ? vm : meq;
    }
    util = smax(util, (float)(meq / Mp));
    float* th0 = a.th0 + 9 * (size_t)t;
    if (meq > Mp) {
        const double r = 1.0 - Mp / meq;
        for (int i = 0; i < 3; i++) {
            th0[3 * i] += (float)(r * d[6 * i + 3]), th0[3 * i + 1] += (float)(r * d[6 * i + 4]);
            d[6 * i + 3] *= 1 - r, d[6 * i + 4] *= 1 - r;
        }
        a.tdmg[t] = smax(a.tdmg[t], 1e-5f);
    }
    for (int i = 0; i < 3; i++) {
        float* t0 = th0 + 3 * i;
        if (t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2] > 0.02f * 0.02f) {
            const float4 r0 = a.r0[3 * (size_t)t + i];
            const qf q = qnorm(qmul(Q(r0.x, r0.y, r0.z, r0.w), qexp(mk(-t0[0], -t0[1], -t0[2]))));
            a.r0[3 * (size_t)t + i] = make_float4(q.x, q.y, q.z, q.w);
            t0[0] = t0[1] = t0[2] = 0;
        }
    }
    return util;
}

// tri_kinematics + eval_tris: the co-rotated frame, the corners' forces and moments, the tangent; PLASTIC: a section
// that yields (the corners' plastic rest rotations off the measured ones, then tri_plastic)
template <bool PLASTIC>
__global__ void k_tris(FemArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nt) return;
    a.ton[t] = 0;
    const uint4 tr = a.tri[t];
    const uint32_t n[3] = {tr.x, tr.y, tr.z};
    const float* X = a.X + 6 * (size_t)t;
    double x[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) x[i][j] = a.xd[n[i] * 3 + j];
    const double av[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
    const double cv[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
    double nv[3] = {av[1] * cv[2] - av[2] * cv[1], av[2] * cv[0] - av[0] * cv[2], av[0] * cv[1] - av[1] * cv[0]};
    const double nl = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]), al = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
    if (!(nl > 1e-12) || !(al > 1e-9) || !isfinite(nl)) return;
    for (int j = 0; j < 3; j++) nv[j] /= nl;
    const double t1[3] = {av[0] / al, av[1] / al, av[2] / al};
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
    double e1d[3], e2d[3];
    for (int j = 0; j < 3; j++) e1d[j] = cp * t1[j] + sp * t2[j], e2d[j] = -sp * t1[j] + cp * t2[j];
    const f3 e1 = mk((float)e1d[0], (float)e1d[1], (float)e1d[2]), e2 = mk((float)e2d[0], (float)e2d[1], (float)e2d[2]);
    const f3 e3 = mk((float)nv[0], (float)nv[1], (float)nv[2]);
    const qf Ec = qconj(from_cols(e1, e2, e3));
    double d[18];
    for (int i = 0; i < 3; i++) {
        d[6 * i] = cp * p[i][0] + sp * p[i][1] - X[2 * i];
        d[6 * i + 1] = -sp * p[i][0] + cp * p[i][1] - X[2 * i + 1];
        d[6 * i + 2] = 0;
        const float4 qn = a.q[n[i]], r0 = a.r0[3 * (size_t)t + i];
        f3 th = qlog(qmul(Ec, qnorm(qmul(Q(qn.x, qn.y, qn.z, qn.w), Q(r0.x, r0.y, r0.z, r0.w)))));
        if (PLASTIC) {
            const float* t0 = a.th0 + 9 * (size_t)t + 3 * i;
            th = blg::sub(th, mk(t0[0], t0[1], t0[2]));
        }
        d[6 * i + 3] = th.x, d[6 * i + 4] = th.y, d[6 * i + 5] = th.z;
    }
    double b3[3], c3[3], A2;
    tri_grad(X, b3, c3, A2);
    if (!(A2 > 0)) return;
    const double nu = a.sec.nu, Dm = (double)a.sec.E / (1 - nu * nu);
    double eps[3], sig[3];
    tri_membrane(b3, c3, A2, d, nu, Dm, eps, sig);
    if (PLASTIC) a.tutil[t] = tri_plastic(a, t, d, b3, c3, &A2, eps, sig);
    const float* Km = a.Kf + (size_t)t * 162;
    const float* Kp = Km + 81;
    const double E[3][3] = {{e1.x, e2.x, e3.x}, {e1.y, e2.y, e3.y}, {e1.z, e2.z, e3.z}};
    double vl[18];
    for (int i = 0; i < 3; i++) {
        const f3 vv = mk(a.v[3 * n[i]], a.v[3 * n[i] + 1], a.v[3 * n[i] + 2]);
        const f3 om = mk(a.w[3 * n[i]], a.w[3 * n[i] + 1], a.w[3 * n[i] + 2]);
        for (int k = 0; k < 3; k++) {
            vl[6 * i + k] = E[0][k] * vv.x + E[1][k] * vv.y + E[2][k] * vv.z;
            vl[6 * i + 3 + k] = E[0][k] * om.x + E[1][k] * om.y + E[2][k] * om.z;
        }
    }
    double fo[18], fl[18];
    {
        double dm[9], dp[9], vm[9], vp[9];
        for (int k = 0; k < 9; k++) {
            const int ci = 6 * (k / 3);
            dm[k] = d[ci + c_mem_dof[k % 3]], dp[k] = d[ci + c_plate_dof[k % 3]];
            vm[k] = vl[ci + c_mem_dof[k % 3]], vp[k] = vl[ci + c_plate_dof[k % 3]];
        }
        double fm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, fp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double um[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, up[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < 9; c++)
            for (int k = 0; k < 9; k++) {
                const double km = (double)Km[c * 9 + k], kp = (double)Kp[c * 9 + k];
                fm[k] += km * dm[c], um[k] += km * vm[c];
                fp[k] += kp * dp[c], up[k] += kp * vp[c];
            }
        for (int k = 0; k < 9; k++) {
            const int ci = 6 * (k / 3);
            fo[ci + c_mem_dof[k % 3]] = fm[k], fl[ci + c_mem_dof[k % 3]] = um[k];
            fo[ci + c_plate_dof[k % 3]] = fp[k], fl[ci + c_plate_dof[k % 3]] = up[k];
        }
    }
    float* te = a.te + 9 * (size_t)t;
    te[0] = e1.x, te[1] = e1.y, te[2] = e1.z, te[3] = e2.x, te[4] = e2.y, te[5] = e2.z, te[6] = e3.x, te[7] = e3.y, te[8] = e3.z;
    for (int i = 0; i < 3; i++) {
        const f3 fw = blg::add(blg::add(blg::mul(e1, (float)fo[6 * i]), blg::mul(e2, (float)fo[6 * i + 1])), blg::mul(e3, (float)fo[6 * i + 2]));
        const f3 tw = blg::add(blg::add(blg::mul(e1, (float)fo[6 * i + 3]), blg::mul(e2, (float)fo[6 * i + 4])), blg::mul(e3, (float)fo[6 * i + 5]));
        a.tfo[9 * (size_t)t + 3 * i] = -fw.x, a.tfo[9 * (size_t)t + 3 * i + 1] = -fw.y, a.tfo[9 * (size_t)t + 3 * i + 2] = -fw.z;
        a.tto[9 * (size_t)t + 3 * i] = -tw.x, a.tto[9 * (size_t)t + 3 * i + 1] = -tw.y, a.tto[9 * (size_t)t + 3 * i + 2] = -tw.z;
    }
    {
        const double sx = sig[0] * a.sec.t, sy = sig[1] * a.sec.t, sxy = sig[2] * a.sec.t;
        const double tr2 = 0.5 * (sx + sy), df = sqrt(0.25 * (sx - sy) * (sx - sy) + sxy * sxy);
        const double l1 = fmax(0.0, tr2 + df), l2 = fmax(0.0, tr2 - df);
        double vx = sxy, vy = l1 - sx;
        if (fabs(vx) + fabs(vy) < 1e-12 * (fabs(sx) + fabs(sy) + 1)) vx = sx >= sy ? 1 : 0, vy = sx >= sy ? 0 : 1;
        const double vlen = sqrt(vx * vx + vy * vy);
        vx /= vlen, vy /= vlen;
        const double S00 = l1 * vx * vx + l2 * vy * vy, S01 = (l1 - l2) * vx * vy, S11 = l1 * vy * vy + l2 * vx * vx;
        const double A = 0.5 * A2;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double gi0 = b3[i] / A2, gi1 = c3[i] / A2, gj0 = b3[j] / A2, gj1 = c3[j] / A2;
                a.tg[9 * (size_t)t + 3 * i + j] = (float)(A * (gi0 * (S00 * gj0 + S01 * gj1) + gi1 * (S01 * gj0 + S11 * gj1)));
            }
    }
    for (int i = 0; i < 3; i++)
        for (int p2 = 0; p2 < 3; p2++) {
            a.tkv[18 * (size_t)t + 6 * i + p2] = E[p2][0] * fl[6 * i] + E[p2][1] * fl[6 * i + 1] + E[p2][2] * fl[6 * i + 2];
            a.tkv[18 * (size_t)t + 6 * i + 3 + p2] = E[p2][0] * fl[6 * i + 3] + E[p2][1] * fl[6 * i + 4] + E[p2][2] * fl[6 * i + 5];
        }
    a.ton[t] = 1;
}

// gather_part's triangle block (src/phys/frame_fem.cpp:948-985): corners (i, j) in world axes, times c, the
// membrane tension's stiffness on the translations' diagonal
__device__ void tri_block(const FemArgs& a, uint32_t t, int i, int j, double c, double h2, double* D) {
    const float* te = a.te + 9 * (size_t)t;
    const double E[3][3] = {{te[0], te[3], te[6]}, {te[1], te[4], te[7]}, {te[2], te[5], te[8]}};
    const float* Km = a.Kf + (size_t)t * 162;
    const float* Kp = Km + 81;
    for (int ra = 0; ra < 2; ra++)
        for (int cb = 0; cb < 2; cb++) {
            double T[3][3];
            for (int p = 0; p < 3; p++) {
                const double E0 = E[p][0], E1 = E[p][1], E2 = E[p][2];
#define M_(s2, t2) ((double)Km[(3 * i + (s2)) * 9 + 3 * j + (t2)])
#define P_(s2, t2) ((double)Kp[(3 * i + (s2)) * 9 + 3 * j + (t2)])
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
#undef M_
#undef P_
            }
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) D[(3 * ra + p) * 6 + 3 * cb + q] += c * (T[p][0] * E[q][0] + T[p][1] * E[q][1] + T[p][2] * E[q][2]);
        }
    const double gg = h2 * a.tg[9 * (size_t)t + 3 * i + j];
    if (gg != 0)
        for (int p = 0; p < 3; p++) D[p * 7] += gg;
}

// Per node: its force (gravity, the load, the triangles' corners in triangle order), and for a free node its
// diagonal block and its right side (gather_part, parts 0 and 1).
__global__ void k_nodes(FemArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nn) return;
    const int fi = a.fidx[i];
    f3 F = blg::mul(mk(a.g[0], a.g[1], a.g[2]), a.mass[i]);
    if (a.loads) F = blg::add(F, mk(a.loads[3 * i], a.loads[3 * i + 1], a.loads[3 * i + 2]));
    f3 T = mk(a.torq[3 * i], a.torq[3 * i + 1], a.torq[3 * i + 2]);
    if (fi < 0) { // (a pinned node: its forces only, for the record)
        for (uint32_t k = a.nc_off[i]; k < a.nc_off[i + 1]; k++) {
            const uint32_t e = a.nc[k];
            if (e >> 31) {
                const uint32_t m = (e & 0x7fffffffu) >> 1, end = e & 1;
                if (!a.mon[m]) continue;
                const float* fo = a.mfo + 9 * (size_t)m;
                F = end ? blg::sub(F, mk(fo[0], fo[1], fo[2])) : blg::add(F, mk(fo[0], fo[1], fo[2]));
                continue;
            }
            const uint32_t t = e / 3, corner = e % 3;
            if (a.ton[t]) F = blg::add(F, mk(a.tfo[9 * t + 3 * corner], a.tfo[9 * t + 3 * corner + 1], a.tfo[9 * t + 3 * corner + 2]));
        }
        a.fnode[3 * i] = F.x, a.fnode[3 * i + 1] = F.y, a.fnode[3 * i + 2] = F.z;
        return;
    }
    const double h = a.h;
    const double h2 = (double)a.theta * a.h * a.h, hd = (double)a.dissipation * a.h * a.h;
    const double c = h2 + (double)a.sec.damping * h, cr = hd + (double)a.sec.damping * h;
    double D[36];
    for (int k = 0; k < 36; k++) D[k] = 0;
    double rd[6] = {0, 0, 0, 0, 0, 0};
    D[0] = D[7] = D[14] = a.mass[i];
    D[21] = D[28] = D[35] = a.inertia[i];
    // (the elements' share alone beside the node's sums: member_f, member_t_, src/phys/frame_fem.cpp:2080-2100)
    f3 Fm = mk(0, 0, 0), Tm = mk(0, 0, 0);
    // its members' ends first (member order), then its triangles' corners (gather_lists, src/phys/frame_fem.cpp:862-898)
    for (uint32_t k = a.nc_off[i]; k < a.nc_off[i + 1]; k++) {
        const uint32_t e = a.nc[k];
        if (e >> 31) {
            const uint32_t m = (e & 0x7fffffffu) >> 1, end = e & 1;
            if (!a.mon[m]) continue;
            const float* fo = a.mfo + 9 * (size_t)m;
            F = end ? blg::sub(F, mk(fo[0], fo[1], fo[2])) : blg::add(F, mk(fo[0], fo[1], fo[2]));
            Fm = end ? blg::sub(Fm, mk(fo[0], fo[1], fo[2])) : blg::add(Fm, mk(fo[0], fo[1], fo[2]));
            T = blg::add(T, end ? mk(fo[6], fo[7], fo[8]) : mk(fo[3], fo[4], fo[5]));
            Tm = blg::add(Tm, end ? mk(fo[6], fo[7], fo[8]) : mk(fo[3], fo[4], fo[5]));
            member_block(a, m, (int)end, 1.0, h2, D);
            const double crm = hd + (double)a.mdamp[m] * h;
            if (crm != 0)
                for (int j = 0; j < 6; j++) rd[j] -= crm * a.mkv[12 * (size_t)m + 6 * end + j];
            continue;
        }
        const uint32_t t = e / 3, corner = e % 3;
        if (!a.ton[t]) continue;
        const f3 tf = mk(a.tfo[9 * t + 3 * corner], a.tfo[9 * t + 3 * corner + 1], a.tfo[9 * t + 3 * corner + 2]);
        const f3 tt = mk(a.tto[9 * t + 3 * corner], a.tto[9 * t + 3 * corner + 1], a.tto[9 * t + 3 * corner + 2]);
        F = blg::add(F, tf), Fm = blg::add(Fm, tf);
        T = blg::add(T, tt), Tm = blg::add(Tm, tt);
        tri_block(a, t, (int)corner, (int)corner, c, h2, D);
        if (cr != 0)
            for (int j = 0; j < 6; j++) rd[j] -= cr * a.tkv[18 * (size_t)t + 6 * corner + j];
    }
    double* Dd = a.Dsys + 36 * (size_t)fi;
    for (int k = 0; k < 36; k++) Dd[k] = D[k];
    double* r = a.rhs + 6 * (size_t)fi;
    // h f - (h - step) f_c + J with no contact forces; then the held impulse (f - f_m - f_c) (-(h - step)) (+-0 for a
    // step of one substep, or NaN after a non-finite force, carried into the next step) and the prediction of the
    // substeps the step spans, f - f_m - f_c and the moments' m - t_m (src/phys/frame_fem.cpp:1026-1042)
    // (a pass again takes the impulses its first pass took: held aside in jsave)
    const double hs = (double)a.h - (double)a.step;
    float* J = a.pass ? a.jsave + 3 * i : a.imp + 3 * i;
    r[0] = h * (double)F.x - hs * 0.0f + J[0], r[1] = h * (double)F.y - hs * 0.0f + J[1];
    r[2] = h * (double)F.z - hs * 0.0f + J[2];
    if (!a.pass) {
        if (a.jsave) a.jsave[3 * i] = J[0], a.jsave[3 * i + 1] = J[1], a.jsave[3 * i + 2] = J[2];
        const f3 pr = blg::sub(blg::sub(F, Fm), mk(0, 0, 0)), tp = blg::sub(T, Tm);
        const f3 Jn = blg::mul(pr, -(float)hs);
        J[0] = Jn.x, J[1] = Jn.y, J[2] = Jn.z;
        a.pred[3 * i] = pr.x, a.pred[3 * i + 1] = pr.y, a.pred[3 * i + 2] = pr.z;
        a.tpred[3 * i] = tp.x, a.tpred[3 * i + 1] = tp.y, a.tpred[3 * i + 2] = tp.z;
    }
    r[3] = h * (double)T.x, r[4] = h * (double)T.y, r[5] = h * (double)T.z;
    for (int j = 0; j < 6; j++) r[j] += rd[j];
    a.fnode[3 * i] = F.x, a.fnode[3 * i + 1] = F.y, a.fnode[3 * i + 2] = F.z;
}

// Per pair of free nodes sharing a triangle: the lower block (later node's row, earlier node's column), its
// triangles in triangle order.
__global__ void k_pairs(FemArgs a) {
    const uint32_t pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= a.npairs) return;
    const uint4 pr = a.pairs[pi];
    const double h2 = (double)a.theta * a.h * a.h;
    const double c = h2 + (double)a.sec.damping * a.h;
    double O[36];
    for (int k = 0; k < 36; k++) O[k] = 0;
    for (uint32_t k = pr.z; k < pr.w; k++) { // (the members' blocks first, then the triangles')
        const uint4 cn = a.pcon[k];
        if (cn.x >> 31) {
            const uint32_t m = cn.x & 0x7fffffffu;
            if (a.mon[m]) member_block(a, m, (int)cn.y, -1.0, h2, O);
        } else if (a.ton[cn.x]) {
            tri_block(a, cn.x, (int)cn.y, (int)cn.z, c, h2, O);
        }
    }
    double* Od = a.Osys + 36 * (size_t)pi;
    for (int k = 0; k < 36; k++) Od[k] = O[k];
}

// ------------------------------------------------------------------------------------------ the multifrontal factor
struct Front {
    uint64_t off;      // first double of its m x m matrix (column-major)
    uint32_t m, c;     // dimension, pivot columns (6 per own node)
    uint32_t own0, ns; // its own nodes: order positions own0 .. own0 + ns
    uint32_t row0, nr; // its row nodes: rows[row0 .. row0 + nr] (order positions, ascending)
    int32_t parent;
    uint32_t level;
    uint64_t voff;     // its m-vector in the substitutions' buffer: the right side, then y (own rows) and the parent's
                       // update (the rows below); in the backward pass, x of its own rows
    uint32_t fb;       // its first flag: one per 64 own columns
    uint32_t pad;
    uint64_t off_l;    // first double of its factor columns (m x c, column-major) in L
};

struct FacArgs {
    double* F;             // the fronts
    double* L;             // the factor's columns of every front (m x c, column-major)
    const Front* fr;
    int* fail;             // bit 1: a pivot not positive; bit 2: a substitution's wait timed out
};

// Scatter of the system into the fronts: entry = (source, 0 diagonal / 1 pair, front, row position, column position)
struct Scat {
    uint32_t src, kind, front, rpos, cpos;
};

// The fronts' lower triangles cleared (every kernel of the factorization reads and writes the lower triangle only):
// task = (front, 16 columns from j0), the rows j .. m of each column.
__global__ __launch_bounds__(256) void k_zero_lower(const Front* fr, const uint2* tasks, uint32_t n, double* F) {
    if (blockIdx.x >= n) return;
    const uint2 t = tasks[blockIdx.x];
    const Front x = fr[t.x];
    const uint32_t j1 = t.y + 16 < x.m ? t.y + 16 : x.m;
    for (uint32_t j = t.y; j < j1; j++) {
        double* col = F + x.off + (uint64_t)j * x.m;
        for (uint32_t r = j + threadIdx.x; r < x.m; r += 256) col[r] = 0.0;
    }
}

__global__ void k_clear_int(int* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

__global__ void k_scatter(FemArgs a, FacArgs f, const Scat* sc, uint32_t n) {
    const uint32_t e = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, l = threadIdx.x & 63;
    if (e >= n || l >= 36) return;
    const Scat s = sc[e];
    const Front fr = f.fr[s.front];
    const double* B = s.kind == 0 ? a.Dsys + 36 * (size_t)s.src : a.Osys + 36 * (size_t)s.src;
    const int p = l / 6, q = l % 6;
    const uint64_t r = 6 * (uint64_t)s.rpos + p, c = 6 * (uint64_t)s.cpos + q;
    if (s.kind == 0 && p < q) return; // (the lower triangle of a diagonal block)
    f.F[fr.off + c * fr.m + r] = B[l];
}

// Extend-add: a child's update matrix (the trailing (6 nr)^2 of its front) into its parent's front. Task = (child,
// row block of 64, column block of 64); map: the child's row nodes -> positions in the parent's front.
struct EaTask {
    uint32_t child, r0, c0, pad;
};
__global__ void k_extend_add(FacArgs f, const EaTask* tk, uint32_t n, const uint32_t* cmap) {
    const uint32_t ti = blockIdx.x;
    if (ti >= n) return;
    const EaTask t = tk[ti];
    const Front ch = f.fr[t.child];
    const Front pa = f.fr[ch.parent];
    const uint32_t M = 6 * ch.nr;
    for (uint32_t k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
        const uint32_t r = t.r0 + k % 64, c = t.c0 + k / 64;
        if (r >= M || c >= M || r < c) continue;
        const double u = f.F[ch.off + (uint64_t)(ch.c + c) * ch.m + ch.c + r];
        const uint32_t pr = 6 * cmap[ch.row0 + r / 6] + r % 6, pc = 6 * cmap[ch.row0 + c / 6] + c % 6;
        f.F[pa.off + (uint64_t)pc * pa.m + pr] += u;
    }
}

// The panel steps of a level, with a look-ahead of one: launch s holds, for every front of the level, the panel
// factorization of columns k = 32 s .. k + w (w <= 32) and the trailing update of the panel before it (columns
// k - 32 ..). A panel task (front, rows r0 .. r0 + 64 of the panel, r0 >= k)
//   1. takes the previous panel's update off its diagonal tile and its rows (from L, the factor's stored columns:
//      that update is the one part of the trailing update of the previous panel not left to the update tasks),
//   2. factors the w x w diagonal tile in one wave (the reference's chol6 order: pivot, the column times 1 / pivot,
//      the rank-1 update; a row per lane in registers, the finished column through LDS),
//   3. solves its rows below the tile against it (X := X L^-T, a row per thread),
//   4. stores its rows (the first task also the tile) as the factor's columns in L, a storage apart from the fronts;
//      L's diagonal holds 1 / L_jj, all the substitutions need of it.
// Every panel task of a front computes the same tile from the same inputs with the same code, bit for bit. An update
// task (front, 64 x 64 tile of columns >= lo2: past the next panel) takes P_r P_c^T of the previous panel off its
// tile on the FP64 matrix cores. Panel tasks write L's columns k.. and read L's columns k - 32.., update tasks read
// L's columns k - 32.. and write the front's columns >= lo2: no task of a launch reads what another writes.
// (The factorization and the substitutions are not bit-exact with the reference - its order of the updates differs -
// so their multiply-adds are fused.)
constexpr int kT = 32;
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t panel_w(const Front& fr, uint32_t k) { return fr.c - k < (uint32_t)kT ? fr.c - k : (uint32_t)kT; }

__device__ __forceinline__ double readlane_d(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l) << 32) |
                                          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l));
}

__device__ void panel_task(const FacArgs& f, uint2 t, uint32_t k, double* sh) {
    double (*Td)[kT + 1] = (double (*)[kT + 1])sh;               // 32 x 33: the diagonal tile
    double (*X)[kT + 1] = (double (*)[kT + 1])(sh + 32 * 33);    // 64 x 33: the task's rows
    double (*Lr)[kT + 1] = (double (*)[kT + 1])(sh + 96 * 33);   // 64 x 33: their previous-panel factor
    double (*Ld)[kT + 1] = (double (*)[kT + 1])(sh + 160 * 33);  // 32 x 33: the tile rows' previous-panel factor
    double* inv = sh + 192 * 33;
    const Front fr = f.fr[t.x];
    const uint32_t m = fr.m, w = panel_w(fr, k), lo = k + w, r0 = t.y;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double* A = f.F + fr.off;
    double* Lf = f.L + fr.off_l;
    const bool prev = k > 0;
    const uint32_t kp = k - kT;
    // 1. loads: the tile (lower), the rows below it, and their previous-panel factor rows; every load issued before
    // the first store to LDS (consecutive threads on consecutive rows of a column: the fronts and L are column-major)
    {
        double vt[kT * kT / 256], vl[kT * kT / 256], vx[64 * kT / 256], vr[64 * kT / 256];
#pragma unroll
        for (int q = 0; q < kT * kT / 256; q++) {
            const uint32_t e = tid + 256 * q, ii = e % kT, jj = e / kT;
            vt[q] = (ii < w && jj <= ii) ? A[(uint64_t)(k + jj) * m + k + ii] : 0.0;
            vl[q] = (prev && ii < w) ? Lf[(uint64_t)(kp + jj) * m + k + ii] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 64 * kT / 256; q++) {
            const uint32_t e = tid + 256 * q, ii = e % 64, jj = e / 64, r = r0 + ii;
            const bool below = r < m && r >= lo;
            vx[q] = (below && jj < w) ? A[(uint64_t)(k + jj) * m + r] : 0.0;
            vr[q] = (prev && below) ? Lf[(uint64_t)(kp + jj) * m + r] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kT * kT / 256; q++) {
            const uint32_t e = tid + 256 * q, ii = e % kT, jj = e / kT;
            Td[ii][jj] = vt[q], Ld[ii][jj] = vl[q];
        }
#pragma unroll
        for (int q = 0; q < 64 * kT / 256; q++) {
            const uint32_t e = tid + 256 * q, ii = e % 64, jj = e / 64;
            X[ii][jj] = vx[q], Lr[ii][jj] = vr[q];
        }
    }
    __syncthreads();
    if (prev) {
        // the tile's lower triangle and the rows, less the previous panel's part: C = U Ld^T with U = [Ld; Lr]
        // (96 x 32) on the matrix cores, 16 x 16 output blocks (6 x 2), three a wave
        const uint32_t li = lane & 15, lk = lane >> 4;
#pragma unroll
        for (int b3 = 0; b3 < 3; b3++) {
            const uint32_t blk = wv + 4 * b3, i0 = (blk >> 1) * 16, j0 = (blk & 1) * 16;
            const uint32_t iu = i0 + li;
            const double* urow = iu < (uint32_t)kT ? Ld[iu] : Lr[iu - kT];
            const double* vrow = Ld[j0 + li];
            dbl4 acc = (dbl4){0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < kT; q += 4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(urow[q + lk], vrow[q + lk], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t i = i0 + lk + 4 * r, j = j0 + li;
                if (j >= w) continue;
                if (i < (uint32_t)kT) {
                    if (i < w && j <= i) Td[i][j] -= acc[r];
                } else {
                    X[i - kT][j] -= acc[r];
                }
            }
        }
        __syncthreads();
    }
    // 2-3. the tile's Cholesky (wave 0: a row per lane in registers, each finished column broadcast lane to lane) and
    // the rows' solve X := X L^-T (wave 1: a row per lane, in LDS, 4 partial sums a dot product) in step: column j
    // of L and 1 / L_jj go to LDS, a barrier, then wave 1 takes column j of its rows (it needs row j of L, final once
    // column j is) while wave 0 factors column j + 1.
    {
        const uint32_t ri = lane & 31;
        const bool xrow = wv == 1 && r0 + lane < m && r0 + lane >= lo;
        double a[kT];
        bool bad = false;
        if (wv == 0)
#pragma unroll
            for (int j = 0; j < kT; j++) a[j] = Td[ri][j];
        double* x = X[lane];
#pragma unroll
        for (int j = 0; j < kT; j++) {
            if (j < (int)w) {
                if (wv == 0) {
                    const double piv = readlane_d(a[j], j);
                    bad |= !(piv > 0);
                    // 1 / sqrt(piv) from the hardware estimate and a Newton step (within an ulp or so: the factor
                    // carries its own rounding anyway), sqrt(piv) = piv / sqrt(piv)
                    const double pv = piv > 0 ? piv : 1.0, y0 = __builtin_amdgcn_rsq(pv);
                    const double id = y0 * (1.5 - (0.5 * pv) * (y0 * y0)), d = pv * id;
                    a[j] = (int)ri == j ? d : a[j] * id;
                    if (lane == 0) inv[j] = id;
                    // (blocked in halves: the first 16 pivots update columns < 16 only; the lower right 16 x 16
                    // block takes L21 L21^T at once after the 16th)
#pragma unroll
                    for (int c = j + 1; c < (j < 16 ? 16 : kT); c++) a[c] -= a[j] * readlane_d(a[j], c);
                    if (lane < kT) Td[ri][j] = j <= (int)ri ? a[j] : 0.0;
                    if (j == 15) {
                        // L21 (rows 16.., columns < 16) through LDS (the previous panel's rows, done with), its
                        // product with itself on the matrix cores, the product back into rows 16..
                        double (*S)[17] = (double (*)[17])&Ld[0][0];
                        double (*Dm)[17] = (double (*)[17])(&Ld[0][0] + 16 * 17);
                        if (lane >= 16 && lane < 32)
#pragma unroll
                            for (int q = 0; q < 16; q++) S[lane - 16][q] = a[q];
                        __builtin_amdgcn_s_waitcnt(0xc07f);
                        __builtin_amdgcn_wave_barrier();
                        const uint32_t li = lane & 15, lk = lane >> 4;
                        dbl4 acc = (dbl4){0, 0, 0, 0};
#pragma unroll
                        for (int kk = 0; kk < 16; kk += 4) {
                            const double v = S[li][kk + lk];
                            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, acc, 0, 0, 0);
                        }
#pragma unroll
                        for (int r = 0; r < 4; r++) Dm[lk + 4 * r][li] = acc[r];
                        __builtin_amdgcn_s_waitcnt(0xc07f);
                        __builtin_amdgcn_wave_barrier();
                        if (ri >= 16)
#pragma unroll
                            for (int c = 16; c < kT; c++) a[c] -= Dm[ri - 16][c - 16];
                    }
                }
                __syncthreads();
                if (xrow) {
                    double s0 = x[j], s1 = 0, s2 = 0, s3 = 0;
                    int mm = 0;
#pragma unroll
                    for (; mm + 4 <= j; mm += 4)
                        s0 -= Td[j][mm] * x[mm], s1 -= Td[j][mm + 1] * x[mm + 1], s2 -= Td[j][mm + 2] * x[mm + 2],
                            s3 -= Td[j][mm + 3] * x[mm + 3];
#pragma unroll
                    for (; mm < j; mm++) s0 -= Td[j][mm] * x[mm];
                    x[j] = ((s0 + s1) + (s2 + s3)) * inv[j];
                }
            }
        }
        if (wv == 0 && bad && lane == 0 && r0 == k) atomicOr(f.fail, 1);
    }
    __syncthreads();
    // 4. store
    for (uint32_t e = tid; e < 64 * kT; e += 256) {
        const uint32_t ii = e % 64, jj = e / 64, r = r0 + ii;
        if (jj >= w || r >= m || r < k) continue;
        if (r >= lo) Lf[(uint64_t)(k + jj) * m + r] = X[ii][jj];
        else if (jj < r - k) Lf[(uint64_t)(k + jj) * m + r] = Td[r - k][jj];
        else if (jj == r - k) Lf[(uint64_t)(k + jj) * m + r] = inv[jj]; // (the diagonal stored as 1 / L_jj)
    }
}

// C -= P_r P_c^T on the matrix cores: a 32 x 32 quarter a wave, 2 x 2 blocks of v_mfma_f64_16x16x4f64. The
// instruction's lanes (measured on gfx950): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; item r of its
// result is D[l / 16 + 4 r][l % 16]. The product is formed transposed (A from P_c, B from P_r: D = (P_r P_c^T)^T), so
// that a lane's items lie down a column of C and 16 lanes cover 16 consecutive rows (C is column-major).
__device__ void update_task(const FacArgs& f, uint4 t, uint32_t kp, double* sh) {
    double (*Pa)[kT + 1] = (double (*)[kT + 1])sh;
    double (*Pb)[kT + 1] = (double (*)[kT + 1])(sh + 64 * 33);
    const Front fr = f.fr[t.x];
    const uint32_t m = fr.m, w = panel_w(fr, kp);
    const uint32_t r0 = t.y, c0 = t.z;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double* Lf = f.L + fr.off_l;
    for (uint32_t e = tid; e < 64 * kT; e += 256) {
        const uint32_t rr = e % 64, kk = e / 64;
        const bool okk = kk < w;
        Pa[rr][kk] = (okk && r0 + rr < m) ? Lf[(uint64_t)(kp + kk) * m + r0 + rr] : 0.0;
        Pb[rr][kk] = (okk && c0 + rr < m) ? Lf[(uint64_t)(kp + kk) * m + c0 + rr] : 0.0;
    }
    __syncthreads();
    const uint32_t wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    dbl4 acc[2][2];
    for (int a0 = 0; a0 < 2; a0++)
        for (int b0 = 0; b0 < 2; b0++) acc[a0][b0] = (dbl4){0, 0, 0, 0};
    const uint32_t li = lane & 15, lk = lane >> 4;
    for (uint32_t kk = 0; kk < (uint32_t)kT; kk += 4) {
        double av[2], bv[2];
        for (int a0 = 0; a0 < 2; a0++) av[a0] = Pa[wr + a0 * 16 + li][kk + lk];
        for (int b0 = 0; b0 < 2; b0++) bv[b0] = Pb[wc + b0 * 16 + li][kk + lk];
        for (int a0 = 0; a0 < 2; a0++)
            for (int b0 = 0; b0 < 2; b0++) acc[a0][b0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b0], av[a0], acc[a0][b0], 0, 0, 0);
    }
    double* C = f.F + fr.off;
    for (int a0 = 0; a0 < 2; a0++)
        for (int b0 = 0; b0 < 2; b0++)
            for (int r = 0; r < 4; r++) {
                const uint32_t row = r0 + wr + a0 * 16 + li, col = c0 + wc + b0 * 16 + lk + 4 * r;
                if (row < m && col < m && row >= col) C[(uint64_t)col * m + row] -= acc[a0][b0][r];
            }
}

// launch s of a level: panel tasks of panel k = 32 s (first), then the update tasks of panel k - 32 (at least 3 waves a
// SIMD: the LDS already holds a CU to 3 workgroups, and the registers must not hold it to fewer)
__global__ __launch_bounds__(256, 3) void k_fstep(FacArgs f, const uint2* pt, uint32_t np, const uint4* ut, uint32_t nu, uint32_t k) {
#pragma clang fp contract(fast)
    __shared__ double sh[192 * 33 + 2 * kT];
    if (blockIdx.x < np) panel_task(f, pt[blockIdx.x], k, sh);
    else if (blockIdx.x < np + nu) update_task(f, ut[blockIdx.x - np], k - kT, sh);
}

// ------------------------------------------------------------------------------------------ substitutions
// Per level, the forward pass (children first) gathers each front's right side (its own rows, then its children's
// updates in child order), then solves it in blocks of 64 rows: a workgroup per block, which takes the solved blocks
// of the front's own columns in order as their flags come up (y of block jb is published with a device-scope flag),
// solves its diagonal block in one wave and publishes its y - or, below the own columns, leaves the parent's update.
// The backward pass (parents first) solves L11^T x = y - L21^T x_rows in blocks of 64 columns, the last block first.
// Every sum runs in a fixed order: the result does not depend on the timing. Waits are bounded (spin_limit polls,
// or another workgroup's error) and raise the error word, so every grid drains.
struct SolArgs {
    const double* F;           // the factor's columns (L)
    const Front* fr;
    double* V;                 // per front m: right side / y / parent's update (forward)
    double* X;                 // per front m: x of the own rows (backward)
    const uint32_t* rows;      // row nodes (order positions) of every front
    const uint32_t* children;  // children of every front (CSR)
    const uint32_t* ch_off;
    const uint32_t* cmap;      // per child row node: its position in the parent's front
    const uint32_t* ord2f;     // order position -> free index
    const double* b;           // right side, 6 per free index
    double* y;                 // 6 per free index
    double* x;                 // 6 per free index
    uint32_t* fflag;           // per front per 64 own columns: forward y published (= epoch)
    uint32_t* bflag;           // ... backward x published
    const uint32_t* epoch;
    int* err;
    uint32_t spin_limit;
    const int* skip;           // (or null) nonzero: every task returns at once (a converged iteration's solve)
};

// lane l's value of v (l uniform across the wave)
__device__ __forceinline__ double readlane_u(double v, uint32_t l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, (int)l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), (int)l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double ld_dev(const double* p) {
    return __builtin_bit_cast(double, __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_dev(double* p, double v) {
    __hip_atomic_store((uint64_t*)p, __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// thread 0 of the workgroup: wait for *flag == epoch; false on an error (its own time-out or another workgroup's)
__device__ bool wait_flag(const SolArgs& s, const uint32_t* flag, uint32_t epoch) {
    uint32_t spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (__hip_atomic_load(s.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > s.spin_limit) {
            __hip_atomic_fetch_or(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

__global__ void k_epoch(uint32_t* e) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t v = *e + 1;
        *e = v ? v : 1;
    }
}

// the right side of every front of the level: own rows from b, zero below, plus the children's updates in order
__global__ __launch_bounds__(256) void k_fgather(SolArgs s, const uint32_t* list, uint32_t n) {
    if (blockIdx.x >= n || (s.skip && *s.skip)) return;
    const uint32_t fi = list[blockIdx.x];
    const Front fr = s.fr[fi];
    double* V = s.V + fr.voff;
    for (uint32_t r = threadIdx.x; r < fr.m; r += blockDim.x)
        V[r] = r < fr.c ? s.b[6 * (size_t)s.ord2f[fr.own0 + r / 6] + r % 6] : 0.0;
    __syncthreads();
    for (uint32_t k = s.ch_off[fi]; k < s.ch_off[fi + 1]; k++) {
        const Front ch = s.fr[s.children[k]];
        for (uint32_t r = threadIdx.x; r < 6 * ch.nr; r += blockDim.x)
            V[6 * s.cmap[ch.row0 + r / 6] + r % 6] += s.V[ch.voff + ch.c + r];
        __syncthreads();
    }
}

// task = (front, row block): rows R .. R + 64 of the front. What does not depend on the other blocks is read ahead
// of the waits: the diagonal block into LDS first, and each column block's tile into registers before its flag is
// awaited, so that a block's turn on the chain is the flag, the block's y, its products and the diagonal solve.
__global__ __launch_bounds__(256) void k_fsolve(SolArgs s, const uint2* tasks, uint32_t n) {
#pragma clang fp contract(fast)
    __shared__ double ys[2][64];
    __shared__ double red[4][64];
    __shared__ double Ld[64][65];
    __shared__ int sbad;
    if (blockIdx.x >= n || (s.skip && *s.skip)) return;
    const uint2 t = tasks[blockIdx.x];
    const Front fr = s.fr[t.x];
    const uint32_t m = fr.m, c = fr.c, R = t.y * 64, nrow = m - R < 64u ? m - R : 64u;
    const uint32_t tid = threadIdx.x, lane = tid & 63, part = tid >> 6;
    const uint32_t epoch = *s.epoch;
    const double* F = s.F + fr.off_l;
    double* V = s.V + fr.voff;
    if (tid == 0) sbad = 0;
    if (R < c) {
        const uint32_t w = c - R < 64u ? c - R : 64u;
        for (uint32_t e = tid; e < 64 * 64; e += 256) {
            const uint32_t i = e % 64, j = e / 64;
            Ld[i][j] = (i < nrow && j < w && i >= j) ? F[(uint64_t)(R + j) * m + R + i] : 0.0;
        }
    }
    const double vinit = lane < nrow ? V[R + lane] : 0.0;
    const uint32_t rown = R + lane;
    const size_t ydst = part == 0 && lane < nrow && rown < c ? 6 * (size_t)s.ord2f[fr.own0 + rown / 6] + rown % 6 : 0;
    double ys_own = 0, pacc = 0;
    const uint32_t ncol = R < c ? R : c, nblk = (ncol + 63) / 64, j0 = part * 16;
    for (uint32_t jb = 0; jb < nblk; jb++) {
        const uint32_t J = jb * 64, w = c - J < 64u ? c - J : 64u;
        double fv[16];
#pragma unroll
        for (int q = 0; q < 16; q++) fv[q] = (lane < nrow && j0 + q < w) ? F[(uint64_t)(J + j0 + q) * m + R + lane] : 0.0;
        if (tid == 0 && !sbad && !wait_flag(s, &s.fflag[fr.fb + jb], epoch)) sbad = 1;
        __syncthreads();
        if (sbad) break;
        // (two y buffers: a wave still in the last block's products reads the other one)
        if (tid < w) ys[jb & 1][tid] = ld_dev(&V[J + tid]);
        __syncthreads();
        if (lane < nrow) {
#pragma unroll
            for (int q = 0; q < 16; q++)
                if (j0 + q < w) pacc += fv[q] * ys[jb & 1][j0 + q];
        }
    }
    red[part][lane] = pacc;
    __syncthreads();
    if (part == 0 && !sbad) {
        double acc = vinit - (((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
        if (R < c) {
            const uint32_t w = c - R < 64u ? c - R : 64u;
            for (uint32_t j = 0; j < w; j++) {
                const double yj = readlane_u(acc, j) * Ld[j][j]; // (L's diagonal holds 1 / L_jj)
                if (lane == j) acc = yj;
                else if (lane > j) acc -= Ld[lane][j] * yj;
            }
        }
        if (lane < nrow) st_dev(&V[R + lane], acc);
        ys_own = acc;
        __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();
    if (tid == 0 && R < c && !sbad) __hip_atomic_store(&s.fflag[fr.fb + t.y], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // (the solution in the system's order after the flag: only later launches read it)
    if (part == 0 && !sbad && lane < nrow && R + lane < c) s.y[ydst] = ys_own;
}

// task = (front, own column block jb), the front's last block first; the diagonal block and each later block's tile
// read ahead of the waits, as in the forward pass
__global__ __launch_bounds__(256) void k_bsolve(SolArgs s, const uint2* tasks, uint32_t n) {
#pragma clang fp contract(fast)
    __shared__ double xb[2][64];
    __shared__ double dots[64];
    __shared__ double Ld[64][65];
    __shared__ int sbad;
    if (blockIdx.x >= n || (s.skip && *s.skip)) return;
    const uint2 t = tasks[blockIdx.x];
    const Front fr = s.fr[t.x];
    const uint32_t m = fr.m, c = fr.c, J = t.y * 64, w = c - J < 64u ? c - J : 64u;
    const uint32_t tid = threadIdx.x, lane = tid & 63, part = tid >> 6;
    const uint32_t epoch = *s.epoch;
    const double* F = s.F + fr.off_l;
    double* X = s.X + fr.voff;
    if (tid == 0) sbad = 0;
    // the own rows' right side (the forward pass's y), read now rather than on the chain
    const double vown = part == 0 && lane < w ? s.V[fr.voff + J + lane] : 0.0;
    const size_t xdst = part == 0 && lane < w ? 6 * (size_t)s.ord2f[fr.own0 + (J + lane) / 6] + (J + lane) % 6 : 0;
    double xs_own = 0;
    // this wave's 16 columns: partial sums of L[r][j] x_r per lane
    double pj[16];
#pragma unroll
    for (int q = 0; q < 16; q++) pj[q] = 0;
    // the rows below the own columns: x known (the ancestors'), gathered 4 096 at a time (all of a front's, on this
    // plate) into the diagonal block's LDS, then the products, lane over rows (rows in increasing order per lane)
    for (uint32_t base = 0; base < m - c; base += 4096) {
        double* xr = &Ld[0][0];
        const uint32_t nb = m - c - base < 4096u ? m - c - base : 4096u;
        for (uint32_t k = tid; k < nb; k += 256) {
            const uint32_t kk = base + k;
            xr[k] = s.x[6 * (size_t)s.ord2f[s.rows[fr.row0 + kk / 6]] + kk % 6];
        }
        __syncthreads();
        for (uint32_t k0 = 0; k0 < nb; k0 += 64) {
            const uint32_t k = k0 + lane;
            if (k < nb) {
                const double xv = xr[k];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t j = part * 16 + q;
                    if (j < w) pj[q] += F[(uint64_t)(J + j) * m + c + base + k] * xv;
                }
            }
        }
        __syncthreads();
    }
    for (uint32_t e = tid; e < 64 * 64; e += 256) {
        const uint32_t i = e % 64, j = e / 64;
        Ld[i][j] = (i < w && j < w && i >= j) ? F[(uint64_t)(J + j) * m + J + i] : 0.0;
    }
    // the later own blocks, the last first, as their flags come up
    const uint32_t nb = (c + 63) / 64;
    for (uint32_t kb = nb; kb-- > t.y + 1;) {
        const uint32_t K = kb * 64, wk = c - K < 64u ? c - K : 64u;
        double fv[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const uint32_t j = part * 16 + q;
            fv[q] = (j < w && lane < wk) ? F[(uint64_t)(J + j) * m + K + lane] : 0.0;
        }
        if (tid == 0 && !sbad && !wait_flag(s, &s.bflag[fr.fb + kb], epoch)) sbad = 1;
        __syncthreads();
        if (sbad) break;
        if (tid < wk) xb[kb & 1][tid] = ld_dev(&X[K + tid]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const uint32_t j = part * 16 + q;
            if (j < w && lane < wk) pj[q] += fv[q] * xb[kb & 1][lane];
        }
    }
    // one sum per column (a fixed tree over the lanes)
#pragma unroll
    for (int q = 0; q < 16; q++) {
        double v = pj[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) dots[part * 16 + q] = v;
    }
    __syncthreads();
    if (part == 0 && !sbad) {
        double acc = lane < w ? vown - dots[lane] : 0.0;
        for (uint32_t j = w; j-- > 0;) {
            const double xj = readlane_u(acc, j) * Ld[j][j]; // (L's diagonal holds 1 / L_jj)
            if (lane == j) acc = xj;
            else if (lane < j) acc -= Ld[j][lane] * xj;
        }
        if (lane < w) st_dev(&X[J + lane], acc);
        xs_own = acc;
        __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();
    if (tid == 0 && !sbad) __hip_atomic_store(&s.bflag[fr.fb + t.y], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // (the solution in the system's order after the flag: only later launches read it)
    if (part == 0 && !sbad && lane < w) s.x[xdst] = xs_own;
}

// ------------------------------------------------------------------------------------------ residual and update
// r = A x - b over the assembled system (diagonal blocks, and each pair block both ways): per free node, its
// pairs (CSR, both directions) - then the partial sums of |r|^2 and |b|^2 per workgroup.
__global__ void k_residual(FemArgs a, const double* x, const uint32_t* adj_off, const uint2* adj, double* part) {
    __shared__ double s1[256], s2[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double rr = 0, bb = 0;
    if (i < a.nf) {
        double y[6];
        const double* D = a.Dsys + 36 * (size_t)i;
        for (int p = 0; p < 6; p++) {
            double acc = 0;
            for (int q = 0; q < 6; q++) acc += (p >= q ? D[p * 6 + q] : D[q * 6 + p]) * x[6 * i + q];
            y[p] = acc;
        }
        for (uint32_t k = adj_off[i]; k < adj_off[i + 1]; k++) {
            const uint2 e = adj[k]; // (pair, other free index | direction << 31)
            const double* O = a.Osys + 36 * (size_t)e.x;
            const uint32_t o = e.y & 0x7fffffffu;
            const bool row_side = (e.y >> 31) != 0; // this node is the pair's row node: O x_other; else O^T x_other
            for (int p = 0; p < 6; p++) {
                double acc = 0;
                for (int q = 0; q < 6; q++) acc += (row_side ? O[p * 6 + q] : O[q * 6 + p]) * x[6 * o + q];
                y[p] += acc;
            }
        }
        for (int p = 0; p < 6; p++) {
            const double r = y[p] - a.rhs[6 * i + p];
            rr += r * r;
            bb += a.rhs[6 * i + p] * a.rhs[6 * i + p];
        }
    }
    s1[threadIdx.x] = rr, s2[threadIdx.x] = bb;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) s1[threadIdx.x] += s1[threadIdx.x + off], s2[threadIdx.x] += s2[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[2 * blockIdx.x] = s1[0], part[2 * blockIdx.x + 1] = s2[0];
}

// y = A x over the assembled system (the lower triangle of the diagonal blocks and each pair block both ways)
__global__ void k_spmv(FemArgs a, const double* x, const uint32_t* adj_off, const uint2* adj, double* y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nf) return;
    double acc[6];
    const double* D = a.Dsys + 36 * (size_t)i;
    for (int p = 0; p < 6; p++) {
        double v = 0;
        for (int q = 0; q < 6; q++) v += (p >= q ? D[p * 6 + q] : D[q * 6 + p]) * x[6 * i + q];
        acc[p] = v;
    }
    for (uint32_t k = adj_off[i]; k < adj_off[i + 1]; k++) {
        const uint2 e = adj[k];
        const double* O = a.Osys + 36 * (size_t)e.x;
        const uint32_t o = e.y & 0x7fffffffu;
        const bool row_side = (e.y >> 31) != 0;
        for (int p = 0; p < 6; p++) {
            double v = 0;
            for (int q = 0; q < 6; q++) v += (row_side ? O[p * 6 + q] : O[q * 6 + p]) * x[6 * o + q];
            acc[p] += v;
        }
    }
    for (int p = 0; p < 6; p++) y[6 * i + p] = acc[p];
}

__global__ void k_sum_parts(const double* part, uint32_t n, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0, b = 0;
    for (uint32_t k = 0; k < n; k++) a += part[2 * k], b += part[2 * k + 1];
    out[0] = a, out[1] = b;
}

// The end of the substep, per node:
//  - solve_impl's update (src/phys/frame_fem.cpp:3000-3027): a pinned node's spin zeroed; a free node's new velocity
//    handed to the integrator as the force dv m / step, unless the solution is not finite (counted; the node's gathered
//    forces then act explicitly), |dv| <= 40 m/s and |w| <= 3000 rad/s (counted). After a factorization that failed
//    (explicit_step) every node's gathered forces act and the held impulses are cleared (:2904-2908);
//  - SoftBody::integrate_nodes (src/phys/softbody.cpp:769-795): v += f inv_mass dt, the 400 m/s clamp, a NaN velocity
//    reset (counted), p += v dt;
//  - FemFrame::sync_positions (:3407-3422): x += h v in double, the orientation turned by w h, the float position
//    taken from the double one when within 1e-4 (1 + |p|), else the double one from it. (So (float)x == p after every
//    step, and FemFrame::begin_forces's resync of x from p, :1867-1872, changes nothing: the host's set_state keeps
//    the two in step.)
// integrate_nodes and sync_positions for node i over the substep (the force F, the spin w already updated)
__device__ void integrate_sync(const FemArgs& a, uint32_t i, f3 F, f3 w) {
    a.w[3 * i] = w.x, a.w[3 * i + 1] = w.y, a.w[3 * i + 2] = w.z;
    a.fnode[3 * i] = F.x, a.fnode[3 * i + 1] = F.y, a.fnode[3 * i + 2] = F.z;
    f3 v = mk(a.v[3 * i], a.v[3 * i + 1], a.v[3 * i + 2]);
    const float im = a.invm[i], h = a.step;
    f3 p = mk(a.pf[3 * i], a.pf[3 * i + 1], a.pf[3 * i + 2]);
    if (im > 0) {
        v = blg::add(v, blg::mul(F, im * h));
        const float v2 = blg::dot(v, v);
        if (v2 > 400.0f * 400.0f) {
            v = blg::mul(v, 400.0f / sqrtf(v2));
        } else if (!(v2 >= 0.0f)) {
            v = mk(0, 0, 0);
            atomicAdd(&a.cnt[2], 1ull);
        }
        p = blg::add(p, blg::mul(v, h));
        a.v[3 * i] = v.x, a.v[3 * i + 1] = v.y, a.v[3 * i + 2] = v.z;
    }
    double* xd = a.xd + 3 * (size_t)i;
    xd[0] += (double)h * v.x, xd[1] += (double)h * v.y, xd[2] += (double)h * v.z;
    if (blg::dot(w, w) > 0) {
        const float4 q0 = a.q[i];
        const qf qn = qnorm(qmul(qexp(blg::mul(w, h)), Q(q0.x, q0.y, q0.z, q0.w)));
        a.q[i] = make_float4(qn.x, qn.y, qn.z, qn.w);
    }
    const double tol = 1e-4 * (1.0 + fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
    if (im > 0 && fabs(xd[0] - p.x) <= tol && fabs(xd[1] - p.y) <= tol && fabs(xd[2] - p.z) <= tol)
        p = mk((float)xd[0], (float)xd[1], (float)xd[2]);
    else
        xd[0] = p.x, xd[1] = p.y, xd[2] = p.z;
    a.pf[3 * i] = p.x, a.pf[3 * i + 1] = p.y, a.pf[3 * i + 2] = p.z;
    a.torq[3 * i] = a.torq[3 * i + 1] = a.torq[3 * i + 2] = 0; // (the moments from outside used up)
}

__global__ void k_update(FemArgs a, const double* x, int explicit_step) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (explicit_step && i == 0) atomicAdd(&a.cnt[0], 1ull);
    if (i >= a.nn) return;
    const int fi = a.fidx[i];
    f3 w = mk(a.w[3 * i], a.w[3 * i + 1], a.w[3 * i + 2]);
    f3 F = mk(a.fnode[3 * i], a.fnode[3 * i + 1], a.fnode[3 * i + 2]);
    if (fi < 0) {
        w = mk(0, 0, 0);
    } else if (explicit_step) {
        a.imp[3 * i] = a.imp[3 * i + 1] = a.imp[3 * i + 2] = 0;
    } else {
        const double* xs = x + 6 * (size_t)fi;
        f3 dv = mk((float)xs[0], (float)xs[1], (float)xs[2]);
        const f3 dw = mk((float)xs[3], (float)xs[4], (float)xs[5]);
        if (!__builtin_isfinite(dv.x + dv.y + dv.z + dw.x + dw.y + dw.z)) {
            atomicAdd(&a.cnt[0], 1ull);
        } else {
            const float kMaxDv = 40.0f, kMaxW = 3000.0f;
            float l = blg::length(dv);
            if (l > kMaxDv) dv = blg::mul(dv, kMaxDv / l), atomicAdd(&a.cnt[1], 1ull);
            F = blg::mul(dv, a.mass[i] / a.step);
            w = blg::add(w, dw);
            l = blg::length(w);
            if (l > kMaxW) w = blg::mul(w, kMaxW / l), atomicAdd(&a.cnt[1], 1ull);
        }
    }
    integrate_sync(a, i, F, w);
}

// A held substep's right side (FemFrame::held_begin / held_rhs_col, src/phys/frame_fem.cpp:2342-2370): the node's
// forces (gravity, the load) and moments from outside beyond the last step's prediction, times the substep, the
// prediction's share of the held impulse paid; the forces kept for the update. r: 6 per free node.
__global__ void k_held_rhs(FemArgs a, double* r) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nn) return;
    f3 F = blg::mul(mk(a.g[0], a.g[1], a.g[2]), a.mass[i]);
    if (a.loads) F = blg::add(F, mk(a.loads[3 * i], a.loads[3 * i + 1], a.loads[3 * i + 2]));
    a.fnode[3 * i] = F.x, a.fnode[3 * i + 1] = F.y, a.fnode[3 * i + 2] = F.z;
    const int fi = a.fidx[i];
    if (fi < 0) return;
    double* rr = r + 6 * (size_t)fi;
    if (!(a.invm[i] > 0)) {
        for (int j = 0; j < 6; j++) rr[j] = 0;
        return;
    }
    const f3 pr = mk(a.pred[3 * i], a.pred[3 * i + 1], a.pred[3 * i + 2]);
    const f3 tp = mk(a.tpred[3 * i], a.tpred[3 * i + 1], a.tpred[3 * i + 2]);
    const f3 d = blg::mul(blg::sub(F, pr), a.step);
    const f3 t = blg::mul(blg::sub(mk(a.torq[3 * i], a.torq[3 * i + 1], a.torq[3 * i + 2]), tp), a.step);
    rr[0] = d.x, rr[1] = d.y, rr[2] = d.z, rr[3] = t.x, rr[4] = t.y, rr[5] = t.z;
    const f3 J = blg::mul(pr, a.step);
    a.imp[3 * i] += J.x, a.imp[3 * i + 1] += J.y, a.imp[3 * i + 2] += J.z;
}

// The end of a held substep (held_vel_col, src/phys/frame_fem.cpp:2371-2395): a free node's velocity change from the
// last factorization's solution as the force dv m / step (a non-finite one: no force), its spin turned, |dv| <= 40 m/s
// and |w| <= 3000 rad/s (counted); without a factorization FemFrame::hold (:2321-2340): the forces held over as
// impulse. Then integrate_nodes and sync_positions.
__global__ void k_held_update(FemArgs a, const double* x, int factored) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nn) return;
    f3 w = mk(a.w[3 * i], a.w[3 * i + 1], a.w[3 * i + 2]);
    f3 F = mk(a.fnode[3 * i], a.fnode[3 * i + 1], a.fnode[3 * i + 2]);
    const int fi = a.fidx[i];
    if (!factored) {
        if (a.invm[i] > 0) {
            const f3 fc = mk(0, 0, 0), J = blg::mul(blg::sub(F, fc), a.step);
            a.imp[3 * i] += J.x, a.imp[3 * i + 1] += J.y, a.imp[3 * i + 2] += J.z;
            const f3 v = blg::add(mk(a.v[3 * i], a.v[3 * i + 1], a.v[3 * i + 2]), blg::mul(fc, a.step * a.invm[i]));
            a.v[3 * i] = v.x, a.v[3 * i + 1] = v.y, a.v[3 * i + 2] = v.z;
            F = mk(0, 0, 0);
        }
    } else if (fi >= 0) {
        const double* xs = x + 6 * (size_t)fi;
        f3 dv = mk((float)xs[0], (float)xs[1], (float)xs[2]);
        const f3 dw = mk((float)xs[3], (float)xs[4], (float)xs[5]);
        if (!__builtin_isfinite(dv.x + dv.y + dv.z + dw.x + dw.y + dw.z)) {
            F = mk(0, 0, 0);
        } else {
            const float kMaxDv = 40.0f, kMaxW = 3000.0f;
            float l = blg::length(dv);
            if (l > kMaxDv) dv = blg::mul(dv, kMaxDv / l), atomicAdd(&a.cnt[1], 1ull);
            F = blg::mul(dv, a.mass[i] / a.step);
            w = blg::add(w, dw);
            l = blg::length(w);
            if (l > kMaxW) w = blg::mul(w, kMaxW / l), atomicAdd(&a.cnt[1], 1ull);
        }
    }
    integrate_sync(a, i, F, w);
}

} // namespace blf

// ============================================================================================== host: the engine
namespace {

using namespace blf;

#define FCHK(x)                                                                                                     \
    do {                                                                                                            \
        hipError_t e_ = (x);                                                                                        \
        if (e_ != hipSuccess) {                                                                                     \
            blg::set_last_error((std::string(#x) + ": " + hipGetErrorString(e_)).c_str());                          \
            return BL_ERR_HIP;                                                                                      \
        }                                                                                                           \
    } while (0)

template <class T>
struct Dv {
    T* p = nullptr;
    size_t n = 0;
    int put(const std::vector<T>& v) {
        if (p) (void)hipFree(p), p = nullptr;
        n = v.size();
        if (!n) return BL_OK;
        if (hipMalloc((void**)&p, sizeof(T) * n) != hipSuccess) return BL_ERR_HIP;
        return hipMemcpy(p, v.data(), sizeof(T) * n, hipMemcpyHostToDevice) == hipSuccess ? BL_OK : BL_ERR_HIP;
    }
    int alloc(size_t cnt) {
        if (p) (void)hipFree(p), p = nullptr;
        n = cnt;
        if (!n) return BL_OK;
        return hipMalloc((void**)&p, sizeof(T) * n) == hipSuccess ? BL_OK : BL_ERR_HIP;
    }
    ~Dv() {
        if (p) (void)hipFree(p);
    }
};

struct Level {
    std::vector<uint32_t> fronts;
    uint32_t max_c = 0;
};

} // namespace

struct bl_fem {
    hipStream_t st = nullptr;
    hipEvent_t ev[5] = {};
    uint32_t nn = 0, nt = 0, nf = 0, npairs = 0;
    uint64_t serial = 0; // unique per engine in the process (an address can be reused after a destroy)
    bool factor_ok = false; // the factor buffer holds the last successful factorization (of an earlier system)
    Section sec{};
    // state and topology
    Dv<double> xd;
    Dv<float> v, w, mass, inertia, X, Kf, te, tfo, tto, tg, loads;
    Dv<float> pf, invm, fnode, imp; // the body's node positions and inverse masses, the forces, the held impulses
    Dv<unsigned long long> cnt;     // solve failures, clamps, NaN resets
    Dv<float> torq;                 // moments from outside for the next step
    uint32_t nm = 0;                // members and their arrays (FemArgs)
    Dv<uint4> mem;
    Dv<float> mL0, mst, mtw, mdamp, mfo;
    Dv<float4> mqa, mqb, me1;
    Dv<uint8_t> mon;
    Dv<double> mW, mkv;
    // plasticity: the members' sections, plastic state, tangents and their bits; the triangles' authored rest shapes,
    // plastic rest rotations, damage, utilisation and curvature rows; the held impulses of a step's first pass; the
    // components (the members' own, their node counts); the passes-again budget and the statistics
    Dv<float> msec, mps, mtan, X0, th0, tdmg, tutil, Bc, jsave;
    Dv<float> pred, tpred;  // the prediction of the smooth forces and moments for the held substeps
    Dv<double> hrhs, hx;    // a held substep's right side and solution
    Dv<uint32_t> mbits, nyield, mcomp, again;
    std::vector<uint32_t> comp_size;
    uint32_t ncomp = 0;
    bool plastic_members = false;
    float repass_credit = 8.0f;
    int64_t repasses = 0, passes = 0;
    Dv<float4> q, r0;
    Dv<double> Kd, tkv, Dsys, rhs, Osys, y, x, xprev, part, sums;
    Dv<float> p0f;     // rest positions of the free nodes (3 each)
    FemArgs step_args{}; // the last step's arguments (the update's)
    Dv<uint8_t> fixed, ton;
    Dv<uint4> tri, pairs, pcon;
    Dv<uint32_t> nc_off, nc, ford, ord2f, rows, children, ch_off, cmap, adj_off;
    Dv<int> fidx, fail;
    Dv<uint2> adj;
    // the factorization
    Dv<Front> fronts;
    Dv<double> F, Lbuf, V, Xs;
    Dv<double> Lkeep;     // a copy of the factor (kept across a factorization that must not replace it)
    bool kept = false;
    Dv<uint32_t> fflag, bflag, epoch;
    Dv<uint2> fs_tasks, bs_tasks;
    Dv<Scat> scat;
    std::vector<Front> hfronts;
    std::vector<Level> levels; // bottom-up
    // task lists: per level the extend-add rounds, per (level, panel) the potrf / trsm / syrk tasks
    struct Step {
        uint32_t k;
        uint64_t pfac, npfac, syrk, nsyrk; // its tasks in pfac_tasks, syrk_tasks
    };
    struct LevelPlan {
        std::vector<std::pair<uint64_t, uint32_t>> ea; // per round: (offset, count)
        std::vector<Step> steps;
        uint64_t list_off;                              // the level's fronts in `flist`
        uint32_t nlist, max_m;
        uint64_t fs_off, bs_off;                        // the substitutions' block tasks
        uint32_t nfs, nbs;
    };
    std::vector<LevelPlan> plan;
    Dv<EaTask> ea_tasks;
    Dv<uint2> ztasks; // (front, first of 16 columns): the clearing of the fronts' lower triangles
    Dv<uint32_t> flist;
    Dv<uint2> pfac_tasks;
    Dv<uint4> syrk_tasks;
    uint32_t nparts = 0;
    double flops = 0;
    int64_t nnz = 0;
    int fronts_max = 0;
    bool have_system = false;

    ~bl_fem() {
        if (st) (void)hipStreamSynchronize(st);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    FemArgs args(const float* dloads, float h, float theta, float dis, const float g[3]) const {
        FemArgs a{};
        a.nn = nn, a.nt = nt, a.nf = nf, a.sec = sec;
        a.xd = xd.p, a.v = v.p, a.w = w.p, a.q = q.p, a.mass = mass.p, a.inertia = inertia.p, a.fixed = fixed.p;
        a.pf = pf.p, a.invm = invm.p, a.fnode = fnode.p, a.imp = imp.p, a.cnt = cnt.p, a.torq = torq.p;
        a.nm = nm, a.mem = mem.p, a.mL0 = mL0.p, a.mqa = mqa.p, a.mqb = mqb.p, a.mst = mst.p, a.mtw = mtw.p;
        a.mdamp = mdamp.p, a.mon = mon.p, a.mfo = mfo.p, a.me1 = me1.p, a.mW = mW.p, a.mkv = mkv.p;
        a.msec = msec.p, a.mps = mps.p, a.mtan = mtan.p, a.mbits = mbits.p, a.nyield = nyield.p, a.mcomp = mcomp.p;
        a.again = again.p, a.jsave = jsave.p, a.pred = pred.p, a.tpred = tpred.p;
        a.tri = tri.p, a.X = X.p, a.r0 = r0.p, a.Kf = Kf.p, a.Kd = Kd.p;
        a.X0 = X0.p, a.th0 = th0.p, a.tdmg = tdmg.p, a.tutil = tutil.p, a.Bc = Bc.p;
        a.te = te.p, a.tfo = tfo.p, a.tto = tto.p, a.tg = tg.p, a.tkv = tkv.p, a.ton = ton.p;
        a.nc_off = nc_off.p, a.nc = nc.p, a.fidx = fidx.p, a.ford = ford.p;
        a.Dsys = Dsys.p, a.rhs = rhs.p, a.Osys = Osys.p, a.pairs = pairs.p, a.pcon = pcon.p, a.npairs = npairs;
        a.loads = dloads;
        if (g) a.g[0] = g[0], a.g[1] = g[1], a.g[2] = g[2];
        a.h = h, a.theta = theta, a.dissipation = dis, a.step = h;
        return a;
    }
    FacArgs fac() const {
        FacArgs f{};
        f.F = F.p, f.L = Lbuf.p, f.fr = fronts.p, f.fail = fail.p;
        return f;
    }
    SolArgs sol(const double* b) const {
        SolArgs s{};
        s.F = Lbuf.p, s.fr = fronts.p, s.V = V.p, s.X = Xs.p, s.rows = rows.p, s.children = children.p;
        s.ch_off = ch_off.p, s.cmap = cmap.p, s.ord2f = ord2f.p, s.b = b, s.y = y.p, s.x = x.p;
        s.fflag = fflag.p, s.bflag = bflag.p, s.epoch = epoch.p, s.err = fail.p, s.spin_limit = 1u << 20;
        return s;
    }
};

namespace {

// Nested dissection of the free nodes by their rest positions: the set split at the median of its longest extent,
// the nodes of the smaller side touching the other side taken out as the separator; sets of at most `leaf` nodes
// are leaves. Returns the supernodes (own nodes) with their children, in post-order.
struct SNode {
    std::vector<uint32_t> own;
    std::vector<uint32_t> children;
};

void dissect(const std::vector<uint32_t>& S, const std::vector<float>& pos, const std::vector<std::vector<uint32_t>>& adj,
             std::vector<int>& side, std::vector<SNode>& out, std::vector<uint32_t>& roots, uint32_t leaf) {
    if (S.size() <= leaf) {
        SNode s;
        s.own = S;
        out.push_back(s);
        roots.push_back((uint32_t)out.size() - 1);
        return;
    }
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i : S)
        for (int k = 0; k < 3; k++) mn[k] = std::min(mn[k], pos[3 * i + k]), mx[k] = std::max(mx[k], pos[3 * i + k]);
    int ax = 0;
    for (int k = 1; k < 3; k++)
        if (mx[k] - mn[k] > mx[ax] - mn[ax]) ax = k;
    std::vector<uint32_t> sorted = S;
    std::stable_sort(sorted.begin(), sorted.end(), [&](uint32_t a, uint32_t b) { return pos[3 * a + ax] < pos[3 * b + ax]; });
    const size_t half = sorted.size() / 2;
    std::vector<uint32_t> L(sorted.begin(), sorted.begin() + half), R(sorted.begin() + half, sorted.end());
    for (uint32_t i : L) side[i] = 1;
    for (uint32_t i : R) side[i] = 2;
    std::vector<uint32_t> bl, br;
    for (uint32_t i : L)
        for (uint32_t j : adj[i])
            if (side[j] == 2) { bl.push_back(i); break; }
    for (uint32_t i : R)
        for (uint32_t j : adj[i])
            if (side[j] == 1) { br.push_back(i); break; }
    for (uint32_t i : S) side[i] = 0;
    const bool use_left = bl.size() <= br.size();
    std::vector<uint32_t>& sep = use_left ? bl : br;
    std::vector<uint8_t> in_sep(0);
    std::vector<uint32_t> L2, R2;
    {
        std::vector<uint32_t> sp = sep;
        std::sort(sp.begin(), sp.end());
        auto in = [&](uint32_t i) { return std::binary_search(sp.begin(), sp.end(), i); };
        for (uint32_t i : L)
            if (!(use_left && in(i))) L2.push_back(i);
        for (uint32_t i : R)
            if (!(!use_left && in(i))) R2.push_back(i);
    }
    std::vector<uint32_t> rl, rr;
    if (!L2.empty()) dissect(L2, pos, adj, side, out, rl, leaf);
    if (!R2.empty()) dissect(R2, pos, adj, side, out, rr, leaf);
    if (sep.empty()) {
        roots.insert(roots.end(), rl.begin(), rl.end());
        roots.insert(roots.end(), rr.begin(), rr.end());
        return;
    }
    SNode s;
    s.own = sep;
    s.children = rl;
    s.children.insert(s.children.end(), rr.begin(), rr.end());
    out.push_back(s);
    roots.push_back((uint32_t)out.size() - 1);
}

int fem_build(bl_fem* f, const bl_fem_desc* d) {
    const uint32_t nn = d->n_nodes, nt = d->n_tris;
    f->nn = nn, f->nt = nt;
    f->sec = Section{d->section.E, d->section.nu, d->section.rho, d->section.t, d->section.drill, d->section.damping,
                     d->section.yield};
    // masses (ShellMesher::tri), rest shapes and frames (FemFrame::finalize), inertia (node_inertia)
    std::vector<float> mass(nn, 0.0f), inertia(nn, 0.0f), X(6 * (size_t)nt, 0.0f), tmass(nt), area0(nt);
    std::vector<float4> r0(3 * (size_t)nt);
    std::vector<uint4> tri(nt);
    const float mpa = d->section.rho * d->section.t;
    auto P = [&](uint32_t i) { return mk(d->p[3 * i], d->p[3 * i + 1], d->p[3 * i + 2]); };
    for (uint32_t k = 0; k < nt; k++) {
        const uint32_t a = d->tris[3 * k], b = d->tris[3 * k + 1], c = d->tris[3 * k + 2];
        if (a >= nn || b >= nn || c >= nn || a == b || b == c || a == c) {
            blg::set_last_error("a triangle's corners are out of range or repeated");
            return BL_ERR_ARG;
        }
        tri[k] = make_uint4(a, b, c, 0);
        const f3 x0 = P(a), x1 = P(b), x2 = P(c);
        const f3 u = blg::sub(x1, x0), vv = blg::sub(x2, x0);
        const f3 nn3 = mk(u.y * vv.z - u.z * vv.y, u.z * vv.x - u.x * vv.z, u.x * vv.y - u.y * vv.x);
        const float a2 = blg::length(nn3);
        tmass[k] = mpa * (0.5f * a2);
        if (!d->mass)
            for (uint32_t v2 : {a, b, c}) mass[v2] += tmass[k] / 3.0f;
        if (!(a2 > 1e-8f)) {
            blg::set_last_error("a degenerate triangle");
            return BL_ERR_ARG;
        }
        const f3 e3 = blg::divs(nn3, a2);
        const f3 d01 = blg::sub(x1, x0);
        const float l01 = blg::length(d01);
        const f3 e1 = l01 > 1e-12f ? blg::mul(d01, 1.0f / l01) : mk(0, 0, 0);
        const f3 e2 = mk(e3.y * e1.z - e3.z * e1.y, e3.z * e1.x - e3.x * e1.z, e3.x * e1.y - e3.y * e1.x);
        const f3 cc = blg::divs(blg::add(blg::add(x0, x1), x2), 3.0f);
        const f3 xs[3] = {x0, x1, x2};
        for (int i = 0; i < 3; i++) X[6 * k + 2 * i] = blg::dot(blg::sub(xs[i], cc), e1), X[6 * k + 2 * i + 1] = blg::dot(blg::sub(xs[i], cc), e2);
        area0[k] = 0.5f * a2;
        const qf E0 = from_cols(e1, e2, e3);
        for (int i = 0; i < 3; i++) {
            const qf r = qnorm(qmul(qconj(Q(0, 0, 0, 1)), E0)); // (the nodes' orientations start at identity)
            r0[3 * (size_t)k + i] = make_float4(r.x, r.y, r.z, r.w);
        }
    }
    if (d->mass) std::memcpy(mass.data(), d->mass, sizeof(float) * nn);
    // the members (FemFrame::finalize, src/phys/frame_fem.cpp:244-256): rest length and frames (x along the member, y
    // towards the world's up, a vertical member's towards x; the nodes' orientations start at identity); their elastic
    // stiffness (stiffness, :1408-1476) and twist weight (twist_weight, :1372-1375), in float as the reference has them
    const uint32_t nm = d->n_members;
    f->nm = nm;
    std::vector<uint4> mem(nm);
    std::vector<float> mL0(nm), mst(8 * (size_t)nm), mtw(nm), mdamp(nm);
    std::vector<float4> mqa(nm), mqb(nm);
    for (uint32_t k = 0; k < nm; k++) {
        const bl_frame_member& e = d->members[k];
        if (e.a >= nn || e.b >= nn || e.section >= d->n_sections || e.end_a > BL_FJ_ELASTIC || e.end_b > BL_FJ_ELASTIC) {
            blg::set_last_error("a member's nodes, section or joints are out of range");
            return BL_ERR_ARG;
        }
        const bl_frame_section& sc = d->sections[e.section];
        const f3 dd = blg::sub(P(e.b), P(e.a));
        const float L0 = blg::length(dd);
        mL0[k] = L0;
        const bool broken = L0 < 1e-5f || e.a == e.b;
        mem[k] = make_uint4(e.a, e.b, e.section, broken ? 1u : 0u);
        mdamp[k] = sc.damping;
        if (broken) continue;
        const f3 e1 = blg::divs(dd, L0);
        const f3 ref = std::fabs(e1.y) < 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
        const f3 rr = blg::sub(ref, blg::mul(e1, blg::dot(e1, ref)));
        const float lr = blg::length(rr);
        const f3 e2 = lr > 1e-12f ? blg::mul(rr, 1.0f / lr) : mk(0, 0, 0);
        const f3 e3 = mk(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
        const qf E0 = from_cols(e1, e2, e3);
        const qf qa = qnorm(qmul(qconj(Q(0, 0, 0, 1)), E0)), qb = qnorm(qmul(qconj(Q(0, 0, 0, 1)), E0));
        mqa[k] = make_float4(qa.x, qa.y, qa.z, qa.w), mqb[k] = make_float4(qb.x, qb.y, qb.z, qb.w);
        // end_flex, condense, ends2, stiffness
        struct Flex {
            bool rel[3];
            float f;
        };
        auto flex = [&](uint32_t j) {
            Flex x{{false, false, false}, 0.0f};
            if (j == BL_FJ_BALL) x.rel[0] = x.rel[1] = x.rel[2] = true;
            else if (j == BL_FJ_HINGE_V) x.rel[2] = true;
            else if (j == BL_FJ_HINGE_H) x.rel[1] = true;
            else if (j == BL_FJ_SWIVEL) x.rel[0] = true;
            else if (j == BL_FJ_ELASTIC) x.f = sc.joint_k > 0 ? 1.0f / sc.joint_k : 0.0f;
            return x;
        };
        auto condense = [](float c[3], bool pa, bool pb) {
            if (pa && pb) c[0] = c[1] = c[2] = 0;
            else if (pa) c[2] = c[0] > 0 ? c[2] - c[1] * c[1] / c[0] : 0, c[0] = c[1] = 0;
            else if (pb) c[0] = c[2] > 0 ? c[0] - c[1] * c[1] / c[2] : 0, c[1] = c[2] = 0;
        };
        const Flex fa = flex(e.end_a), fb = flex(e.end_b);
        float* st = &mst[8 * (size_t)k];
        st[0] = sc.axial * sc.E * sc.A / L0;
        st[1] = sc.G * sc.J / L0;
        if (fa.rel[0] || fb.rel[0]) st[1] = 0;
        else if (fa.f + fb.f > 0) st[1] = 1.0f / (1.0f / st[1] + fa.f + fb.f);
        auto bend = [&](float I, float As, float* c, int axis) {
            const float phi = As > 0 ? 12.0f * sc.E * I / (sc.G * As * L0 * L0) : 0.0f;
            const float k0 = sc.E * I / (L0 * (1.0f + phi));
            c[0] = c[2] = (4.0f + phi) * k0;
            c[1] = (2.0f - phi) * k0;
            const bool ra = fa.rel[axis], rb = fb.rel[axis];
            if (ra || rb) {
                condense(c, ra, rb);
                if (ra != rb) {
                    float& kk = ra ? c[2] : c[0];
                    const float fl = ra ? fb.f : fa.f;
                    if (fl > 0 && kk > 0) kk = 1.0f / (1.0f / kk + fl);
                }
                return;
            }
            if (fa.f <= 0 && fb.f <= 0) return;
            const double det = (double)c[0] * c[2] - (double)c[1] * c[1];
            if (!(det > 0)) return;
            const double F00 = c[2] / det + fa.f, F01 = -c[1] / det, F11 = c[0] / det + fb.f;
            const double d2 = F00 * F11 - F01 * F01;
            c[0] = (float)(F11 / d2), c[1] = (float)(-F01 / d2), c[2] = (float)(F00 / d2);
        };
        bend(sc.Iy, sc.As_z, st + 2, 1);
        bend(sc.Iz, sc.As_y, st + 5, 2);
        const bool ta = e.end_a == BL_FJ_BALL || e.end_a == BL_FJ_SWIVEL, tb = e.end_b == BL_FJ_BALL || e.end_b == BL_FJ_SWIVEL;
        mtw[k] = ta == tb ? 0.5f : ta ? 1.0f : 0.0f;
    }
    // node_inertia (:3075-3086): the members' halves (m L^2 / 78 per end), then a corner's share of its triangles, the
    // node's floor
    for (uint32_t k = 0; k < nm; k++) {
        if (mem[k].w) continue;
        const bl_frame_section& sc = d->sections[d->members[k].section];
        const float em = sc.rho * sc.A * mL0[k];
        inertia[mem[k].x] += em * mL0[k] * mL0[k] / 78.0f;
        inertia[mem[k].y] += em * mL0[k] * mL0[k] / 78.0f;
    }
    for (uint32_t k = 0; k < nt; k++)
        for (int c = 0; c < 3; c++) inertia[d->tris[3 * k + c]] += tmass[k] * area0[k] / 108.0f;
    for (uint32_t i = 0; i < nn; i++) inertia[i] = std::max(inertia[i], std::max(1e-6f, mass[i] * 1e-4f));
    // node -> incidences: its members' ends in member order (1 << 31 | 2 member + end), then its triangles' corners
    // in triangle order (3 triangle + corner)
    std::vector<uint32_t> nc_off(nn + 1, 0), nc(2 * (size_t)nm + 3 * (size_t)nt);
    for (uint32_t k = 0; k < nm; k++) nc_off[mem[k].x + 1]++, nc_off[mem[k].y + 1]++;
    for (uint32_t k = 0; k < nt; k++)
        for (int c = 0; c < 3; c++) nc_off[d->tris[3 * k + c] + 1]++;
    for (uint32_t i = 0; i < nn; i++) nc_off[i + 1] += nc_off[i];
    {
        std::vector<uint32_t> fill(nc_off.begin(), nc_off.end() - 1);
        for (uint32_t k = 0; k < nm; k++)
            for (uint32_t end = 0; end < 2; end++) nc[fill[end ? mem[k].y : mem[k].x]++] = 0x80000000u | (2 * k + end);
        for (uint32_t k = 0; k < nt; k++)
            for (int c = 0; c < 3; c++) nc[fill[d->tris[3 * k + c]]++] = 3 * k + (uint32_t)c;
    }
    // the free nodes, their graph
    std::vector<int> fidx(nn, -1);
    std::vector<uint32_t> freen;
    for (uint32_t i = 0; i < nn; i++)
        if (!(d->fixed && d->fixed[i])) fidx[i] = (int)freen.size(), freen.push_back(i);
    const uint32_t nf = (uint32_t)freen.size();
    f->nf = nf;
    std::vector<std::vector<uint32_t>> adj(nf);
    for (uint32_t k = 0; k < nm; k++) {
        const int a = fidx[mem[k].x], b = fidx[mem[k].y];
        if (!mem[k].w && a >= 0 && b >= 0 && a != b) adj[a].push_back((uint32_t)b), adj[b].push_back((uint32_t)a);
    }
    for (uint32_t k = 0; k < nt; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const int a = fidx[d->tris[3 * k + i]], b = fidx[d->tris[3 * k + j]];
                if (i != j && a >= 0 && b >= 0) adj[a].push_back((uint32_t)b);
            }
    for (auto& l : adj) std::sort(l.begin(), l.end()), l.erase(std::unique(l.begin(), l.end()), l.end());
    // nested dissection, post-order: the order positions
    std::vector<float> fpos(3 * (size_t)nf);
    for (uint32_t i = 0; i < nf; i++)
        for (int k = 0; k < 3; k++) fpos[3 * i + k] = d->p[3 * freen[i] + k];
    std::vector<SNode> sn;
    std::vector<uint32_t> roots;
    std::vector<int> side(nf, 0);
    std::vector<uint32_t> all(nf);
    for (uint32_t i = 0; i < nf; i++) all[i] = i;
    if (nf) dissect(all, fpos, adj, side, sn, roots, 8);
    // (dissect appends children before parents: the vector is a post-order)
    std::vector<uint32_t> order, opos(nf);
    std::vector<uint32_t> sown0(sn.size());
    for (size_t s = 0; s < sn.size(); s++) {
        sown0[s] = (uint32_t)order.size();
        for (uint32_t i : sn[s].own) opos[i] = (uint32_t)order.size(), order.push_back(i);
    }
    // the structure: a supernode's rows - later nodes adjacent to its own, and its children's rows past it
    std::vector<int> parent(sn.size(), -1);
    for (size_t s = 0; s < sn.size(); s++)
        for (uint32_t c : sn[s].children) parent[c] = (int)s;
    std::vector<std::vector<uint32_t>> srows(sn.size());
    for (size_t s = 0; s < sn.size(); s++) {
        const uint32_t last = sown0[s] + (uint32_t)sn[s].own.size();
        std::vector<uint32_t> r;
        for (uint32_t i : sn[s].own)
            for (uint32_t j : adj[i])
                if (opos[j] >= last) r.push_back(opos[j]);
        for (uint32_t c : sn[s].children)
            for (uint32_t o : srows[c])
                if (o >= last) r.push_back(o);
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        srows[s] = r;
    }
    // fronts, their storage and levels (depth from the leaves)
    std::vector<Front> fr(sn.size());
    std::vector<uint32_t> rows, cmap;
    std::vector<uint32_t> level(sn.size(), 0);
    uint64_t off = 0, voff = 0, loff = 0;
    uint32_t fb = 0;
    double flops = 0;
    int64_t nnz = 0;
    int maxm = 0;
    for (size_t s = 0; s < sn.size(); s++) {
        Front& x = fr[s];
        x.ns = (uint32_t)sn[s].own.size(), x.nr = (uint32_t)srows[s].size();
        x.c = 6 * x.ns, x.m = 6 * (x.ns + x.nr);
        x.own0 = sown0[s], x.row0 = (uint32_t)rows.size();
        rows.insert(rows.end(), srows[s].begin(), srows[s].end());
        x.parent = parent[s];
        x.off = off, off += (uint64_t)x.m * x.m;
        x.off_l = loff, loff += (uint64_t)x.m * x.c;
        x.voff = voff, voff += x.m;
        x.fb = fb, fb += (x.c + 63) / 64;
        for (uint32_t c : sn[s].children) level[s] = std::max(level[s], level[c] + 1);
        x.level = level[s];
        const double c = x.c, r = 6.0 * x.nr;
        flops += c * c * c / 3.0 + c * c * r + c * r * r;
        nnz += (int64_t)(c * (c + 1) / 2 + c * r);
        maxm = std::max(maxm, (int)x.m);
    }
    // each child's row nodes -> positions in its parent's front (own nodes first, then rows)
    cmap.assign(rows.size(), 0);
    for (size_t s = 0; s < sn.size(); s++) {
        if (fr[s].parent < 0) continue;
        const Front& p = fr[fr[s].parent];
        for (uint32_t k = 0; k < fr[s].nr; k++) {
            const uint32_t o = rows[fr[s].row0 + k];
            uint32_t pos;
            if (o >= p.own0 && o < p.own0 + p.ns) pos = o - p.own0;
            else {
                const auto it = std::lower_bound(rows.begin() + p.row0, rows.begin() + p.row0 + p.nr, o);
                if (it == rows.begin() + p.row0 + p.nr || *it != o) {
                    blg::set_last_error("symbolic factorization: a child's row is missing from its parent");
                    return BL_ERR_STATE;
                }
                pos = p.ns + (uint32_t)(it - (rows.begin() + p.row0));
            }
            cmap[fr[s].row0 + k] = pos;
        }
    }
    f->flops = flops, f->nnz = nnz, f->fronts_max = maxm;
    // the pairs (lower blocks) and their triangles, in triangle order
    std::vector<uint4> pairs, pcon;
    {
        std::vector<std::vector<uint4>> con; // per pair
        std::vector<uint64_t> keys;
        // (row free, col free, member m or nm + triangle, corners / the member's block): a pair's members first, then
        // its triangles (gather_lists, src/phys/frame_fem.cpp:883-896)
        std::vector<uint4> raw;
        for (uint32_t k = 0; k < nm; k++) {
            const int a = fidx[mem[k].x], b = fidx[mem[k].y];
            if (mem[k].w || a < 0 || b < 0 || a == b) continue;
            // (the row is the later node: its block of the member's matrix, ba when b is later, else ab)
            if (opos[b] > opos[a]) raw.push_back(make_uint4((uint32_t)b, (uint32_t)a, k, 2u));
            else raw.push_back(make_uint4((uint32_t)a, (uint32_t)b, k, 3u));
        }
        for (uint32_t k = 0; k < nt; k++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    if (i == j) continue;
                    const int a = fidx[d->tris[3 * k + i]], b = fidx[d->tris[3 * k + j]];
                    if (a < 0 || b < 0) continue;
                    if (opos[a] <= opos[b]) continue; // (row: the later node in the order)
                    raw.push_back(make_uint4((uint32_t)a, (uint32_t)b, nm + k, (uint32_t)(i | (j << 8))));
                }
        std::stable_sort(raw.begin(), raw.end(), [](const uint4& x, const uint4& y) {
            return x.x != y.x ? x.x < y.x : x.y != y.y ? x.y < y.y : x.z < y.z;
        });
        for (size_t e = 0; e < raw.size();) {
            size_t e2 = e;
            while (e2 < raw.size() && raw[e2].x == raw[e].x && raw[e2].y == raw[e].y) e2++;
            pairs.push_back(make_uint4(raw[e].x, raw[e].y, (uint32_t)pcon.size(), (uint32_t)(pcon.size() + (e2 - e))));
            for (size_t q2 = e; q2 < e2; q2++)
                pcon.push_back(raw[q2].z < nm ? make_uint4(0x80000000u | raw[q2].z, raw[q2].w, 0, 0)
                                              : make_uint4(raw[q2].z - nm, raw[q2].w & 0xff, raw[q2].w >> 8, 0));
            e = e2;
        }
    }
    f->npairs = (uint32_t)pairs.size();
    // the scatter list: diagonal blocks and pair blocks into the fronts
    std::vector<Scat> scat;
    std::vector<uint32_t> front_of(nf);
    for (size_t s = 0; s < sn.size(); s++)
        for (uint32_t i : sn[s].own) front_of[i] = (uint32_t)s;
    for (uint32_t i = 0; i < nf; i++) {
        const uint32_t s = front_of[i];
        scat.push_back(Scat{i, 0, s, opos[i] - fr[s].own0, opos[i] - fr[s].own0});
    }
    for (size_t pi = 0; pi < pairs.size(); pi++) {
        const uint32_t rnode = pairs[pi].x, cnode = pairs[pi].y;
        const uint32_t s = front_of[cnode];
        const Front& x = fr[s];
        const uint32_t cpos = opos[cnode] - x.own0;
        uint32_t rpos;
        if (opos[rnode] < x.own0 + x.ns) rpos = opos[rnode] - x.own0;
        else {
            const auto it = std::lower_bound(rows.begin() + x.row0, rows.begin() + x.row0 + x.nr, opos[rnode]);
            rpos = x.ns + (uint32_t)(it - (rows.begin() + x.row0));
        }
        scat.push_back(Scat{(uint32_t)pi, 1, s, rpos, cpos});
    }
    // residual adjacency: per free node its pairs, both directions
    std::vector<uint32_t> adj_off(nf + 1, 0);
    std::vector<uint2> adjl;
    {
        std::vector<std::vector<uint2>> per(nf);
        for (size_t pi = 0; pi < pairs.size(); pi++) {
            per[pairs[pi].x].push_back(make_uint2((uint32_t)pi, pairs[pi].y | 0x80000000u));
            per[pairs[pi].y].push_back(make_uint2((uint32_t)pi, pairs[pi].x));
        }
        for (uint32_t i = 0; i < nf; i++) {
            adjl.insert(adjl.end(), per[i].begin(), per[i].end());
            adj_off[i + 1] = (uint32_t)adjl.size();
        }
    }
    // levels and task lists
    uint32_t nlev = 0;
    for (const Front& x : fr) nlev = std::max(nlev, x.level + 1);
    f->levels.assign(nlev, Level());
    for (size_t s = 0; s < fr.size(); s++) {
        f->levels[fr[s].level].fronts.push_back((uint32_t)s);
        f->levels[fr[s].level].max_c = std::max(f->levels[fr[s].level].max_c, fr[s].c);
    }
    std::vector<EaTask> ea;
    std::vector<uint2> zt;
    for (uint32_t s = 0; s < (uint32_t)fr.size(); s++)
        for (uint32_t j0 = 0; j0 < fr[s].m; j0 += 16) zt.push_back(make_uint2(s, j0));
    std::vector<uint32_t> flist;
    std::vector<uint2> pfac;
    std::vector<uint4> syrk;
    std::vector<uint2> fs_tasks, bs_tasks;
    std::vector<uint32_t> children, ch_off(fr.size() + 1, 0);
    for (size_t s = 0; s < fr.size(); s++) {
        children.insert(children.end(), sn[s].children.begin(), sn[s].children.end());
        ch_off[s + 1] = (uint32_t)children.size();
    }
    f->plan.assign(nlev, bl_fem::LevelPlan());
    for (uint32_t L = 0; L < nlev; L++) {
        bl_fem::LevelPlan& lp = f->plan[L];
        const Level& lv = f->levels[L];
        lp.list_off = flist.size(), lp.nlist = (uint32_t)lv.fronts.size(), lp.max_m = 0;
        for (uint32_t s : lv.fronts) flist.push_back(s), lp.max_m = std::max(lp.max_m, fr[s].m);
        // extend-add rounds: the r-th child of every front of the level
        for (uint32_t rnd = 0;; rnd++) {
            const uint64_t o = ea.size();
            for (uint32_t s : lv.fronts) {
                if (rnd >= sn[s].children.size()) continue;
                const uint32_t ch = sn[s].children[rnd];
                const uint32_t M = 6 * fr[ch].nr;
                for (uint32_t r0 = 0; r0 < M; r0 += 64)
                    for (uint32_t c0 = 0; c0 <= r0; c0 += 64) ea.push_back(EaTask{ch, r0, c0, 0});
            }
            if (ea.size() == o) {
                bool any = false;
                for (uint32_t s : lv.fronts) any |= rnd < sn[s].children.size();
                if (!any) break;
            }
            lp.ea.push_back({o, (uint32_t)(ea.size() - o)});
        }
        lp.fs_off = fs_tasks.size(), lp.bs_off = bs_tasks.size();
        for (uint32_t s : lv.fronts) {
            for (uint32_t rb = 0; rb * 64 < fr[s].m; rb++) fs_tasks.push_back(make_uint2(s, rb));
            for (uint32_t jb = (fr[s].c + 63) / 64; jb-- > 0;) bs_tasks.push_back(make_uint2(s, jb));
        }
        lp.nfs = (uint32_t)(fs_tasks.size() - lp.fs_off), lp.nbs = (uint32_t)(bs_tasks.size() - lp.bs_off);
        // launch s: the panel tasks of panel k = 32 s, the update tasks of panel k - 32 (columns past panel k)
        for (uint32_t k = 0; k < lv.max_c + kT; k += kT) {
            bl_fem::Step stp{};
            stp.k = k;
            stp.pfac = pfac.size(), stp.syrk = syrk.size();
            for (uint32_t s : lv.fronts) {
                const Front& x = fr[s];
                if (k < x.c)
                    for (uint32_t r0 = k; r0 < x.m; r0 += 64) pfac.push_back(make_uint2(s, r0));
                if (k >= kT && k - kT < x.c) {
                    const uint32_t kp = k - kT, lop = kp + std::min<uint32_t>(kT, x.c - kp);
                    const uint32_t lo2 = lop < x.c ? lop + std::min<uint32_t>(kT, x.c - lop) : lop;
                    for (uint32_t c0 = lo2; c0 < x.m; c0 += 64)
                        for (uint32_t r0 = c0; r0 < x.m; r0 += 64) syrk.push_back(make_uint4(s, r0, c0, 0));
                }
            }
            stp.npfac = pfac.size() - stp.pfac, stp.nsyrk = syrk.size() - stp.syrk;
            if (stp.npfac + stp.nsyrk) lp.steps.push_back(stp);
        }
    }
    // ord2f: order position -> free index
    std::vector<float> p0f(3 * (size_t)nf);
    for (uint32_t i = 0; i < nf; i++)
        for (int k = 0; k < 3; k++) p0f[3 * i + k] = d->p[3 * freen[i] + k];
    std::vector<uint32_t> ord2f(order.begin(), order.end()), ford(nf);
    for (uint32_t i = 0; i < nf; i++) ford[i] = opos[i];
    f->hfronts = fr;
    // device arrays
    std::vector<double> xd(3 * (size_t)nn);
    for (uint32_t i = 0; i < 3 * nn; i++) xd[i] = d->p[i];
    std::vector<float4> q(nn, make_float4(0, 0, 0, 1));
    std::vector<uint8_t> fixed(nn, 0);
    if (d->fixed) std::memcpy(fixed.data(), d->fixed, nn);
    // the body's Node::p and inv_mass (ShellMesher::finish, src/phys/fem_shell.cpp:48-53: 0 for a pinned node)
    std::vector<float> pf(d->p, d->p + 3 * (size_t)nn), invm(nn);
    for (uint32_t i = 0; i < nn; i++) invm[i] = fixed[i] || mass[i] <= 0 ? 0.0f : 1.0f / mass[i];
    // the components (FemFrame::analyse, src/phys/frame_fem.cpp:557-566, 594-599: the frame's nodes joined by members
    // and triangles, in the order of their first node), each member's, their node counts
    std::vector<uint32_t> up(nn), cid(nn, UINT32_MAX), mcomp(nm);
    for (uint32_t i = 0; i < nn; i++) up[i] = i;
    auto find = [&](uint32_t x) {
        while (up[x] != x) x = up[x] = up[up[x]];
        return x;
    };
    for (uint32_t k = 0; k < nm; k++) up[find(mem[k].x)] = find(mem[k].y);
    for (uint32_t k = 0; k < nt; k++) {
        const uint32_t* n = &d->tris[3 * k];
        up[find(n[0])] = find(n[1]), up[find(n[1])] = find(n[2]);
    }
    f->comp_size.clear();
    for (uint32_t i = 0; i < nn; i++) {
        const uint32_t r = find(i);
        if (cid[r] == UINT32_MAX) cid[r] = (uint32_t)f->comp_size.size(), f->comp_size.push_back(0);
        f->comp_size[cid[r]]++;
    }
    f->ncomp = (uint32_t)f->comp_size.size();
    for (uint32_t k = 0; k < nm; k++) mcomp[k] = cid[find(mem[k].x)];
    static_assert(sizeof(bl_frame_section) == 17 * sizeof(float), "bl_frame_section is 17 floats");
    std::vector<float> msec(17 * (size_t)d->n_sections);
    if (d->n_sections) std::memcpy(msec.data(), d->sections, sizeof(float) * msec.size());
    f->plastic_members = false;
    for (uint32_t k = 0; k < nm; k++) f->plastic_members |= d->sections[d->members[k].section].Np > 0;
    const bool shell_yield = d->section.yield > 0;
    int rc;
    if ((rc = f->msec.put(msec)) || (rc = f->mps.put(std::vector<float>(16 * (size_t)nm, 0.0f))) ||
        (rc = f->mtan.alloc(17 * (size_t)nm)) || (rc = f->mbits.put(std::vector<uint32_t>(nm, 0u))) ||
        (rc = f->nyield.put(std::vector<uint32_t>(2, 0u))) || (rc = f->mcomp.put(mcomp)) ||
        (rc = f->again.put(std::vector<uint32_t>(f->ncomp, 0u))) || (rc = f->X0.put(X)) ||
        (rc = f->th0.put(std::vector<float>(9 * (size_t)nt, 0.0f))) || (rc = f->tdmg.put(std::vector<float>(nt, 0.0f))) ||
        (rc = f->tutil.put(std::vector<float>(nt, 0.0f))) || (rc = f->Bc.alloc(shell_yield ? 54 * (size_t)nt : 0)) ||
        (rc = f->jsave.alloc(3 * (size_t)nn)) || (rc = f->pred.put(std::vector<float>(3 * (size_t)nn, 0.0f))) ||
        (rc = f->tpred.put(std::vector<float>(3 * (size_t)nn, 0.0f))))
        return rc;
    if ((rc = f->mem.put(mem)) || (rc = f->mL0.put(mL0)) || (rc = f->mqa.put(mqa)) || (rc = f->mqb.put(mqb)) ||
        (rc = f->mst.put(mst)) || (rc = f->mtw.put(mtw)) || (rc = f->mdamp.put(mdamp)) || (rc = f->mon.alloc(nm)) ||
        (rc = f->mfo.alloc(9 * (size_t)nm)) || (rc = f->me1.alloc(nm)) || (rc = f->mW.alloc(144 * (size_t)nm)) ||
        (rc = f->mkv.alloc(12 * (size_t)nm)) || (rc = f->torq.put(std::vector<float>(3 * (size_t)nn, 0.0f))))
        return rc;
    if ((rc = f->pf.put(pf)) || (rc = f->invm.put(invm)) || (rc = f->fnode.put(std::vector<float>(3 * (size_t)nn, 0.0f))) ||
        (rc = f->imp.put(std::vector<float>(3 * (size_t)nn, 0.0f))) ||
        (rc = f->cnt.put(std::vector<unsigned long long>(3, 0ull))))
        return rc;
    if ((rc = f->xd.put(xd)) || (rc = f->v.put(std::vector<float>(3 * (size_t)nn, 0.0f))) ||
        (rc = f->w.put(std::vector<float>(3 * (size_t)nn, 0.0f))) || (rc = f->q.put(q)) || (rc = f->mass.put(mass)) ||
        (rc = f->inertia.put(inertia)) || (rc = f->fixed.put(fixed)) || (rc = f->tri.put(tri)) || (rc = f->X.put(X)) ||
        (rc = f->r0.put(r0)) || (rc = f->Kf.alloc(162 * (size_t)nt)) || (rc = f->Kd.alloc(324 * (size_t)nt)) ||
        (rc = f->te.alloc(9 * (size_t)nt)) || (rc = f->tfo.alloc(9 * (size_t)nt)) || (rc = f->tto.alloc(9 * (size_t)nt)) ||
        (rc = f->tg.alloc(9 * (size_t)nt)) || (rc = f->tkv.alloc(18 * (size_t)nt)) || (rc = f->ton.alloc(nt)) ||
        (rc = f->nc_off.put(nc_off)) || (rc = f->nc.put(nc)) || (rc = f->fidx.put(fidx)) || (rc = f->ford.put(ford)) ||
        (rc = f->ord2f.put(ord2f)) || (rc = f->Dsys.alloc(36 * (size_t)nf)) || (rc = f->rhs.alloc(6 * (size_t)nf)) ||
        (rc = f->Osys.alloc(36 * pairs.size())) || (rc = f->pairs.put(pairs)) || (rc = f->pcon.put(pcon)) ||
        (rc = f->y.alloc(6 * (size_t)nf)) || (rc = f->x.put(std::vector<double>(6 * (size_t)nf, 0.0))) ||
        (rc = f->xprev.put(std::vector<double>(6 * (size_t)nf, 0.0))) || (rc = f->p0f.put(p0f)) || (rc = f->loads.alloc(3 * (size_t)nn)) ||
        (rc = f->rows.put(rows)) || (rc = f->cmap.put(cmap)) || (rc = f->children.put(children)) ||
        (rc = f->ch_off.put(ch_off)) || (rc = f->fronts.put(fr)) || (rc = f->F.alloc(off)) || (rc = f->V.alloc(voff)) ||
        (rc = f->Xs.alloc(voff)) || (rc = f->fflag.put(std::vector<uint32_t>(fb, 0u))) ||
        (rc = f->bflag.put(std::vector<uint32_t>(fb, 0u))) || (rc = f->epoch.put(std::vector<uint32_t>(1, 0u))) ||
        (rc = f->fs_tasks.put(fs_tasks)) || (rc = f->bs_tasks.put(bs_tasks)) ||
        (rc = f->scat.put(scat)) || (rc = f->ea_tasks.put(ea)) || (rc = f->ztasks.put(zt)) || (rc = f->pfac_tasks.put(pfac)) || (rc = f->syrk_tasks.put(syrk)) ||
        (rc = f->flist.put(flist)) || (rc = f->Lbuf.alloc(loff)) ||
        (rc = f->adj_off.put(adj_off)) || (rc = f->adj.put(adjl)) || (rc = f->fail.alloc(1)))
        return rc;
    f->nparts = (nf + 255) / 256;
    if ((rc = f->part.alloc(2 * (size_t)f->nparts + 2)) || (rc = f->sums.alloc(2))) return rc;
    FemArgs a = f->args(nullptr, 0, 0, 0, nullptr);
    if (nt) hipLaunchKernelGGL(k_tri_k, dim3((nt + 63) / 64), dim3(64), 0, f->st, a);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(f->st));
    return BL_OK;
}

// L L^T x = b (b on the device, 6 per free index; x into `out`, f->x when null): forward by levels from the leaves,
// backward from the root. With `skip`, every launch returns at once when *skip is set on the device.
void enqueue_solve(bl_fem* f, const double* b, double* out = nullptr, const int* skip = nullptr) {
    const hipStream_t st = f->st;
    SolArgs so = f->sol(b);
    if (out) so.x = out;
    so.skip = skip;
    hipLaunchKernelGGL(k_epoch, dim3(1), dim3(64), 0, st, f->epoch.p);
    for (size_t L = 0; L < f->plan.size(); L++) {
        const bl_fem::LevelPlan& lp = f->plan[L];
        hipLaunchKernelGGL(k_fgather, dim3(lp.nlist), dim3(256), 0, st, so, f->flist.p + lp.list_off, lp.nlist);
        hipLaunchKernelGGL(k_fsolve, dim3(lp.nfs), dim3(256), 0, st, so, f->fs_tasks.p + lp.fs_off, lp.nfs);
    }
    for (size_t L = f->plan.size(); L-- > 0;) {
        const bl_fem::LevelPlan& lp = f->plan[L];
        hipLaunchKernelGGL(k_bsolve, dim3(lp.nbs), dim3(256), 0, st, so, f->bs_tasks.p + lp.bs_off, lp.nbs);
    }
}

int fem_factor_solve(bl_fem* f, float* ms_factor, float* ms_solve) {
    FacArgs fa = f->fac();
    const hipStream_t st = f->st;
    FCHK(hipEventRecord(f->ev[1], st));
    if (f->ztasks.n)
        hipLaunchKernelGGL(k_zero_lower, dim3((uint32_t)f->ztasks.n), dim3(256), 0, st, f->fronts.p, f->ztasks.p, (uint32_t)f->ztasks.n, f->F.p);
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, f->fail.p, 1u);
    FemArgs a = f->args(nullptr, 0, 0, 0, nullptr);
    const uint32_t nsc = (uint32_t)f->scat.n;
    if (nsc) hipLaunchKernelGGL(k_scatter, dim3((nsc + 3) / 4), dim3(256), 0, st, a, fa, f->scat.p, nsc);
    for (size_t L = 0; L < f->plan.size(); L++) {
        const bl_fem::LevelPlan& lp = f->plan[L];
        for (const auto& r : lp.ea)
            if (r.second) hipLaunchKernelGGL(k_extend_add, dim3(r.second), dim3(256), 0, st, fa, f->ea_tasks.p + r.first, r.second, f->cmap.p);
        for (const bl_fem::Step& s : lp.steps) {
            hipLaunchKernelGGL(k_fstep, dim3((uint32_t)(s.npfac + s.nsyrk)), dim3(256), 0, st, fa, f->pfac_tasks.p + s.pfac,
                               (uint32_t)s.npfac, f->syrk_tasks.p + s.syrk, (uint32_t)s.nsyrk, s.k);
        }
    }
    FCHK(hipGetLastError());
    FCHK(hipEventRecord(f->ev[2], st));
    // substitutions: forward bottom-up, backward top-down
    enqueue_solve(f, f->rhs.p);
    FCHK(hipGetLastError());
    FCHK(hipEventRecord(f->ev[3], st));
    FCHK(hipEventSynchronize(f->ev[3]));
    if (ms_factor) FCHK(hipEventElapsedTime(ms_factor, f->ev[1], f->ev[2]));
    if (ms_solve) FCHK(hipEventElapsedTime(ms_solve, f->ev[2], f->ev[3]));
    return BL_OK;
}

} // namespace

namespace blf {

FemView fem_view(bl_fem* f) {
    FemView v{};
    v.nf = f->nf, v.D = f->Dsys.p, v.O = f->Osys.p, v.pairs = f->pairs.p, v.npairs = f->npairs;
    v.adj_off = f->adj_off.p, v.adj = f->adj.p, v.rhs = f->rhs.p, v.x = f->x.p, v.xprev = f->xprev.p;
    v.p0 = f->p0f.p, v.st = f->st, v.serial = f->serial;
    return v;
}

int fem_assemble(bl_fem* f, const float* loads, float h, float theta, float dissipation, const float g[3], float* ms,
                 float step) {
    const hipStream_t st = f->st;
    const float* dl = nullptr;
    // (theta and the dissipation as FemFrame::solve_begin takes them, src/phys/frame_fem.cpp:2524-2525)
    dissipation = std::min(1.0f, std::max(0.0f, dissipation));
    theta = std::min(1.0f, std::max(0.25f + 0.5f * dissipation, theta));
    if (loads) {
        FCHK(hipMemcpyAsync(f->loads.p, loads, sizeof(float) * 3 * f->nn, hipMemcpyHostToDevice, st));
        dl = f->loads.p;
    }
    FCHK(hipMemcpyAsync(f->xprev.p, f->x.p, sizeof(double) * 6 * f->nf, hipMemcpyDeviceToDevice, st));
    FemArgs a = f->args(dl, h, theta, dissipation, g);
    if (step > 0) a.step = step;
    // the passes-again budget (FemFrame::begin_forces, src/phys/frame_fem.cpp:1863-1864): one more every fourth step,
    // at most 8; with less than one left the yielding members start from their elastic tangents
    f->repass_credit = std::min(8.0f, f->repass_credit + 0.25f);
    a.elastic_step = f->repass_credit < 1.0f ? 1 : 0;
    a.pass = 0;
    f->step_args = a;
    FCHK(hipEventRecord(f->ev[0], st));
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, (int*)f->nyield.p, 2u);
    if (f->nm) hipLaunchKernelGGL(k_members, dim3((f->nm + 63) / 64), dim3(64), 0, st, a);
    if (f->nt && f->sec.yield > 0) hipLaunchKernelGGL(k_tris<true>, dim3((f->nt + 63) / 64), dim3(64), 0, st, a);
    else if (f->nt) hipLaunchKernelGGL(k_tris<false>, dim3((f->nt + 63) / 64), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_nodes, dim3((f->nn + 63) / 64), dim3(64), 0, st, a);
    if (f->npairs) hipLaunchKernelGGL(k_pairs, dim3((f->npairs + 63) / 64), dim3(64), 0, st, a);
    FCHK(hipGetLastError());
    FCHK(hipEventRecord(f->ev[1], st));
    f->have_system = true;
    if (ms) {
        FCHK(hipEventSynchronize(f->ev[1]));
        FCHK(hipEventElapsedTime(ms, f->ev[0], f->ev[1]));
    }
    return BL_OK;
}

// solve_impl's passes again (src/phys/frame_fem.cpp:2959-2996) after the first pass's solve: while a member that
// yielded at the step's start is unloaded by the pass's solution (k_mem_repass: its elastic tangent, its matrix
// again), the system is assembled, factored and solved again, at most three passes; a pass again of a component of
// 60 nodes or more spends the budget. *passes counts the systems solved, *repasses the components' passes again.
int fem_repasses(bl_fem* f, float* ms_factor, float* ms_solve, int* fail, int* passes, int* repasses) {
    const hipStream_t st = f->st;
    uint32_t ny[2];
    FCHK(hipMemcpy(ny, f->nyield.p, sizeof(ny), hipMemcpyDeviceToHost));
    if (!ny[0]) return BL_OK;
    FemArgs a = f->step_args;
    std::vector<uint32_t> again(f->ncomp);
    for (int pass = 1; pass <= 2; pass++) {
        hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(256), 0, st, (int*)f->again.p, f->ncomp);
        hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, (int*)f->nyield.p + 1, 1u);
        hipLaunchKernelGGL(k_mem_repass, dim3((f->nm + 63) / 64), dim3(64), 0, st, a, (const double*)f->x.p);
        FCHK(hipGetLastError());
        FCHK(hipMemcpyAsync(ny, f->nyield.p, sizeof(ny), hipMemcpyDeviceToHost, st));
        FCHK(hipMemcpyAsync(again.data(), f->again.p, sizeof(uint32_t) * f->ncomp, hipMemcpyDeviceToHost, st));
        FCHK(hipStreamSynchronize(st));
        for (uint32_t c = 0; c < f->ncomp; c++)
            if (again[c] && f->comp_size[c] >= 60) (*repasses)++, f->repass_credit -= 1.0f;
        if (!ny[1]) break;
        a.pass = pass;
        hipLaunchKernelGGL(k_nodes, dim3((f->nn + 63) / 64), dim3(64), 0, st, a);
        if (f->npairs) hipLaunchKernelGGL(k_pairs, dim3((f->npairs + 63) / 64), dim3(64), 0, st, a);
        FCHK(hipGetLastError());
        (*passes)++;
        float mf = 0, ms = 0;
        int rc = fem_cholesky(f, &mf, &ms, fail);
        if (rc) return rc;
        if (ms_factor) *ms_factor += mf;
        if (ms_solve) *ms_solve += ms;
        if (*fail) break;
    }
    return BL_OK;
}

bool fem_plastic(const bl_fem* f) { return f->plastic_members || f->sec.yield > 0; }

int fem_cholesky(bl_fem* f, float* ms_factor, float* ms_solve, int* fail) {
    f->factor_ok = false;
    int rc = fem_factor_solve(f, ms_factor, ms_solve);
    if (rc) return rc;
    FCHK(hipMemcpy(fail, f->fail.p, sizeof(int), hipMemcpyDeviceToHost));
    f->factor_ok = *fail == 0;
    return BL_OK;
}

bool fem_has_factor(bl_fem* f) { return f->factor_ok; }

int fem_keep_factor(bl_fem* f) {
    f->kept = false;
    if (!f->factor_ok || !f->Lbuf.n) return BL_OK;
    int rc;
    if (f->Lkeep.n != f->Lbuf.n && (rc = f->Lkeep.alloc(f->Lbuf.n))) return rc;
    FCHK(hipMemcpyAsync(f->Lkeep.p, f->Lbuf.p, sizeof(double) * f->Lbuf.n, hipMemcpyDeviceToDevice, f->st));
    f->kept = true;
    return BL_OK;
}

int fem_restore_factor(bl_fem* f) {
    if (!f->kept) return BL_OK;
    FCHK(hipMemcpyAsync(f->Lbuf.p, f->Lkeep.p, sizeof(double) * f->Lbuf.n, hipMemcpyDeviceToDevice, f->st));
    f->kept = false;
    f->factor_ok = true;
    return BL_OK;
}

int fem_precond(bl_fem* f, const double* r, double* z, const int* skip) {
    enqueue_solve(f, r, z, skip);
    FCHK(hipGetLastError());
    return BL_OK;
}

int fem_solve_error(bl_fem* f, int* err) {
    FCHK(hipMemcpyAsync(err, f->fail.p, sizeof(int), hipMemcpyDeviceToHost, f->st));
    FCHK(hipStreamSynchronize(f->st));
    return BL_OK;
}

int fem_residual(bl_fem* f, const double* x, double* rel) {
    const hipStream_t st = f->st;
    hipLaunchKernelGGL(k_residual, dim3(f->nparts), dim3(256), 0, st, f->step_args, x, f->adj_off.p, f->adj.p, f->part.p);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, f->part.p, f->nparts, f->sums.p);
    double sums[2] = {0, 0};
    FCHK(hipMemcpyAsync(sums, f->sums.p, sizeof(sums), hipMemcpyDeviceToHost, st));
    FCHK(hipStreamSynchronize(st));
    *rel = sums[1] > 0 ? std::sqrt(sums[0] / sums[1]) : std::sqrt(sums[0]);
    return BL_OK;
}

void fem_spmv(bl_fem* f, const double* x, double* y) {
    hipLaunchKernelGGL(k_spmv, dim3((f->nf + 255) / 256), dim3(256), 0, f->st, f->step_args, x, f->adj_off.p, f->adj.p, y);
}

int fem_update(bl_fem* f, float* ms, bool explicit_step) {
    const hipStream_t st = f->st;
    FCHK(hipEventRecord(f->ev[3], st));
    hipLaunchKernelGGL(k_update, dim3((f->nn + 255) / 256), dim3(256), 0, st, f->step_args, f->x.p, explicit_step ? 1 : 0);
    FCHK(hipGetLastError());
    FCHK(hipEventRecord(f->ev[4], st));
    FCHK(hipEventSynchronize(f->ev[4]));
    if (ms) FCHK(hipEventElapsedTime(ms, f->ev[3], f->ev[4]));
    return BL_OK;
}

void fem_report_static(bl_fem* f, bl_fem_report* rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->factor_flops = f->flops;
    rep->supernodes = (int32_t)f->hfronts.size();
    rep->fronts_max = f->fronts_max;
    rep->factor_nnz = f->nnz;
}

} // namespace blf

// ---------------------------------------------------------------------------------------------- the no-allocation solve
// The substitutions' scratch of bl_fem_solve_noalloc: a file-scope device array, one copy per GPU, held until the
// process ends. Layout: [epoch u32, error word i32, pad to 256 B][forward flags u32 x nflag][backward flags u32 x
// nflag][pad to 256 B][y f64 x 6 nf][V f64 x nv][X f64 x nv]. The flags hold the epoch of the solve that published
// them, so stale ones of earlier calls (of any engine) never read as published and need no clearing; the error word
// is cleared by a kernel at each call.
constexpr size_t kNoAllocScratchBytes = 32u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (hipGetSymbolAddress once per device, with it current).
int scratch_address(void** out) {
    int dev = 0;
    FCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        blg::set_last_error("device index out of range for the no-allocation scratch");
        return BL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        FCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) {
            blg::set_last_error("hipGetSymbolAddress returned no address for the scratch array");
            return BL_ERR_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return BL_OK;
}

// the scratch bytes the engine's substitutions need
size_t noalloc_bytes(const bl_fem* f) {
    const size_t head = 256, flags = ((2 * f->fflag.n * sizeof(uint32_t) + 255) / 256) * 256;
    return head + flags + sizeof(double) * (6 * (size_t)f->nf + 2 * f->V.n);
}

} // namespace

extern "C" {

int bl_fem_create(bl_fem** out, const bl_fem_desc* d) {
    if (!out || !d || (d->n_nodes && !d->p) || (d->n_tris && !d->tris) || (d->n_members && (!d->members || !d->sections)))
        return BL_ERR_ARG;
    *out = nullptr;
    if (d->n_nodes == 0 || (d->n_tris == 0 && d->n_members == 0)) return BL_ERR_ARG;
    if (d->n_members && !d->mass) {
        blg::set_last_error("a body with members passes its node masses");
        return BL_ERR_ARG;
    }
    bl_fem* f = new bl_fem();
    static std::atomic<uint64_t> next_serial{1};
    f->serial = next_serial.fetch_add(1);
    if (hipStreamCreate(&f->st) != hipSuccess) {
        delete f;
        return BL_ERR_HIP;
    }
    for (hipEvent_t& e : f->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete f;
            return BL_ERR_HIP;
        }
    int rc = fem_build(f, d);
    void* sp = nullptr;
    if (!rc) rc = scratch_address(&sp); // (the no-allocation solve's scratch on this device, once)
    if (rc) {
        delete f;
        return rc;
    }
    *out = f;
    return BL_OK;
}

void bl_fem_destroy(bl_fem* f) { delete f; }

int bl_fem_step(bl_fem* f, const float* loads, float h, float theta, float dissipation, const float g[3],
                bl_fem_report* rep) {
    return bl_fem_step_span(f, loads, h, h, theta, dissipation, g, rep);
}

int bl_fem_step_span(bl_fem* f, const float* loads, float h, float step, float theta, float dissipation,
                     const float g[3], bl_fem_report* rep) {
    if (!f || !g || !(h > 0) || !(step > 0) || !(step <= h)) return BL_ERR_ARG;
    float msa = 0, msf = 0, mss = 0, msu = 0;
    int fail = 0, rc;
    double res = 0;
    if ((rc = blf::fem_assemble(f, loads, h, theta, dissipation, g, &msa, step))) return rc;
    if ((rc = blf::fem_cholesky(f, &msf, &mss, &fail))) return rc;
    int passes = 1, repasses = 0;
    if (!fail && f->plastic_members && (rc = blf::fem_repasses(f, &msf, &mss, &fail, &passes, &repasses))) return rc;
    f->passes += passes, f->repasses += repasses;
    if ((rc = blf::fem_residual(f, f->x.p, &res))) return rc;
    if (fail & 2) { // (the device's own trouble, not the system's: the state is left as it was)
        blg::set_last_error("a substitution's wait for a block timed out");
        return BL_ERR_STATE;
    }
    // a pivot not positive: the gathered forces act explicitly this once, counted (the reference's solve_failures)
    if ((rc = blf::fem_update(f, &msu, (fail & 1) != 0))) return rc;
    if (rep) {
        blf::fem_report_static(f, rep);
        rep->residual = res;
        rep->assemble_ms = msa, rep->factor_ms = msf, rep->solve_ms = mss, rep->update_ms = msu;
        rep->failures = fail ? 1 : 0;
        rep->passes = passes, rep->repasses = repasses;
    }
    return BL_OK;
}

int bl_fem_step_held(bl_fem* f, const float* loads, float step, const float g[3]) {
    if (!f || !g || !(step > 0)) return BL_ERR_ARG;
    const hipStream_t st = f->st;
    const float* dl = nullptr;
    if (loads) {
        FCHK(hipMemcpyAsync(f->loads.p, loads, sizeof(float) * 3 * f->nn, hipMemcpyHostToDevice, st));
        dl = f->loads.p;
    }
    if (f->hrhs.n != 6 * (size_t)f->nf) {
        int rc;
        if ((rc = f->hrhs.alloc(6 * (size_t)f->nf)) || (rc = f->hx.alloc(6 * (size_t)f->nf))) return rc;
    }
    FemArgs a = f->args(dl, step, 0, 0, g);
    const bool factored = blf::fem_has_factor(f);
    const uint32_t nb = (f->nn + 255) / 256;
    if (factored) {
        hipLaunchKernelGGL(blf::k_held_rhs, dim3(nb), dim3(256), 0, st, a, f->hrhs.p);
        hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, f->fail.p, 1u);
        enqueue_solve(f, f->hrhs.p, f->hx.p);
    } else {
        hipLaunchKernelGGL(blf::k_held_rhs, dim3(nb), dim3(256), 0, st, a, f->hrhs.p); // (the forces alone)
    }
    hipLaunchKernelGGL(blf::k_held_update, dim3(nb), dim3(256), 0, st, a, (const double*)f->hx.p, factored ? 1 : 0);
    FCHK(hipGetLastError());
    if (factored) {
        int err = 0;
        int rc = blf::fem_solve_error(f, &err);
        if (rc) return rc;
        if (err) {
            blg::set_last_error("a held substep's substitution timed out");
            return BL_ERR_STATE;
        }
    } else {
        FCHK(hipStreamSynchronize(st));
    }
    return BL_OK;
}

int bl_fem_get_plastic(bl_fem* f, float* tris, float* members, int64_t counters[2]) {
    if (!f) return BL_ERR_ARG;
    FCHK(hipStreamSynchronize(f->st));
    const size_t nt = f->nt, nm = f->nm;
    if (tris && nt) {
        std::vector<float> X(6 * nt), th0(9 * nt), dmg(nt), util(nt);
        std::vector<float4> r0(3 * nt);
        FCHK(hipMemcpy(X.data(), f->X.p, sizeof(float) * X.size(), hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(th0.data(), f->th0.p, sizeof(float) * th0.size(), hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(dmg.data(), f->tdmg.p, sizeof(float) * nt, hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(util.data(), f->tutil.p, sizeof(float) * nt, hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(r0.data(), f->r0.p, sizeof(float4) * r0.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nt; k++) {
            float* o = tris + BL_FEM_TRI_STATE * k;
            std::memcpy(o, &X[6 * k], 6 * sizeof(float));
            std::memcpy(o + 6, &th0[9 * k], 9 * sizeof(float));
            for (int i = 0; i < 3; i++) {
                const float4 r = r0[3 * k + i];
                o[15 + 4 * i] = r.x, o[16 + 4 * i] = r.y, o[17 + 4 * i] = r.z, o[18 + 4 * i] = r.w;
            }
            o[27] = dmg[k], o[28] = util[k];
        }
    }
    if (members && nm) {
        std::vector<float> ps(16 * nm);
        std::vector<float4> qa(nm), qb(nm);
        std::vector<uint32_t> bits(nm);
        FCHK(hipMemcpy(ps.data(), f->mps.p, sizeof(float) * ps.size(), hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(qa.data(), f->mqa.p, sizeof(float4) * nm, hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(qb.data(), f->mqb.p, sizeof(float4) * nm, hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(bits.data(), f->mbits.p, sizeof(uint32_t) * nm, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nm; k++) {
            float* o = members + BL_FEM_MEM_STATE * k;
            std::memcpy(o, &ps[16 * k], 16 * sizeof(float));
            o[16] = qa[k].x, o[17] = qa[k].y, o[18] = qa[k].z, o[19] = qa[k].w;
            o[20] = qb[k].x, o[21] = qb[k].y, o[22] = qb[k].z, o[23] = qb[k].w;
            o[24] = (float)(bits[k] & 0xffu), o[25] = (float)((bits[k] >> 8) & 0xffu), o[26] = (float)((bits[k] >> 16) & 0xffu);
        }
    }
    if (counters) counters[0] = f->repasses, counters[1] = f->passes;
    return BL_OK;
}

int bl_fem_get_state(bl_fem* f, double* x, float* v, float* w, float* q) {
    if (!f) return BL_ERR_ARG;
    FCHK(hipStreamSynchronize(f->st));
    if (x) FCHK(hipMemcpy(x, f->xd.p, sizeof(double) * 3 * f->nn, hipMemcpyDeviceToHost));
    if (v) FCHK(hipMemcpy(v, f->v.p, sizeof(float) * 3 * f->nn, hipMemcpyDeviceToHost));
    if (w) FCHK(hipMemcpy(w, f->w.p, sizeof(float) * 3 * f->nn, hipMemcpyDeviceToHost));
    if (q) FCHK(hipMemcpy(q, f->q.p, sizeof(float) * 4 * f->nn, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_fem_set_state(bl_fem* f, const double* x, const float* v, const float* w, const float* q) {
    if (!f) return BL_ERR_ARG;
    FCHK(hipStreamSynchronize(f->st));
    if (x) {
        FCHK(hipMemcpy(f->xd.p, x, sizeof(double) * 3 * f->nn, hipMemcpyHostToDevice));
        std::vector<float> pf(3 * (size_t)f->nn); // (the body's float positions follow: (float)x == p)
        for (size_t i = 0; i < pf.size(); i++) pf[i] = (float)x[i];
        FCHK(hipMemcpy(f->pf.p, pf.data(), sizeof(float) * pf.size(), hipMemcpyHostToDevice));
    }
    if (v) FCHK(hipMemcpy(f->v.p, v, sizeof(float) * 3 * f->nn, hipMemcpyHostToDevice));
    if (w) FCHK(hipMemcpy(f->w.p, w, sizeof(float) * 3 * f->nn, hipMemcpyHostToDevice));
    if (q) FCHK(hipMemcpy(f->q.p, q, sizeof(float) * 4 * f->nn, hipMemcpyHostToDevice));
    return BL_OK;
}

int bl_fem_add_torques(bl_fem* f, const float* torques) {
    if (!f || !torques) return BL_ERR_ARG;
    std::vector<float> t(3 * (size_t)f->nn);
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipMemcpy(t.data(), f->torq.p, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < t.size(); i++) t[i] += torques[i];
    FCHK(hipMemcpy(f->torq.p, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
    return BL_OK;
}

int bl_fem_get_nodes(bl_fem* f, bl_node* nodes, float* force, int64_t counters[3]) {
    if (!f) return BL_ERR_ARG;
    FCHK(hipStreamSynchronize(f->st));
    if (nodes) {
        std::vector<float> p(3 * (size_t)f->nn), v(3 * (size_t)f->nn), im(f->nn), m(f->nn);
        FCHK(hipMemcpy(p.data(), f->pf.p, sizeof(float) * p.size(), hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(v.data(), f->v.p, sizeof(float) * v.size(), hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(im.data(), f->invm.p, sizeof(float) * im.size(), hipMemcpyDeviceToHost));
        FCHK(hipMemcpy(m.data(), f->mass.p, sizeof(float) * m.size(), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < f->nn; i++) {
            bl_node& n = nodes[i];
            n.p[0] = p[3 * i], n.p[1] = p[3 * i + 1], n.p[2] = p[3 * i + 2], n.inv_mass = im[i];
            n.v[0] = v[3 * i], n.v[1] = v[3 * i + 1], n.v[2] = v[3 * i + 2], n.mass = m[i];
        }
    }
    if (force) FCHK(hipMemcpy(force, f->fnode.p, sizeof(float) * 3 * f->nn, hipMemcpyDeviceToHost));
    if (counters) {
        unsigned long long c[3];
        FCHK(hipMemcpy(c, f->cnt.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int k = 0; k < 3; k++) counters[k] = (int64_t)c[k];
    }
    return BL_OK;
}

int bl_fem_last_system(bl_fem* f, double* A, double* rhs, double* x, uint32_t* nfo) {
    if (!f) return BL_ERR_ARG;
    if (!f->have_system) return BL_ERR_STATE;
    FCHK(hipStreamSynchronize(f->st));
    const uint32_t nf = f->nf;
    if (nfo) *nfo = nf;
    if (A) {
        std::vector<double> D(36 * (size_t)nf), O(36 * (size_t)f->npairs);
        std::vector<uint4> pairs(f->npairs);
        FCHK(hipMemcpy(D.data(), f->Dsys.p, sizeof(double) * D.size(), hipMemcpyDeviceToHost));
        if (f->npairs) {
            FCHK(hipMemcpy(O.data(), f->Osys.p, sizeof(double) * O.size(), hipMemcpyDeviceToHost));
            FCHK(hipMemcpy(pairs.data(), f->pairs.p, sizeof(uint4) * pairs.size(), hipMemcpyDeviceToHost));
        }
        const size_t N = 6 * (size_t)nf;
        std::memset(A, 0, sizeof(double) * N * N);
        for (uint32_t i = 0; i < nf; i++)
            for (int p = 0; p < 6; p++)
                for (int qq = 0; qq < 6; qq++) {
                    const double vv = p >= qq ? D[36 * (size_t)i + p * 6 + qq] : D[36 * (size_t)i + qq * 6 + p];
                    A[(6 * (size_t)i + p) * N + 6 * i + qq] = vv;
                }
        for (uint32_t k = 0; k < f->npairs; k++) {
            const uint32_t r = pairs[k].x, c = pairs[k].y;
            for (int p = 0; p < 6; p++)
                for (int qq = 0; qq < 6; qq++) {
                    A[(6 * (size_t)r + p) * N + 6 * c + qq] = O[36 * (size_t)k + p * 6 + qq];
                    A[(6 * (size_t)c + qq) * N + 6 * r + p] = O[36 * (size_t)k + p * 6 + qq];
                }
        }
    }
    if (rhs) FCHK(hipMemcpy(rhs, f->rhs.p, sizeof(double) * 6 * nf, hipMemcpyDeviceToHost));
    if (x) FCHK(hipMemcpy(x, f->x.p, sizeof(double) * 6 * nf, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_fem_solve(bl_fem* f, const double* b, double* x) {
    if (!f || !b || !x) return BL_ERR_ARG;
    if (!f->have_system) return BL_ERR_STATE;
    const hipStream_t st = f->st;
    Dv<double> bd;
    std::vector<double> hb(b, b + 6 * (size_t)f->nf);
    int rc;
    if ((rc = bd.put(hb))) return rc;
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, f->fail.p, 1u);
    enqueue_solve(f, bd.p);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(st));
    int fail = 0;
    FCHK(hipMemcpy(&fail, f->fail.p, sizeof(int), hipMemcpyDeviceToHost));
    if (fail) {
        blg::set_last_error("a substitution's wait for a block timed out");
        return BL_ERR_STATE;
    }
    FCHK(hipMemcpy(x, f->x.p, sizeof(double) * 6 * f->nf, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_fem_solve_noalloc(bl_fem* f, const double* b, double* x, uint64_t* scratch_bytes) {
    if (!f) return BL_ERR_ARG;
    const size_t need = noalloc_bytes(f);
    if (scratch_bytes) *scratch_bytes = need;
    if (!b || !x) return scratch_bytes ? BL_OK : BL_ERR_ARG;
    if (need > kNoAllocScratchBytes) {
        blg::set_last_error(("the system needs " + std::to_string(need) + " bytes of substitution scratch; the "
                             "no-allocation solve holds " + std::to_string(kNoAllocScratchBytes)).c_str());
        return BL_ERR_ARG;
    }
    if (!f->have_system) return BL_ERR_STATE;
    void* base = nullptr;
    int rc;
    if ((rc = scratch_address(&base))) return rc;
    unsigned char* sp = static_cast<unsigned char*>(base);
    const size_t nflag = f->fflag.n;
    const hipStream_t st = f->st;
    SolArgs so = f->sol(b);
    uint32_t* epoch = reinterpret_cast<uint32_t*>(sp);
    so.epoch = epoch;
    so.err = reinterpret_cast<int*>(sp + 4);
    so.fflag = reinterpret_cast<uint32_t*>(sp + 256);
    so.bflag = so.fflag + nflag;
    double* dp = reinterpret_cast<double*>(sp + 256 + ((2 * nflag * sizeof(uint32_t) + 255) / 256) * 256);
    so.y = dp;
    so.V = dp + 6 * (size_t)f->nf;
    so.X = so.V + f->V.n;
    so.x = x;
    so.skip = nullptr;
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, so.err, 1u);
    hipLaunchKernelGGL(k_epoch, dim3(1), dim3(64), 0, st, epoch);
    for (size_t L = 0; L < f->plan.size(); L++) {
        const bl_fem::LevelPlan& lp = f->plan[L];
        hipLaunchKernelGGL(k_fgather, dim3(lp.nlist), dim3(256), 0, st, so, f->flist.p + lp.list_off, lp.nlist);
        hipLaunchKernelGGL(k_fsolve, dim3(lp.nfs), dim3(256), 0, st, so, f->fs_tasks.p + lp.fs_off, lp.nfs);
    }
    for (size_t L = f->plan.size(); L-- > 0;) {
        const bl_fem::LevelPlan& lp = f->plan[L];
        hipLaunchKernelGGL(k_bsolve, dim3(lp.nbs), dim3(256), 0, st, so, f->bs_tasks.p + lp.bs_off, lp.nbs);
    }
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(st));
    int fail = 0;
    FCHK(hipMemcpy(&fail, so.err, sizeof(int), hipMemcpyDeviceToHost));
    if (fail) {
        blg::set_last_error("a substitution's wait for a block timed out");
        return BL_ERR_STATE;
    }
    return BL_OK;
}

int bl_fem_tri_k(bl_fem* f, uint32_t tri, double* K) {
    if (!f || !K || tri >= f->nt) return BL_ERR_ARG;
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipMemcpy(K, f->Kd.p + 324 * (size_t)tri, sizeof(double) * 324, hipMemcpyDeviceToHost));
    return BL_OK;
}

} // extern "C"
