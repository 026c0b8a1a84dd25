// The learned warm start and preconditioner of the implicit FEM solve: a 4-layer message-passing network over the
// mesh's free nodes, its training and inference on the bfloat16 matrix cores (FP32 accumulation), and the PCG solve
// it feeds, with the supernodal Cholesky as the fallback and the yardstick of its error. The PCG's preconditioner is
// block Jacobi, block Jacobi scaled by the network, or the engine's last Cholesky factor (factor reuse).
//
// The network (no counterpart in the reference; its math is stated once more, in NumPy, by the oracle):
//   features f_i (32 wide, 18 used): t_i / tau, xprev_i / sigma, b_i / beta, with t_i = D_i^-1 b_i (D_i the node's
//     6x6 diagonal block), tau / sigma / beta the root mean squares over the nodes of t / xprev / b, translations
//     (dof 0-2) and rotations (dof 3-5) apart; sigma falls back to tau where xprev is zero
//   h0 = relu(f W_e + b_e); 4 layers: P = h W_m, agg_i = mean_{j in N(i)} relu(P_j + rel_ij U + b_m),
//     h <- h + relu([h, agg] W_u + b_u); o = h W_d + b_d (16 wide)
//   warm start x0_i = sigma * o_i[0:6]; preconditioner S D^-1 S with s_i = exp(clip(o_i[6:12], -3, 3))
//   loss: mean (o[0:6] - x*/sigma)^2 + mean (k z - e)^2 with z = s * D^-1 (s * r0) / sigma, e = (x* - x0) / sigma,
//     k = <z, e> / <z, z> (the preconditioned residual against the warm start's error, up to the best global factor:
//     CG does not see a preconditioner's scale); r0, e, k constants; x* the Cholesky solution
// Products run on v_mfma_f32_16x16x16bf16_1k with both operands rounded to bfloat16 (nearest even) and FP32 sums.
// The instruction's lanes (measured on gfx950): lane l supplies A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16]
// (e = 0..3); item r of its result is D[4 (l / 16) + r][l % 16].
#include "bl_world.h"
#include "bl_fem_api.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#pragma clang fp contract(off)

namespace blnn {

constexpr int H = 64, FIN = 32, FOUT = 16, NL = 4;
constexpr uint32_t kChunk = 128; // rows of a weight-gradient partial

typedef short short4_ __attribute__((ext_vector_type(4)));
typedef float float4_ __attribute__((ext_vector_type(4)));

// parameter offsets (the oracle's param_shapes order)
struct Layout {
    uint32_t We, be, Wm[NL], U[NL], bm[NL], Wu[NL], bu[NL], Wd, bd, total;
    Layout() {
        uint32_t o = 0;
        We = o, o += FIN * H;
        be = o, o += H;
        for (int l = 0; l < NL; l++) {
            Wm[l] = o, o += H * H;
            U[l] = o, o += 4 * H;
            bm[l] = o, o += H;
            Wu[l] = o, o += 2 * H * H;
            bu[l] = o, o += H;
        }
        Wd = o, o += H * FOUT;
        bd = o, o += FOUT;
        total = o;
    }
};

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------------------------------------ Y = act(X W + b) [+ R]
struct GemmArgs {
    const float* X; uint32_t ldx;  // M x K, row-major; columns k >= k1 come from X2 (the concatenation [X, X2])
    const float* X2; uint32_t ldx2, k1;
    const float* W; uint32_t ldw; int wt; // element (k, n): wt ? W[n ldw + k] : W[k ldw + n]
    const float* bias;             // N, or null
    const float* R;                // M x N residual (ld ldy), or null
    float* Y; uint32_t ldy;        // M x N
    float* Z;                      // M x N pre-activation (ld ldy), or null
    uint32_t M, K, N;
    int relu;
};

// 64 rows a workgroup (16 a wave), all N (<= 128) columns, K <= 128
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ uint16_t Xs[64][128 + 4];
    __shared__ uint16_t Ws[128][128 + 4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, m0 = blockIdx.x * 64;
    for (uint32_t e = tid; e < 64 * g.K; e += 256) {
        const uint32_t r = e / g.K, k = e % g.K, m = m0 + r;
        float x = 0.0f;
        if (m < g.M) x = k < g.k1 ? g.X[(size_t)m * g.ldx + k] : g.X2[(size_t)m * g.ldx2 + (k - g.k1)];
        Xs[r][k] = f2bf(x);
    }
    for (uint32_t e = tid; e < g.K * g.N; e += 256) {
        const uint32_t k = e / g.N, n = e % g.N;
        Ws[n][k] = f2bf(g.wt ? g.W[(size_t)n * g.ldw + k] : g.W[(size_t)k * g.ldw + n]);
    }
    __syncthreads();
    float4_ acc[8];
    for (int t = 0; t < 8; t++) acc[t] = (float4_){0, 0, 0, 0};
    const uint32_t li = lane & 15, lk = 4 * (lane >> 4), nt = g.N / 16;
    for (uint32_t k0 = 0; k0 < g.K; k0 += 16) {
        short4_ a;
        for (int e = 0; e < 4; e++) a[e] = (short)Xs[16 * wv + li][k0 + lk + e];
        for (uint32_t t = 0; t < nt; t++) {
            short4_ b;
            for (int e = 0; e < 4; e++) b[e] = (short)Ws[16 * t + li][k0 + lk + e];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
    for (uint32_t t = 0; t < nt; t++)
        for (int r = 0; r < 4; r++) {
            const uint32_t m = m0 + 16 * wv + lk + r, n = 16 * t + li;
            if (m >= g.M) continue;
            float z = acc[t][r];
            if (g.bias) z += g.bias[n];
            if (g.Z) g.Z[(size_t)m * g.ldy + n] = z;
            float y = g.relu ? fmaxf(z, 0.0f) : z;
            if (g.R) y = g.R[(size_t)m * g.ldy + n] + y;
            g.Y[(size_t)m * g.ldy + n] = y;
        }
}

// ------------------------------------------------------------------------------------------ dW = X^T dY, db = sum dY
struct WgradArgs {
    const float* X; uint32_t ldx; const float* X2; uint32_t ldx2, k1; // M x K (the concatenation as above)
    const float* dY; uint32_t ldd;                                     // M x N
    float* part;                                                       // [chunks][K][N]
    float* colpart;                                                    // [chunks][N], or null
    uint32_t M, K, N;
};

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs g) {
    __shared__ uint16_t XT[128][kChunk + 4];
    __shared__ uint16_t DT[64][kChunk + 4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, c = blockIdx.x, m0 = c * kChunk;
    for (uint32_t e = tid; e < kChunk * g.K; e += 256) {
        const uint32_t r = e / g.K, k = e % g.K, m = m0 + r;
        float x = 0.0f;
        if (m < g.M) x = k < g.k1 ? g.X[(size_t)m * g.ldx + k] : g.X2[(size_t)m * g.ldx2 + (k - g.k1)];
        XT[k][r] = f2bf(x);
    }
    for (uint32_t e = tid; e < kChunk * g.N; e += 256) {
        const uint32_t r = e / g.N, n = e % g.N, m = m0 + r;
        DT[n][r] = f2bf(m < g.M ? g.dY[(size_t)m * g.ldd + n] : 0.0f);
    }
    if (g.colpart && tid < g.N) {
        float s = 0.0f;
        for (uint32_t r = 0; r < kChunk && m0 + r < g.M; r++) s += g.dY[(size_t)(m0 + r) * g.ldd + tid];
        g.colpart[(size_t)c * g.N + tid] = s;
    }
    __syncthreads();
    const uint32_t li = lane & 15, lk = 4 * (lane >> 4);
    const uint32_t kt = g.K / 16, ntl = g.N / 16, ntiles = kt * ntl;
    for (uint32_t t = wv; t < ntiles; t += 4) {
        const uint32_t tk = t / ntl, tn = t % ntl;
        float4_ acc = {0, 0, 0, 0};
        for (uint32_t r0 = 0; r0 < kChunk; r0 += 16) {
            short4_ a, b;
            for (int e = 0; e < 4; e++) a[e] = (short)XT[16 * tk + li][r0 + lk + e], b[e] = (short)DT[16 * tn + li][r0 + lk + e];
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        for (int r = 0; r < 4; r++) g.part[((size_t)c * g.K + 16 * tk + lk + r) * g.N + 16 * tn + li] = acc[r];
    }
}

// the chunks' partials summed in chunk order (into dst, and the column sums into dcol)
__global__ void k_wreduce(const float* part, const float* colpart, uint32_t chunks, uint32_t K, uint32_t N, float* dst,
                          float* dcol) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K * N) {
        float s = 0.0f;
        for (uint32_t c = 0; c < chunks; c++) s += part[(size_t)c * K * N + i];
        dst[i] = s;
    }
    if (dcol && i < N) {
        float s = 0.0f;
        for (uint32_t c = 0; c < chunks; c++) s += colpart[(size_t)c * N + i];
        dcol[i] = s;
    }
}

// ------------------------------------------------------------------------------------------ messages
struct Graph {
    const uint32_t* off; // per node: its neighbours nbr[off[i] .. off[i + 1]]
    const uint32_t* nbr;
    const uint32_t* rev; // per edge (i <- j): the edge (j <- i)
    const float4* rel;   // per edge (i <- j): (p_j - p_i) / the mean edge length
    uint32_t n;
};

// agg_i = 1/deg_i sum_j relu(P_j + rel_ij U + b_m); a thread per (node, feature)
__global__ void k_msg(Graph G, const float* P, const float* U, const float* bm, float* agg) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, i = t / H, c = t % H;
    if (i >= G.n) return;
    const float u0 = U[c], u1 = U[H + c], u2 = U[2 * H + c], u3 = U[3 * H + c], b = bm[c];
    float s = 0.0f;
    const uint32_t e0 = G.off[i], e1 = G.off[i + 1];
    for (uint32_t e = e0; e < e1; e++) {
        const float4 r = G.rel[e];
        const float q = P[(size_t)G.nbr[e] * H + c] + (((r.x * u0 + r.y * u1) + r.z * u2) + r.w * u3) + b;
        s += fmaxf(q, 0.0f);
    }
    agg[(size_t)i * H + c] = e1 > e0 ? s / (float)(e1 - e0) : 0.0f;
}

// backward: dP_j = sum over the edges (i <- j) of dagg_i / deg_i [q > 0]; per (node, feature) partials of dU (4)
// and db_m over the node's incoming edges (i <- j, summed as the edges j -> i leave j)
__global__ void k_msg_bwd(Graph G, const float* P, const float* U, const float* bm, const float* dagg, float* dP,
                          float* upart) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, j = t / H, c = t % H;
    if (j >= G.n) return;
    const float u0 = U[c], u1 = U[H + c], u2 = U[2 * H + c], u3 = U[3 * H + c], b = bm[c];
    const float pj = P[(size_t)j * H + c];
    float s = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f, gb = 0.0f;
    for (uint32_t e = G.off[j]; e < G.off[j + 1]; e++) {
        const uint32_t i = G.nbr[e], er = G.rev[e]; // the edge (i <- j)
        const float4 r = G.rel[er];
        const float q = pj + (((r.x * u0 + r.y * u1) + r.z * u2) + r.w * u3) + b;
        if (q > 0.0f) {
            const float di = dagg[(size_t)i * H + c] / (float)(G.off[i + 1] - G.off[i]);
            s += di, g0 += r.x * di, g1 += r.y * di, g2 += r.z * di, g3 += r.w * di, gb += di;
        }
    }
    dP[(size_t)j * H + c] = s;
    float* u = upart + (size_t)j * 5 * H;
    u[c] = g0, u[H + c] = g1, u[2 * H + c] = g2, u[3 * H + c] = g3, u[4 * H + c] = gb;
}

// the per-node partials of dU and db_m summed in node order
__global__ void k_ureduce(const float* upart, uint32_t n, float* dU, float* dbm) {
    const uint32_t i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= 5 * H) return;
    float s = 0.0f;
    for (uint32_t j = 0; j < n; j++) s += upart[(size_t)j * 5 * H + i];
    if (i < 4 * H) dU[i] = s;
    else dbm[i - 4 * H] = s;
}

// ------------------------------------------------------------------------------------------ elementwise pieces
// dst = src * (z > 0)  (n elements)
__global__ void k_relu_bwd(const float* src, const float* z, float* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = z[i] > 0.0f ? src[i] : 0.0f;
}
// dst = a + b[:, :64 of a row of 128]  (dh_prev = dh + dcat[:, :H]); dagg = dcat[:, H:]
__global__ void k_split_cat(const float* dh, const float* dcat, float* dh_out, float* dagg, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, i = t / H, c = t % H;
    if (i >= n) return;
    dh_out[(size_t)i * H + c] = dh[(size_t)i * H + c] + dcat[(size_t)i * 2 * H + c];
    dagg[(size_t)i * H + c] = dcat[(size_t)i * 2 * H + H + c];
}
__global__ void k_add(const float* a, const float* b, float* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a[i] + b[i];
}

// ------------------------------------------------------------------------------------------ features, loss
// per node: D_i^-1 (Gauss-Jordan with partial pivoting on the symmetrized block, FP64), t = D^-1 b, and the partial
// sums of t^2 and b^2 (translations, rotations) per workgroup
__global__ void k_feat1(blf::FemView v, double* Dinv, double* t, double* part) {
    __shared__ double s[6][256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double tt = 0, tr = 0, bt = 0, br = 0, xt = 0, xr = 0;
    if (i < v.nf) {
        double M[6][12];
        const double* D = v.D + 36 * (size_t)i;
        for (int p = 0; p < 6; p++)
            for (int q = 0; q < 6; q++) M[p][q] = p >= q ? D[p * 6 + q] : D[q * 6 + p], M[p][6 + q] = p == q ? 1.0 : 0.0;
        for (int c = 0; c < 6; c++) {
            int piv = c;
            for (int r = c + 1; r < 6; r++)
                if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
            if (piv != c)
                for (int q = 0; q < 12; q++) {
                    const double x = M[c][q];
                    M[c][q] = M[piv][q], M[piv][q] = x;
                }
            const double id = 1.0 / M[c][c];
            for (int q = 0; q < 12; q++) M[c][q] *= id;
            for (int r = 0; r < 6; r++)
                if (r != c) {
                    const double fct = M[r][c];
                    for (int q = 0; q < 12; q++) M[r][q] -= fct * M[c][q];
                }
        }
        double* Di = Dinv + 36 * (size_t)i;
        for (int p = 0; p < 6; p++)
            for (int q = 0; q < 6; q++) Di[p * 6 + q] = M[p][6 + q];
        const double* b = v.rhs + 6 * (size_t)i;
        for (int p = 0; p < 6; p++) {
            double x = 0;
            for (int q = 0; q < 6; q++) x += M[p][6 + q] * b[q];
            t[6 * (size_t)i + p] = x;
            const double xp = v.xprev[6 * (size_t)i + p];
            if (p < 3) tt += x * x, bt += b[p] * b[p], xt += xp * xp;
            else tr += x * x, br += b[p] * b[p], xr += xp * xp;
        }
    }
    s[0][threadIdx.x] = tt, s[1][threadIdx.x] = tr, s[2][threadIdx.x] = bt, s[3][threadIdx.x] = br;
    s[4][threadIdx.x] = xt, s[5][threadIdx.x] = xr;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off)
            for (int k = 0; k < 6; k++) s[k][threadIdx.x] += s[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[6 * blockIdx.x + k] = s[k][0];
}

// sigma (6), beta (6), tau (6) from the partials (one thread, in block order)
__global__ void k_scales(const double* part, uint32_t nblk, uint32_t n, double* scale) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = 0; b < nblk; b++)
        for (int k = 0; k < 6; k++) a[k] += part[6 * b + k];
    const double st = sqrt(a[0] / (3.0 * n)), sr = sqrt(a[1] / (3.0 * n));
    const double bt = sqrt(a[2] / (3.0 * n)) + 1e-30, br = sqrt(a[3] / (3.0 * n)) + 1e-30;
    double xt = sqrt(a[4] / (3.0 * n)), xr = sqrt(a[5] / (3.0 * n));
    xt = xt > 0 ? xt : st, xr = xr > 0 ? xr : sr;
    for (int k = 0; k < 3; k++) {
        scale[k] = xt + 1e-30, scale[3 + k] = xr + 1e-30;
        scale[6 + k] = bt, scale[9 + k] = br;
        scale[12 + k] = st + 1e-30, scale[15 + k] = sr + 1e-30;
    }
}

__global__ void k_feat2(blf::FemView v, const double* t, const double* scale, const double* xprev, float* f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.nf) return;
    float* fi = f + (size_t)i * FIN;
    for (int k = 0; k < 6; k++) {
        fi[k] = (float)(t[6 * (size_t)i + k] / scale[12 + k]);
        fi[6 + k] = (float)(xprev[6 * (size_t)i + k] / scale[k]);
        fi[12 + k] = (float)(v.rhs[6 * (size_t)i + k] / scale[6 + k]);
    }
    for (int k = 18; k < FIN; k++) fi[k] = 0.0f;
}

// the warm start x0 = sigma o[0:6] and the preconditioner's scales s = exp(clip(o[6:12], -3, 3))
__global__ void k_outputs(const float* o, const double* scale, uint32_t n, double* x0, double* s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 6; k++) {
        x0[6 * (size_t)i + k] = (double)o[(size_t)i * FOUT + k] * scale[k];
        const float l = fminf(fmaxf(o[(size_t)i * FOUT + 6 + k], -3.0f), 3.0f);
        s[6 * (size_t)i + k] = (double)expf(l);
    }
}

// the loss in two passes. Per node: yhat = o[0:6], x0 = sigma yhat, r0 = b - A x0, e = (x* - x0) / sigma,
// s = exp(clip(o[6:12])), u = D^-1 (s r0), z = s u / sigma.
struct LossNode {
    double yh[6], ys[6], r0[6], e[6], s[6], lo[6], u[6], z[6];
};
__device__ void loss_node(const blf::FemView& v, uint32_t i, const float* o, const double* scale, const double* Dinv,
                          const double* Ax0, const double* xstar, LossNode& q) {
    const float* oi = o + (size_t)i * FOUT;
    const double* Di = Dinv + 36 * (size_t)i;
    for (int k = 0; k < 6; k++) {
        q.yh[k] = oi[k];
        const double x0 = q.yh[k] * scale[k];
        q.ys[k] = xstar[6 * (size_t)i + k] / scale[k];
        q.r0[k] = v.rhs[6 * (size_t)i + k] - Ax0[6 * (size_t)i + k];
        q.e[k] = (xstar[6 * (size_t)i + k] - x0) / scale[k];
        q.lo[k] = oi[6 + k];
        q.s[k] = exp(fmin(fmax(q.lo[k], -3.0), 3.0));
    }
    for (int p = 0; p < 6; p++) {
        double a = 0;
        for (int r = 0; r < 6; r++) a += Di[p * 6 + r] * (q.s[r] * q.r0[r]);
        q.u[p] = a;
        q.z[p] = q.s[p] * a / scale[p];
    }
}

// pass 1: dL1/do[0:6]; per workgroup: L1, <z, e>, <z, z> -> part[3 b ..]
__global__ void k_loss1(blf::FemView v, const float* o, const double* scale, const double* Dinv, const double* Ax0,
                        const double* xstar, float* dO, double* part) {
    __shared__ double s3[3][256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double inv6n = 1.0 / (6.0 * v.nf);
    double l1 = 0, ze = 0, zz = 0;
    if (i < v.nf) {
        LossNode q;
        loss_node(v, i, o, scale, Dinv, Ax0, xstar, q);
        for (int k = 0; k < 6; k++) {
            l1 += (q.yh[k] - q.ys[k]) * (q.yh[k] - q.ys[k]);
            dO[(size_t)i * FOUT + k] = (float)(2.0 * (q.yh[k] - q.ys[k]) * inv6n);
            ze += q.z[k] * q.e[k], zz += q.z[k] * q.z[k];
        }
        for (int k = 12; k < FOUT; k++) dO[(size_t)i * FOUT + k] = 0.0f;
    }
    s3[0][threadIdx.x] = l1 * inv6n, s3[1][threadIdx.x] = ze, s3[2][threadIdx.x] = zz;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off)
            for (int k = 0; k < 3; k++) s3[k][threadIdx.x] += s3[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[3 * blockIdx.x + k] = s3[k][0];
}

// L1 and kappa = <z, e> / <z, z> from pass 1's partials (block order) -> lsum[0], lsum[2]
__global__ void k_kappa(const double* part, uint32_t nblk, double* lsum) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0, b = 0, c = 0;
    for (uint32_t k = 0; k < nblk; k++) a += part[3 * k], b += part[3 * k + 1], c += part[3 * k + 2];
    lsum[0] = a, lsum[2] = c > 0 ? b / c : 0.0;
}

// pass 2: dL2/do[6:12] and the per-workgroup L2 -> part2[b]
__global__ void k_loss2(blf::FemView v, const float* o, const double* scale, const double* Dinv, const double* Ax0,
                        const double* xstar, const double* lsum, float* dO, double* part2) {
    __shared__ double s1[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double inv6n = 1.0 / (6.0 * v.nf), kap = lsum[2];
    double l2 = 0;
    if (i < v.nf) {
        LossNode q;
        loss_node(v, i, o, scale, Dinv, Ax0, xstar, q);
        double dz[6];
        for (int k = 0; k < 6; k++) {
            const double d = kap * q.z[k] - q.e[k];
            l2 += d * d;
            dz[k] = 2.0 * kap * d * inv6n;
        }
        const double* Di = Dinv + 36 * (size_t)i;
        for (int m = 0; m < 6; m++) {
            double a = 0;
            for (int k = 0; k < 6; k++) a += dz[k] * q.s[k] / scale[k] * Di[k * 6 + m];
            const double ds = dz[m] * q.u[m] / scale[m] + q.r0[m] * a;
            dO[(size_t)i * FOUT + 6 + m] = (q.lo[m] > -3.0 && q.lo[m] < 3.0) ? (float)(ds * q.s[m]) : 0.0f;
        }
    }
    s1[threadIdx.x] = l2 * inv6n;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) s1[threadIdx.x] += s1[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part2[blockIdx.x] = s1[0];
}

__global__ void k_l2sum(const double* part2, uint32_t nblk, double* lsum) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0;
    for (uint32_t k = 0; k < nblk; k++) a += part2[k];
    lsum[1] = a;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master weights
__global__ void k_adam(float* p, const float* g, float* m, float* v2, uint32_t n, float lr, float bc1, float bc2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v2[i] + 0.001f * gi * gi;
    m[i] = mi, v2[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

// ------------------------------------------------------------------------------------------ PCG (FP64)
// z = S D^-1 S r (S = 1 without the network's scales); partial sums of r.z and r.r per workgroup
__global__ void k_pcg_prec(uint32_t n, const double* Dinv, const double* s, const double* r, double* z, double* part,
                           const int* done) {
    __shared__ double a1[256], a2[256];
    if (*done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, rr = 0;
    if (i < n) {
        double sr[6];
        for (int k = 0; k < 6; k++) sr[k] = (s ? s[6 * (size_t)i + k] : 1.0) * r[6 * (size_t)i + k];
        const double* Di = Dinv + 36 * (size_t)i;
        for (int p = 0; p < 6; p++) {
            double a = 0;
            for (int q = 0; q < 6; q++) a += Di[p * 6 + q] * sr[q];
            const double zp = (s ? s[6 * (size_t)i + p] : 1.0) * a;
            z[6 * (size_t)i + p] = zp;
            rz += r[6 * (size_t)i + p] * zp;
            rr += r[6 * (size_t)i + p] * r[6 * (size_t)i + p];
        }
    }
    a1[threadIdx.x] = rz, a2[threadIdx.x] = rr;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) a1[threadIdx.x] += a1[threadIdx.x + off], a2[threadIdx.x] += a2[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[2 * blockIdx.x] = a1[0], part[2 * blockIdx.x + 1] = a2[0];
}

// partial sums of p.Ap
__global__ void k_pcg_dot(uint32_t n6, const double* a, const double* b, double* part, const int* done) {
    __shared__ double s[256];
    if (*done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    s[threadIdx.x] = i < n6 ? a[i] * b[i] : 0.0;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// the scalars of an iteration. mode 0: the start (rz0, ||r0||, convergence); mode 1: alpha = rz / p.Ap; mode 2: beta and
// the convergence test after the update. sc: [rz, alpha, beta, ||b||^2, tol^2, iterations, ||r||^2]
// The sums of the partials (one or two interleaved series) by one workgroup of 256: each thread a strided subset
// in order, then a fixed tree - the same order every time. Thread 0 gets the totals.
__device__ __forceinline__ void block_sum2(const double* part, uint32_t n, int stride, double& a, double& b) {
    __shared__ double s1[256], s2[256];
    double x = 0, y = 0;
    for (uint32_t k = threadIdx.x; k < n; k += 256) x += part[stride * k], y += stride > 1 ? part[stride * k + 1] : 0.0;
    s1[threadIdx.x] = x, s2[threadIdx.x] = y;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) s1[threadIdx.x] += s1[threadIdx.x + off], s2[threadIdx.x] += s2[threadIdx.x + off];
        __syncthreads();
    }
    a = s1[0], b = s2[0];
}

__global__ __launch_bounds__(256) void k_pcg_scalar(const double* part, uint32_t nblk, int stride, int mode, double* sc, int* done, int* iters) {
    if (*done) return;
    double a, b;
    block_sum2(part, nblk, stride, a, b);
    if (threadIdx.x != 0) return;
    if (mode == 1) {
        sc[1] = a != 0 ? sc[0] / a : 0.0;
        return;
    }
    sc[6] = b;
    if (mode == 0) {
        sc[0] = a;
    } else {
        sc[2] = sc[0] != 0 ? a / sc[0] : 0.0;
        sc[0] = a;
        *iters += 1;
    }
    if (b <= sc[4] * sc[3]) *done = 1;
}

// The scalars of PCG with the factor as its preconditioner (the residual tested before the preconditioner is
// applied, so a converged iteration skips its solve). mode 0: rz0; 1: alpha = rz / p.Ap (one more iteration);
// 2: ||r||^2 and the test; 3: beta = rz' / rz. sc as k_pcg_scalar's.
__global__ __launch_bounds__(256) void k_fac_scalar(const double* part, uint32_t nblk, int mode, double* sc, int* done, int* iters) {
    if (*done) return;
    double a, b;
    block_sum2(part, nblk, 1, a, b);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        sc[0] = a;
    } else if (mode == 1) {
        sc[1] = a != 0 ? sc[0] / a : 0.0;
        *iters += 1;
    } else if (mode == 2) {
        sc[6] = a;
        if (a <= sc[4] * sc[3]) *done = 1;
    } else {
        sc[2] = sc[0] != 0 ? a / sc[0] : 0.0;
        sc[0] = a;
    }
}

// x += alpha p, r -= alpha Ap
__global__ void k_pcg_xr(uint32_t n6, const double* sc, double* x, double* r, const double* p, const double* Ap, const int* done) {
    if (*done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n6) return;
    x[i] += sc[1] * p[i];
    r[i] -= sc[1] * Ap[i];
}
// p = z + beta p (mode 1) or p = z (mode 0)
__global__ void k_pcg_p(uint32_t n6, const double* sc, const double* z, double* p, int first, const int* done) {
    if (*done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n6) return;
    p[i] = first ? z[i] : z[i] + sc[2] * p[i];
}
__global__ void k_sub(uint32_t n6, const double* a, const double* b, double* dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n6) dst[i] = a[i] - b[i];
}
__global__ void k_clear_i(int* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

// ||b||^2 (the partials summed in block_sum2's fixed order) into sc[3], tol^2 into sc[4]
__global__ __launch_bounds__(256) void k_norm_init(const double* part, uint32_t nblk, double* sc, double tol2) {
    double a, b;
    block_sum2(part, nblk, 1, a, b);
    if (threadIdx.x == 0) sc[3] = a, sc[4] = tol2;
}

// y = A x over the engine's assembled system (as its own product), skipped once the solve has converged
__global__ void k_spmv_g(blf::FemView v, const double* x, double* y, const int* done) {
    if (*done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.nf) return;
    double acc[6];
    const double* D = v.D + 36 * (size_t)i;
    for (int p = 0; p < 6; p++) {
        double a = 0;
        for (int q = 0; q < 6; q++) a += (p >= q ? D[p * 6 + q] : D[q * 6 + p]) * x[6 * i + q];
        acc[p] = a;
    }
    for (uint32_t k = v.adj_off[i]; k < v.adj_off[i + 1]; k++) {
        const uint2 e = v.adj[k];
        const double* O = v.O + 36 * (size_t)e.x;
        const uint32_t o = e.y & 0x7fffffffu;
        const bool row_side = (e.y >> 31) != 0;
        for (int p = 0; p < 6; p++) {
            double a = 0;
            for (int q = 0; q < 6; q++) a += (row_side ? O[p * 6 + q] : O[q * 6 + p]) * x[6 * o + q];
            acc[p] += a;
        }
    }
    for (int p = 0; p < 6; p++) y[6 * i + p] = acc[p];
}

__global__ void k_copy_d(const double* src, double* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
__global__ void k_zero_d(double* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = 0.0;
}

} // namespace blnn

// ============================================================================================== host
namespace {

using namespace blnn;

#define GCHK(x)                                                                                                     \
    do {                                                                                                            \
        hipError_t e_ = (x);                                                                                        \
        if (e_ != hipSuccess) {                                                                                     \
            blg::set_last_error((std::string(#x) + ": " + hipGetErrorString(e_)).c_str());                          \
            return BL_ERR_HIP;                                                                                      \
        }                                                                                                           \
    } while (0)

template <class T>
struct DB {
    T* p = nullptr;
    size_t n = 0;
    int ensure(size_t cnt) {
        if (cnt <= n && p) return BL_OK;
        if (p) (void)hipFree(p), p = nullptr;
        n = cnt;
        return hipMalloc((void**)&p, sizeof(T) * (cnt ? cnt : 1)) == hipSuccess ? BL_OK : BL_ERR_HIP;
    }
    ~DB() {
        if (p) (void)hipFree(p);
    }
};

inline dim3 grid(size_t n, uint32_t b = 256) { return dim3((uint32_t)((n + b - 1) / b)); }

} // namespace

struct bl_gnn {
    Layout L;
    DB<float> params, grads, m, v2;
    uint32_t adam_t = 0;
    // the mesh's graph and the workspace (for the engine they were built for)
    uint64_t mesh = 0; // the engine's serial
    uint32_t n = 0, ne = 0, nblk = 0, nchunks = 0;
    DB<uint32_t> off, nbr, rev;
    DB<float4> rel;
    DB<double> Dinv, t, part, part2, scale, x0, s, Ax0, sums, pcg_r, pcg_z, pcg_p, pcg_Ap, pcg_sc, pcg_part;
    DB<int> pcg_flag;
    DB<float> f, z0, h[NL + 1], P[NL], agg[NL], zu[NL], o, dO, dh, dh2, dcat, dzu, dP, dagg, upart, wpart, cpart;
    hipEvent_t ev[4] = {};

    ~bl_gnn() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    int setup(bl_fem* fm) {
        const blf::FemView v = blf::fem_view(fm);
        if (mesh == v.serial) return BL_OK;
        mesh = 0;
        n = v.nf;
        // the graph from the engine's pair lists (both directions), in the node's pair order
        std::vector<uint32_t> aoff(n + 1);
        GCHK(hipStreamSynchronize(v.st));
        GCHK(hipMemcpy(aoff.data(), v.adj_off, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost));
        std::vector<uint2> adj(aoff[n]);
        if (!adj.empty()) GCHK(hipMemcpy(adj.data(), v.adj, sizeof(uint2) * adj.size(), hipMemcpyDeviceToHost));
        std::vector<float> p0(3 * (size_t)n);
        GCHK(hipMemcpy(p0.data(), v.p0, sizeof(float) * p0.size(), hipMemcpyDeviceToHost));
        std::vector<uint32_t> hoff(n + 1, 0), hnbr;
        std::vector<std::vector<uint32_t>> lists(n);
        for (uint32_t i = 0; i < n; i++) {
            for (uint32_t k = aoff[i]; k < aoff[i + 1]; k++) lists[i].push_back(adj[k].y & 0x7fffffffu);
            std::sort(lists[i].begin(), lists[i].end());
            hnbr.insert(hnbr.end(), lists[i].begin(), lists[i].end());
            hoff[i + 1] = (uint32_t)hnbr.size();
        }
        ne = (uint32_t)hnbr.size();
        std::vector<uint32_t> hrev(ne);
        double ell = 0;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t e = hoff[i]; e < hoff[i + 1]; e++) {
                const uint32_t j = hnbr[e];
                const auto it = std::lower_bound(hnbr.begin() + hoff[j], hnbr.begin() + hoff[j + 1], i);
                hrev[e] = (uint32_t)(it - hnbr.begin());
                double d2 = 0;
                for (int k = 0; k < 3; k++) d2 += ((double)p0[3 * j + k] - p0[3 * i + k]) * ((double)p0[3 * j + k] - p0[3 * i + k]);
                ell += sqrt(d2);
            }
        ell = ne ? ell / ne : 1.0;
        std::vector<float4> hrel(ne);
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t e = hoff[i]; e < hoff[i + 1]; e++) {
                const uint32_t j = hnbr[e];
                hrel[e] = make_float4((float)(((double)p0[3 * j] - p0[3 * i]) / ell), (float)(((double)p0[3 * j + 1] - p0[3 * i + 1]) / ell),
                                      (float)(((double)p0[3 * j + 2] - p0[3 * i + 2]) / ell), 0.0f);
            }
        nblk = (n + 255) / 256;
        nchunks = (n + kChunk - 1) / kChunk;
        int rc;
        if ((rc = off.ensure(n + 1)) || (rc = nbr.ensure(ne)) || (rc = rev.ensure(ne)) || (rc = rel.ensure(ne)) ||
            (rc = Dinv.ensure(36 * (size_t)n)) || (rc = t.ensure(6 * (size_t)n)) || (rc = part.ensure(6 * (size_t)nblk + 6)) ||
            (rc = part2.ensure((size_t)nblk + 1)) || (rc = scale.ensure(18)) || (rc = x0.ensure(6 * (size_t)n)) || (rc = s.ensure(6 * (size_t)n)) ||
            (rc = Ax0.ensure(6 * (size_t)n)) || (rc = sums.ensure(4)) || (rc = pcg_r.ensure(6 * (size_t)n)) ||
            (rc = pcg_z.ensure(6 * (size_t)n)) || (rc = pcg_p.ensure(6 * (size_t)n)) || (rc = pcg_Ap.ensure(6 * (size_t)n)) ||
            (rc = pcg_sc.ensure(8)) || (rc = pcg_part.ensure(2 * (size_t)(6 * n / 256 + 2) + 2 * nblk)) || (rc = pcg_flag.ensure(2)) ||
            (rc = f.ensure((size_t)n * FIN)) || (rc = z0.ensure((size_t)n * H)) || (rc = o.ensure((size_t)n * FOUT)) ||
            (rc = dO.ensure((size_t)n * FOUT)) || (rc = dh.ensure((size_t)n * H)) || (rc = dh2.ensure((size_t)n * H)) ||
            (rc = dcat.ensure((size_t)n * 2 * H)) || (rc = dzu.ensure((size_t)n * H)) || (rc = dP.ensure((size_t)n * H)) ||
            (rc = dagg.ensure((size_t)n * H)) || (rc = upart.ensure((size_t)n * 5 * H)) ||
            (rc = wpart.ensure((size_t)nchunks * 2 * H * H)) || (rc = cpart.ensure((size_t)nchunks * H)))
            return rc;
        for (int l = 0; l <= NL; l++)
            if ((rc = h[l].ensure((size_t)n * H))) return rc;
        for (int l = 0; l < NL; l++)
            if ((rc = P[l].ensure((size_t)n * H)) || (rc = agg[l].ensure((size_t)n * H)) || (rc = zu[l].ensure((size_t)n * H))) return rc;
        GCHK(hipMemcpy(off.p, hoff.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
        if (ne) {
            GCHK(hipMemcpy(nbr.p, hnbr.data(), sizeof(uint32_t) * ne, hipMemcpyHostToDevice));
            GCHK(hipMemcpy(rev.p, hrev.data(), sizeof(uint32_t) * ne, hipMemcpyHostToDevice));
            GCHK(hipMemcpy(rel.p, hrel.data(), sizeof(float4) * ne, hipMemcpyHostToDevice));
        }
        mesh = v.serial;
        return BL_OK;
    }

    Graph graph() const { return Graph{off.p, nbr.p, rev.p, rel.p, n}; }

    void gemm(hipStream_t st, const float* X, uint32_t ldx, const float* X2, uint32_t ldx2, uint32_t k1, const float* W,
              uint32_t ldw, int wt, const float* bias, const float* R, float* Y, uint32_t ldy, float* Z, uint32_t K,
              uint32_t N, int relu) {
        GemmArgs g{X, ldx, X2, ldx2, k1, W, ldw, wt, bias, R, Y, ldy, Z, n, K, N, relu};
        hipLaunchKernelGGL(k_gemm, grid(n, 64), dim3(256), 0, st, g);
    }
    void wgrad(hipStream_t st, const float* X, uint32_t ldx, const float* X2, uint32_t ldx2, uint32_t k1, const float* dY,
               uint32_t ldd, uint32_t K, uint32_t N, float* dW, float* db) {
        WgradArgs g{X, ldx, X2, ldx2, k1, dY, ldd, wpart.p, db ? cpart.p : nullptr, n, K, N};
        hipLaunchKernelGGL(k_wgrad, dim3(nchunks), dim3(256), 0, st, g);
        hipLaunchKernelGGL(k_wreduce, grid((size_t)K * N), dim3(256), 0, st, wpart.p, cpart.p, nchunks, K, N, dW, db);
    }

    // features of the engine's current system (its right side, blocks and previous solution)
    void features(hipStream_t st, const blf::FemView& v) {
        hipLaunchKernelGGL(k_feat1, dim3(nblk), dim3(256), 0, st, v, Dinv.p, t.p, part.p);
        hipLaunchKernelGGL(k_scales, dim3(1), dim3(64), 0, st, part.p, nblk, n, scale.p);
        hipLaunchKernelGGL(k_feat2, dim3(nblk), dim3(256), 0, st, v, t.p, scale.p, v.xprev, f.p);
    }

    void forward(hipStream_t st) {
        const float* W = params.p;
        gemm(st, f.p, FIN, nullptr, 0, FIN, W + L.We, H, 0, W + L.be, nullptr, h[0].p, H, z0.p, FIN, H, 1);
        for (int l = 0; l < NL; l++) {
            gemm(st, h[l].p, H, nullptr, 0, H, W + L.Wm[l], H, 0, nullptr, nullptr, P[l].p, H, nullptr, H, H, 0);
            hipLaunchKernelGGL(k_msg, grid((size_t)n * H), dim3(256), 0, st, graph(), P[l].p, W + L.U[l], W + L.bm[l], agg[l].p);
            gemm(st, h[l].p, H, agg[l].p, H, H, W + L.Wu[l], H, 0, W + L.bu[l], h[l].p, h[l + 1].p, H, zu[l].p, 2 * H, H, 1);
        }
        gemm(st, h[NL].p, H, nullptr, 0, H, W + L.Wd, FOUT, 0, W + L.bd, nullptr, o.p, FOUT, nullptr, H, FOUT, 0);
    }

    void backward(hipStream_t st) {
        const float* W = params.p;
        float* G = grads.p;
        const size_t nh = (size_t)n * H;
        wgrad(st, h[NL].p, H, nullptr, 0, H, dO.p, FOUT, H, FOUT, G + L.Wd, G + L.bd);
        gemm(st, dO.p, FOUT, nullptr, 0, FOUT, W + L.Wd, FOUT, 1, nullptr, nullptr, dh.p, H, nullptr, FOUT, H, 0);
        for (int l = NL - 1; l >= 0; l--) {
            hipLaunchKernelGGL(k_relu_bwd, grid(nh), dim3(256), 0, st, dh.p, zu[l].p, dzu.p, (uint32_t)nh);
            wgrad(st, h[l].p, H, agg[l].p, H, H, dzu.p, H, 2 * H, H, G + L.Wu[l], G + L.bu[l]);
            gemm(st, dzu.p, H, nullptr, 0, H, W + L.Wu[l], H, 1, nullptr, nullptr, dcat.p, 2 * H, nullptr, H, 2 * H, 0);
            hipLaunchKernelGGL(k_split_cat, grid(nh), dim3(256), 0, st, dh.p, dcat.p, dh2.p, dagg.p, n);
            hipLaunchKernelGGL(k_msg_bwd, grid(nh), dim3(256), 0, st, graph(), P[l].p, W + L.U[l], W + L.bm[l], dagg.p, dP.p, upart.p);
            hipLaunchKernelGGL(k_ureduce, dim3(2), dim3(256), 0, st, upart.p, n, G + L.U[l], G + L.bm[l]);
            wgrad(st, h[l].p, H, nullptr, 0, H, dP.p, H, H, H, G + L.Wm[l], nullptr);
            // dh = dh_prev + dP W_m^T
            gemm(st, dP.p, H, nullptr, 0, H, W + L.Wm[l], H, 1, nullptr, dh2.p, dh.p, H, nullptr, H, H, 0);
        }
        hipLaunchKernelGGL(k_relu_bwd, grid(nh), dim3(256), 0, st, dh.p, z0.p, dzu.p, (uint32_t)nh);
        wgrad(st, f.p, FIN, nullptr, 0, FIN, dzu.p, H, FIN, H, G + L.We, G + L.be);
    }

    // the loss of the current system against the engine's solution x* (its x), gradients into grads
    int loss_grads(bl_fem* fm, double losses[2]) {
        const blf::FemView v = blf::fem_view(fm);
        const hipStream_t st = v.st;
        features(st, v);
        forward(st);
        hipLaunchKernelGGL(k_outputs, dim3(nblk), dim3(256), 0, st, o.p, scale.p, n, x0.p, s.p);
        blf::fem_spmv(fm, x0.p, Ax0.p);
        hipLaunchKernelGGL(k_loss1, dim3(nblk), dim3(256), 0, st, v, o.p, scale.p, Dinv.p, Ax0.p, v.x, dO.p, part.p);
        hipLaunchKernelGGL(k_kappa, dim3(1), dim3(64), 0, st, part.p, nblk, sums.p);
        hipLaunchKernelGGL(k_loss2, dim3(nblk), dim3(256), 0, st, v, o.p, scale.p, Dinv.p, Ax0.p, v.x, sums.p, dO.p, part2.p);
        hipLaunchKernelGGL(k_l2sum, dim3(1), dim3(64), 0, st, part2.p, nblk, sums.p);
        backward(st);
        GCHK(hipGetLastError());
        GCHK(hipMemcpyAsync(losses, sums.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        GCHK(hipStreamSynchronize(st));
        return BL_OK;
    }

    // x0 and s for the current system (inference)
    void infer(bl_fem* fm) {
        const blf::FemView v = blf::fem_view(fm);
        features(v.st, v);
        forward(v.st);
        hipLaunchKernelGGL(k_outputs, dim3(nblk), dim3(256), 0, v.st, o.p, scale.p, n, x0.p, s.p);
    }

    // PCG on the engine's current system from x (in place), preconditioner S D^-1 S (S = 1 when sv is null); the
    // block inverses must be in Dinv. K iterations are enqueued at once: each kernel reads the device's converged flag
    // and does nothing once it is up, so the host waits once.
    int pcg(bl_fem* fm, double* x, const double* sv, int max_iter, double tol, int* iters, int* converged) {
        const blf::FemView v = blf::fem_view(fm);
        const hipStream_t st = v.st;
        const uint32_t n6 = 6 * n, nb6 = (n6 + 255) / 256;
        double* sc = pcg_sc.p;
        int* done = pcg_flag.p;
        int* it = pcg_flag.p + 1;
        hipLaunchKernelGGL(k_clear_i, dim3(1), dim3(64), 0, st, pcg_flag.p, 2u);
        hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, v.rhs, v.rhs, pcg_part.p, done);
        hipLaunchKernelGGL(k_norm_init, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, sc, tol * tol);
        hipLaunchKernelGGL(k_spmv_g, dim3(nblk), dim3(256), 0, st, v, (const double*)x, pcg_Ap.p, done);
        hipLaunchKernelGGL(k_sub, grid(n6), dim3(256), 0, st, n6, v.rhs, pcg_Ap.p, pcg_r.p);
        hipLaunchKernelGGL(k_pcg_prec, dim3(nblk), dim3(256), 0, st, n, Dinv.p, sv, pcg_r.p, pcg_z.p, pcg_part.p, done);
        hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nblk, 2, 0, sc, done, it);
        hipLaunchKernelGGL(k_pcg_p, grid(n6), dim3(256), 0, st, n6, sc, pcg_z.p, pcg_p.p, 1, done);
        for (int k = 0; k < max_iter; k++) {
            hipLaunchKernelGGL(k_spmv_g, dim3(nblk), dim3(256), 0, st, v, (const double*)pcg_p.p, pcg_Ap.p, done);
            hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, pcg_p.p, pcg_Ap.p, pcg_part.p, done);
            hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, 1, 1, sc, done, it);
            hipLaunchKernelGGL(k_pcg_xr, grid(n6), dim3(256), 0, st, n6, sc, x, pcg_r.p, pcg_p.p, pcg_Ap.p, done);
            hipLaunchKernelGGL(k_pcg_prec, dim3(nblk), dim3(256), 0, st, n, Dinv.p, sv, pcg_r.p, pcg_z.p, pcg_part.p, done);
            hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nblk, 2, 2, sc, done, it);
            hipLaunchKernelGGL(k_pcg_p, grid(n6), dim3(256), 0, st, n6, sc, pcg_z.p, pcg_p.p, 0, done);
        }
        GCHK(hipGetLastError());
        int hv[2];
        GCHK(hipMemcpyAsync(hv, pcg_flag.p, sizeof(hv), hipMemcpyDeviceToHost, st));
        GCHK(hipStreamSynchronize(st));
        *converged = hv[0], *iters = hv[1];
        return BL_OK;
    }

    // PCG preconditioned by the engine's last factorization: z = (L L^T)^-1 r by the supernodal substitutions. All
    // max_iter iterations are enqueued at once; after the residual test passes every launch returns at once.
    int pcg_factor(bl_fem* fm, double* x, int max_iter, double tol, int* iters, int* converged) {
        const blf::FemView v = blf::fem_view(fm);
        const hipStream_t st = v.st;
        const uint32_t n6 = 6 * n, nb6 = (n6 + 255) / 256;
        double* sc = pcg_sc.p;
        int* done = pcg_flag.p;
        int* it = pcg_flag.p + 1;
        int rc;
        hipLaunchKernelGGL(k_clear_i, dim3(1), dim3(64), 0, st, pcg_flag.p, 2u);
        hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, v.rhs, v.rhs, pcg_part.p, done);
        hipLaunchKernelGGL(k_norm_init, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, sc, tol * tol);
        hipLaunchKernelGGL(k_spmv_g, dim3(nblk), dim3(256), 0, st, v, (const double*)x, pcg_Ap.p, done);
        hipLaunchKernelGGL(k_sub, grid(n6), dim3(256), 0, st, n6, v.rhs, pcg_Ap.p, pcg_r.p);
        hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, pcg_r.p, pcg_r.p, pcg_part.p, done);
        hipLaunchKernelGGL(k_fac_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, 2, sc, done, it);
        if ((rc = blf::fem_precond(fm, pcg_r.p, pcg_z.p, done))) return rc;
        hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, pcg_r.p, pcg_z.p, pcg_part.p, done);
        hipLaunchKernelGGL(k_fac_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, 0, sc, done, it);
        hipLaunchKernelGGL(k_pcg_p, grid(n6), dim3(256), 0, st, n6, sc, pcg_z.p, pcg_p.p, 1, done);
        for (int k = 0; k < max_iter; k++) {
            hipLaunchKernelGGL(k_spmv_g, dim3(nblk), dim3(256), 0, st, v, (const double*)pcg_p.p, pcg_Ap.p, done);
            hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, pcg_p.p, pcg_Ap.p, pcg_part.p, done);
            hipLaunchKernelGGL(k_fac_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, 1, sc, done, it);
            hipLaunchKernelGGL(k_pcg_xr, grid(n6), dim3(256), 0, st, n6, sc, x, pcg_r.p, pcg_p.p, pcg_Ap.p, done);
            hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, pcg_r.p, pcg_r.p, pcg_part.p, done);
            hipLaunchKernelGGL(k_fac_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, 2, sc, done, it);
            if (k + 1 == max_iter) break; // (the last iteration's search direction is never used)
            if ((rc = blf::fem_precond(fm, pcg_r.p, pcg_z.p, done))) return rc;
            hipLaunchKernelGGL(k_pcg_dot, dim3(nb6), dim3(256), 0, st, n6, pcg_r.p, pcg_z.p, pcg_part.p, done);
            hipLaunchKernelGGL(k_fac_scalar, dim3(1), dim3(256), 0, st, pcg_part.p, nb6, 3, sc, done, it);
            hipLaunchKernelGGL(k_pcg_p, grid(n6), dim3(256), 0, st, n6, sc, pcg_z.p, pcg_p.p, 0, done);
        }
        GCHK(hipGetLastError());
        int hv[2];
        GCHK(hipMemcpyAsync(hv, pcg_flag.p, sizeof(hv), hipMemcpyDeviceToHost, st));
        GCHK(hipStreamSynchronize(st));
        int err = 0;
        if ((rc = blf::fem_solve_error(fm, &err))) return rc;
        if (err) {
            blg::set_last_error("a substitution's wait for a block timed out");
            return BL_ERR_STATE;
        }
        *converged = hv[0], *iters = hv[1];
        return BL_OK;
    }
};

// ============================================================================================== C ABI
namespace {
double rel_err(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0, nb = 0;
    for (size_t k = 0; k < a.size(); k++) d += (a[k] - b[k]) * (a[k] - b[k]), nb += b[k] * b[k];
    return nb > 0 ? std::sqrt(d / nb) : std::sqrt(d);
}
} // namespace

extern "C" {

int bl_gnn_param_count(void) { return (int)Layout().total; }

int bl_gnn_create(bl_gnn** out, const float* params, uint32_t n) {
    if (!out || !params) return BL_ERR_ARG;
    *out = nullptr;
    bl_gnn* g = new bl_gnn();
    if (n != g->L.total) {
        delete g;
        blg::set_last_error("parameter count differs from the network's");
        return BL_ERR_ARG;
    }
    int rc;
    if ((rc = g->params.ensure(n)) || (rc = g->grads.ensure(n)) || (rc = g->m.ensure(n)) || (rc = g->v2.ensure(n))) {
        delete g;
        return rc;
    }
    for (hipEvent_t& e : g->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete g;
            return BL_ERR_HIP;
        }
    if ((rc = bl_gnn_set_params(g, params, n))) {
        delete g;
        return rc;
    }
    *out = g;
    return BL_OK;
}

void bl_gnn_destroy(bl_gnn* g) {
    if (g) (void)hipDeviceSynchronize();
    delete g;
}

int bl_gnn_get_params(bl_gnn* g, float* out, uint32_t n) {
    if (!g || !out || n != g->L.total) return BL_ERR_ARG;
    GCHK(hipDeviceSynchronize());
    GCHK(hipMemcpy(out, g->params.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_gnn_set_params(bl_gnn* g, const float* p, uint32_t n) {
    if (!g || !p || n != g->L.total) return BL_ERR_ARG;
    GCHK(hipDeviceSynchronize());
    GCHK(hipMemcpy(g->params.p, p, sizeof(float) * n, hipMemcpyHostToDevice));
    std::vector<float> z(n, 0.0f);
    GCHK(hipMemcpy(g->m.p, z.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    GCHK(hipMemcpy(g->v2.p, z.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    g->adam_t = 0;
    return BL_OK;
}

int bl_gnn_gradients(bl_gnn* g, bl_fem* f, float* grads, uint32_t n, double losses[2]) {
    if (!g || !f || (grads && n != g->L.total)) return BL_ERR_ARG;
    int rc;
    double l[2];
    if ((rc = g->setup(f)) || (rc = g->loss_grads(f, l))) return rc;
    if (losses) losses[0] = l[0], losses[1] = l[1];
    if (grads) GCHK(hipMemcpy(grads, g->grads.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_gnn_train_step(bl_gnn* g, bl_fem* f, float lr, double losses[2]) {
    if (!g || !f || !(lr > 0)) return BL_ERR_ARG;
    int rc;
    double l[2];
    if ((rc = g->setup(f)) || (rc = g->loss_grads(f, l))) return rc;
    g->adam_t++;
    const float bc1 = 1.0f - std::pow(0.9f, (float)g->adam_t), bc2 = 1.0f - std::pow(0.999f, (float)g->adam_t);
    const hipStream_t st = blf::fem_view(f).st;
    hipLaunchKernelGGL(k_adam, grid(g->L.total), dim3(256), 0, st, g->params.p, g->grads.p, g->m.p, g->v2.p, g->L.total, lr, bc1, bc2);
    GCHK(hipGetLastError());
    GCHK(hipStreamSynchronize(st));
    if (losses) losses[0] = l[0], losses[1] = l[1];
    return BL_OK;
}

int bl_gnn_infer(bl_gnn* g, bl_fem* f, double* x0, double* s) {
    if (!g || !f) return BL_ERR_ARG;
    int rc;
    if ((rc = g->setup(f))) return rc;
    g->infer(f);
    const hipStream_t st = blf::fem_view(f).st;
    GCHK(hipStreamSynchronize(st));
    if (x0) GCHK(hipMemcpy(x0, g->x0.p, sizeof(double) * 6 * g->n, hipMemcpyDeviceToHost));
    if (s) GCHK(hipMemcpy(s, g->s.p, sizeof(double) * 6 * g->n, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_fem_step_pcg(bl_fem* f, bl_gnn* g, const float* loads, float h, float theta, float dissipation, const float gr[3],
                    const bl_pcg_opts* o, bl_pcg_report* rep) {
    if (!f || !g || !gr || !o || !(h > 0) || o->max_iter < 0 || !(o->tol > 0)) return BL_ERR_ARG;
    if (o->precond < BL_PRECOND_BLOCK_JACOBI || o->precond > BL_PRECOND_FACTOR || o->warm < BL_WARM_ZERO || o->warm > BL_WARM_NETWORK)
        return BL_ERR_ARG;
    const bool net_warm = o->warm == BL_WARM_NETWORK, net_prec = o->precond == BL_PRECOND_NETWORK;
    const bool fac_prec = o->precond == BL_PRECOND_FACTOR;
    if (blf::fem_plastic(f)) { // (its passes again need the direct solve: bl_fem_step)
        blg::set_last_error("the PCG step takes bodies of elastic sections only");
        return BL_ERR_UNSUPPORTED;
    }
    int rc;
    bl_pcg_report r{};
    float ms = 0;
    if ((rc = blf::fem_assemble(f, loads, h, theta, dissipation, gr, &r.assemble_ms))) return rc;
    if ((rc = g->setup(f))) return rc;
    const blf::FemView v = blf::fem_view(f);
    const hipStream_t st = v.st;
    const uint32_t n6 = 6 * g->n;
    // the warm start and the preconditioner
    GCHK(hipEventRecord(g->ev[0], st));
    if (net_warm || net_prec) g->infer(f);
    else if (!fac_prec) g->features(st, v); // (the block inverses)
    if (o->warm == BL_WARM_ZERO) hipLaunchKernelGGL(k_zero_d, grid(n6), dim3(256), 0, st, v.x, n6);
    else if (net_warm) hipLaunchKernelGGL(k_copy_d, grid(n6), dim3(256), 0, st, (const double*)g->x0.p, v.x, n6);
    // (BL_WARM_PREVIOUS: the engine's x still holds the last step's solution)
    std::vector<double> hx0;
    if (o->check_error) {
        hx0.resize(n6);
        GCHK(hipMemcpyAsync(hx0.data(), v.x, sizeof(double) * n6, hipMemcpyDeviceToHost, st));
    }
    GCHK(hipEventRecord(g->ev[1], st));
    int iters = 0, conv = 0;
    const bool had_factor = blf::fem_has_factor(f);
    if (fac_prec) {
        // (no factor yet: the step is a Cholesky step)
        if (had_factor && (rc = g->pcg_factor(f, v.x, o->max_iter, o->tol, &iters, &conv))) return rc;
    } else if ((rc = g->pcg(f, v.x, net_prec ? g->s.p : nullptr, o->max_iter, o->tol, &iters, &conv))) {
        return rc;
    }
    GCHK(hipEventRecord(g->ev[2], st));
    GCHK(hipEventSynchronize(g->ev[2]));
    GCHK(hipEventElapsedTime(&r.gnn_ms, g->ev[0], g->ev[1]));
    GCHK(hipEventElapsedTime(&r.pcg_ms, g->ev[1], g->ev[2]));
    r.iterations = iters, r.converged = conv;
    std::vector<double> hxp;
    // the fallback: PCG missed tol (or, preconditioned by the factor, there was none yet)
    const bool fallback = !conv && (o->fallback || (fac_prec && !had_factor));
    if (o->check_error || fallback) {
        // the supernodal Cholesky: the fallback, and the yardstick of the warm start's and the PCG's error (then
        // the factor the PCG used stays in force)
        const bool keep = !fallback && fac_prec;
        if (o->check_error) {
            hxp.resize(n6);
            GCHK(hipMemcpy(hxp.data(), v.x, sizeof(double) * n6, hipMemcpyDeviceToHost));
        }
        if (keep && (rc = blf::fem_keep_factor(f))) return rc;
        float fms = 0, sms = 0;
        int fail = 0;
        if ((rc = blf::fem_cholesky(f, &fms, &sms, &fail))) return rc;
        r.fallback_ms = fms + sms;
        if (fail) {
            blg::set_last_error("a pivot of the factorization is not positive");
            return BL_ERR_STATE;
        }
        if (keep && (rc = blf::fem_restore_factor(f))) return rc;
        if (o->check_error) {
            std::vector<double> hs(n6);
            GCHK(hipMemcpy(hs.data(), v.x, sizeof(double) * n6, hipMemcpyDeviceToHost));
            r.warm_error = rel_err(hx0, hs);
            r.pcg_error = rel_err(hxp, hs);
            if (!fallback) GCHK(hipMemcpy(v.x, hxp.data(), sizeof(double) * n6, hipMemcpyHostToDevice));
        }
        r.fallback_used = fallback ? 1 : 0;
    }
    if ((rc = blf::fem_residual(f, v.x, &r.residual))) return rc;
    if ((rc = blf::fem_update(f, &r.update_ms))) return rc;
    r.total_ms = r.assemble_ms + r.gnn_ms + r.pcg_ms + r.fallback_ms + r.update_ms;
    (void)ms;
    if (rep) *rep = r;
    return BL_OK;
}

} // extern "C"
