"""Test infrastructure only: the NumPy restatement of the learned warm start / preconditioner of the PCG solve (the
checker of beamlab_amd/csrc/bl_gnn.hip). The reference has no learned part; this module states the network, its loss
and its gradients, and is itself pinned by finite differences (tests/test_gnn_oracle.py).

The network over the free nodes of a shell mesh (an edge where two free nodes share a triangle):
    features f_i (32, of which 18 used): t_i / tau, xprev_i / sigma, b_i / beta, with t_i = D_i^-1 b_i (D_i the
        node's 6x6 diagonal block), tau / sigma / beta the root mean squares over the nodes of t / xprev / b, taken
        apart for the translations (dof 0-2) and the rotations (dof 3-5); sigma falls back to tau where xprev is zero
    h0 = relu(f W_e + b_e)                                           64 wide
    4 layers: P = h W_m;  agg_i = 1/deg_i sum_{j in N(i)} relu(P_j + rel_ij U + b_m)
              h <- h + relu([h, agg] W_u + b_u)
    o = h W_d + b_d                                                  16 wide: o[0:6] the warm start, o[6:12] log s
    x0_i = sigma * o_i[0:6];   s_i = exp(clip(o_i[6:12], -3, 3))
The PCG's preconditioner is S D^-1 S (S = diag(s), D the block diagonal).
Loss (the mean over the nodes' 6 dofs):
    L1 = mean (o[0:6] - x*/sigma)^2                                  the warm start against the Cholesky solution
    L2 = mean (k z - e)^2, z = s * D^-1 (s * r0) / sigma, e = (x* - x0) / sigma, k = <z, e> / <z, z>
         the preconditioned residual of the warm start against its error, up to the best global factor k (CG does
         not see a preconditioner's scale); r0 = b - A x0, e and k taken as constants
The matrix products take their operands rounded to bfloat16 (round to nearest even) and sum in FP32, as the
kernels do on the matrix cores; `bf16=False` computes in FP64 without rounding (the finite-difference check).
"""
import numpy as np

H, FIN, FOUT, NL = 64, 32, 16, 4


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(x), x, out)


def param_shapes():
    shapes = [("W_e", (FIN, H)), ("b_e", (H,))]
    for l in range(NL):
        shapes += [(f"W_m{l}", (H, H)), (f"U{l}", (4, H)), (f"b_m{l}", (H,)), (f"W_u{l}", (2 * H, H)), (f"b_u{l}", (H,))]
    shapes += [("W_d", (H, FOUT)), ("b_d", (FOUT,))]
    return shapes


def init_params(seed=0):
    """He-normal weights, zero biases; the last layer scaled down (the network starts near a zero warm start)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in param_shapes():
        if name.startswith("b"):
            p[name] = np.zeros(shp, np.float32)
        else:
            fan = shp[0]
            w = rng.standard_normal(shp) * np.sqrt(2.0 / fan)
            if name == "W_d":
                w *= 0.1
            p[name] = w.astype(np.float32)
    return p


def flat(p):
    return np.concatenate([p[n].ravel() for n, _ in param_shapes()]).astype(np.float32)


def unflat(v):
    p, o = {}, 0
    for n, shp in param_shapes():
        k = int(np.prod(shp))
        p[n] = np.asarray(v[o:o + k], np.float32).reshape(shp)
        o += k
    return p


def features(D, b, xprev):
    """f (n x 32), sigma (6), from the blocks D (n x 6 x 6), the right side b and the previous solution (n x 6)."""
    Dinv = np.linalg.inv(D)
    t = np.einsum("nij,nj->ni", Dinv, b)
    def rms2(v):
        return np.array([np.sqrt(np.mean(v[:, :3] ** 2)), np.sqrt(np.mean(v[:, 3:] ** 2))])
    st, sb, sx = rms2(t), rms2(b), rms2(xprev)
    sx = np.where(sx > 0, sx, st)
    tau, sigma, beta = np.repeat(st + 1e-30, 3), np.repeat(sx + 1e-30, 3), np.repeat(sb + 1e-30, 3)
    f = np.zeros((len(b), FIN))
    f[:, 0:6] = t / tau
    f[:, 6:12] = xprev / sigma
    f[:, 12:18] = b / beta
    return f.astype(np.float32), sigma, Dinv


def _mm(a, w, use_bf16):
    if use_bf16:
        return (bf16(a).astype(np.float64) @ bf16(w).astype(np.float64)).astype(np.float32)
    return a @ w


def forward(p, f, nbr_off, nbr, rel, use_bf16=True):
    """The network's output o (n x 16) and the intermediates the backward pass needs."""
    dt = np.float32 if use_bf16 else np.float64
    n = len(f)
    deg = np.diff(nbr_off).astype(dt)
    dst = np.repeat(np.arange(n), np.diff(nbr_off))
    cache = {"f": f}
    z0 = _mm(f.astype(dt), p["W_e"].astype(dt), use_bf16) + p["b_e"]
    h = np.maximum(z0, 0).astype(dt)
    cache["z0"] = z0
    for l in range(NL):
        P = _mm(h, p[f"W_m{l}"].astype(dt), use_bf16)
        q = P[nbr] + (rel.astype(dt) @ p[f"U{l}"].astype(dt)) + p[f"b_m{l}"]
        msg = np.maximum(q, 0)
        agg = np.zeros((n, H), dt)
        np.add.at(agg, dst, msg)
        agg = agg / np.maximum(deg, 1)[:, None]
        cat = np.concatenate([h, agg], 1)
        zu = _mm(cat, p[f"W_u{l}"].astype(dt), use_bf16) + p[f"b_u{l}"]
        cache[f"h{l}"], cache[f"q{l}"], cache[f"cat{l}"], cache[f"zu{l}"] = h, q, cat, zu
        h = (h + np.maximum(zu, 0)).astype(dt)
    cache["hL"] = h
    o = _mm(h, p["W_d"].astype(dt), use_bf16) + p["b_d"]
    return o.astype(dt), cache


def loss_and_grads(p, f, nbr_off, nbr, rel, D, Dinv, b, sigma, xstar, A_apply, use_bf16=True):
    """(L1, L2, grads) of one sample; A_apply(x) = A x (FP64) for the residual of the warm start."""
    dt = np.float32 if use_bf16 else np.float64
    n = len(f)
    o, c = forward(p, f, nbr_off, nbr, rel, use_bf16)
    yhat = o[:, 0:6].astype(np.float64)
    ystar = xstar / sigma
    x0 = yhat * sigma
    L1 = np.mean((yhat - ystar) ** 2)
    lo = o[:, 6:12].astype(np.float64)
    s = np.exp(np.clip(lo, -3, 3))
    r0 = b - A_apply(x0)
    e = (xstar - x0) / sigma
    u = np.einsum("nij,nj->ni", Dinv, s * r0)
    z = s * u / sigma
    kap = np.sum(z * e) / np.sum(z * z)
    L2 = np.mean((kap * z - e) ** 2)
    do = np.zeros((n, FOUT), np.float64)
    do[:, 0:6] = 2 * (yhat - ystar) / (6 * n)
    dz = 2 * kap * (kap * z - e) / (6 * n)
    ds = dz * u / sigma + r0 * np.einsum("nk,nkm->nm", dz * s / sigma, Dinv)
    inside = (lo > -3) & (lo < 3)
    do[:, 6:12] = np.where(inside, ds * s, 0.0)
    g = backward(p, c, do.astype(dt), nbr_off, nbr, rel, use_bf16)
    return L1, L2, g


def backward(p, c, do, nbr_off, nbr, rel, use_bf16=True):
    dt = do.dtype
    n = len(do)
    deg = np.maximum(np.diff(nbr_off), 1).astype(dt)
    dst = np.repeat(np.arange(n), np.diff(nbr_off))
    g = {}
    hL = c["hL"]
    g["W_d"] = _mm(hL.T, do, use_bf16)
    g["b_d"] = do.sum(0)
    dh = _mm(do, p["W_d"].T.astype(dt), use_bf16)
    for l in reversed(range(NL)):
        zu = c[f"zu{l}"]
        dzu = dh * (zu > 0)
        g[f"W_u{l}"] = _mm(c[f"cat{l}"].T, dzu, use_bf16)
        g[f"b_u{l}"] = dzu.sum(0)
        dcat = _mm(dzu, p[f"W_u{l}"].T.astype(dt), use_bf16)
        dh_prev = dh + dcat[:, :H]
        dagg = dcat[:, H:] / deg[:, None]
        q = c[f"q{l}"]
        dq = dagg[dst] * (q > 0)
        g[f"U{l}"] = rel.T.astype(dt) @ dq
        g[f"b_m{l}"] = dq.sum(0)
        dP = np.zeros((n, H), dt)
        np.add.at(dP, nbr, dq)
        g[f"W_m{l}"] = _mm(c[f"h{l}"].T, dP, use_bf16)
        dh = dh_prev + _mm(dP, p[f"W_m{l}"].T.astype(dt), use_bf16)
    dz0 = dh * (c["z0"] > 0)
    g["W_e"] = _mm(c["f"].T.astype(dt), dz0, use_bf16)
    g["b_e"] = dz0.sum(0)
    return g
