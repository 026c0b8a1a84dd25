"""The learned warm start's NumPy oracle (oracle/gnn_oracle.py) pinned by the mathematics: its analytic gradients
against central finite differences of its own loss (FP64, no rounding), on a plate's graph with a random SPD system;
and its bfloat16 mode within bfloat16 precision of the FP64 one."""
import numpy as np
import pytest

from beamlab_amd import fem
from oracle import gnn_oracle as G


def _graph(nu=6, nv=5):
    p, t, fx = fem.plate_mesh(nu, nv)
    free = np.nonzero(fx == 0)[0]
    fidx = -np.ones(len(p), int)
    fidx[free] = np.arange(len(free))
    nb = [set() for _ in free]
    for tri in t:
        for a in range(3):
            for c in range(3):
                i, j = fidx[tri[a]], fidx[tri[c]]
                if a != c and i >= 0 and j >= 0:
                    nb[i].add(j)
    off = np.zeros(len(free) + 1, np.int64)
    lst = []
    for i, s in enumerate(nb):
        lst += sorted(s)
        off[i + 1] = len(lst)
    nbr = np.array(lst, np.int64)
    dst = np.repeat(np.arange(len(free)), np.diff(off))
    ell = np.mean(np.linalg.norm(p[free[nbr]] - p[free[dst]], axis=1))
    rel = np.zeros((len(nbr), 4))
    rel[:, :3] = (p[free[nbr]] - p[free[dst]]) / ell
    return len(free), off, nbr, rel


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((6 * n, 6 * n)) / np.sqrt(6 * n)
    A = B @ B.T + np.eye(6 * n)
    D = np.stack([A[6 * i:6 * i + 6, 6 * i:6 * i + 6] for i in range(n)])
    b = rng.standard_normal((n, 6))
    xstar = np.linalg.solve(A, b.ravel()).reshape(n, 6)
    xprev = xstar + 0.3 * rng.standard_normal((n, 6))
    return A, D, b, xstar, xprev


def test_gradients_match_finite_differences():
    n, off, nbr, rel = _graph()
    A, D, b, xstar, xprev = _problem(n, 3)
    f, sigma, Dinv = G.features(D, b, xprev)
    p = {k: v.astype(np.float64) for k, v in G.init_params(5).items()}
    for i, k in enumerate(sorted(p)):  # (biases off zero: every path of the loss is live)
        p[k] = p[k] + 0.05 * np.random.default_rng(100 + i).standard_normal(p[k].shape)
    apply = lambda x: (A @ x.ravel()).reshape(-1, 6)
    L1, L2, g = G.loss_and_grads(p, f.astype(np.float64), off, nbr, rel, D, Dinv, b, sigma, xstar, apply, use_bf16=False)
    assert L1 > 0 and L2 > 0
    # the loss with the warm start's residual and the correction's target frozen at p (as the loss defines them)
    o0, _ = G.forward(p, f.astype(np.float64), off, nbr, rel, use_bf16=False)
    x0 = o0[:, 0:6] * sigma
    r0 = b - apply(x0)
    e = (xstar - x0) / sigma
    s0 = np.exp(np.clip(o0[:, 6:12], -3, 3))
    z0 = s0 * np.einsum("nij,nj->ni", Dinv, s0 * r0) / sigma
    kap = np.sum(z0 * e) / np.sum(z0 * z0)

    def loss(q):
        o, _ = G.forward(q, f.astype(np.float64), off, nbr, rel, use_bf16=False)
        l1 = np.mean((o[:, 0:6] - xstar / sigma) ** 2)
        s = np.exp(np.clip(o[:, 6:12], -3, 3))
        z = s * np.einsum("nij,nj->ni", Dinv, s * r0) / sigma
        return l1 + np.mean((kap * z - e) ** 2)

    assert loss(p) == pytest.approx(L1 + L2, rel=1e-12)
    rng = np.random.default_rng(9)
    gmax = max(np.abs(v).max() for v in g.values())
    checked = 0
    for name, shp in G.param_shapes():
        for _ in range(3):
            idx = tuple(int(rng.integers(0, s)) for s in shp)
            an = g[name][idx]
            # central differences at three steps: truncation grows with h, roundoff as h shrinks, and a ReLU kink
            # within h of a pre-activation spoils a step; a right gradient agrees at one of them at least
            errs = []
            for h0 in (1e-4, 1e-5, 1e-6):
                h = h0 * max(1.0, abs(p[name][idx]))
                q1 = {k: v.copy() for k, v in p.items()}
                q2 = {k: v.copy() for k, v in p.items()}
                q1[name][idx] += h
                q2[name][idx] -= h
                fd = (loss(q1) - loss(q2)) / (2 * h)
                # (relative, with a floor of 1e-6 of the largest component for the small ones)
                errs.append(abs(fd - an) / (max(abs(fd), abs(an)) + 0.1 * gmax))
            assert min(errs) <= 1e-5, (name, idx, an, errs)
            checked += 1
    assert checked == 3 * len(G.param_shapes())


def test_bf16_mode_close_to_fp64():
    n, off, nbr, rel = _graph(9, 7)
    A, D, b, xstar, xprev = _problem(n, 4)
    f, sigma, Dinv = G.features(D, b, xprev)
    p = G.init_params(1)
    apply = lambda x: (A @ x.ravel()).reshape(-1, 6)
    _, _, g16 = G.loss_and_grads(p, f, off, nbr, rel, D, Dinv, b, sigma, xstar, apply, use_bf16=True)
    p64 = {k: v.astype(np.float64) for k, v in p.items()}
    _, _, g64 = G.loss_and_grads(p64, f.astype(np.float64), off, nbr, rel, D, Dinv, b, sigma, xstar, apply, use_bf16=False)
    # bfloat16 keeps 8 bits of mantissa: the whole gradient within 5 % (1.3 % measured); each parameter's within 25 %
    # of its own size plus 1 % of the whole's (the edge offsets' weights U, with the smallest gradients, differ most)
    d = np.concatenate([(g16[k] - g64[k]).ravel() for k in g64])
    ref = np.concatenate([g64[k].ravel() for k in g64])
    assert np.linalg.norm(d) <= 0.05 * np.linalg.norm(ref)
    for k in g64:
        assert np.linalg.norm(g16[k] - g64[k]) <= 0.25 * np.linalg.norm(g64[k]) + 0.01 * np.linalg.norm(ref), k


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-40, 3.0e38], np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6)  # ties to even
    assert r[3] == -2.5 and (r.view(np.uint32) & 0xFFFF).max() == 0
