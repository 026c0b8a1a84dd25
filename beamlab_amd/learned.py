"""The learned warm start and preconditioner of the implicit FEM solve (bl_gnn_*, bl_fem_step_pcg): the network's
weights, its training on the engine's own Cholesky solutions, and the PCG step it feeds.

Layer order of the weights (float32, flat): W_e (32 x 64), b_e; for each of the 4 layers W_m (64 x 64), U (4 x 64),
b_m, W_u (128 x 64), b_u; W_d (64 x 16), b_d - the layout bl_gnn_create takes.
"""
import ctypes
import os

import numpy as np

from . import abi

HIDDEN, FEATURES, OUTPUTS, LAYERS = 64, 32, 16, 4
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gnn_plate.npz")


def shapes():
    s = [("W_e", (FEATURES, HIDDEN)), ("b_e", (HIDDEN,))]
    for l in range(LAYERS):
        s += [(f"W_m{l}", (HIDDEN, HIDDEN)), (f"U{l}", (4, HIDDEN)), (f"b_m{l}", (HIDDEN,)),
              (f"W_u{l}", (2 * HIDDEN, HIDDEN)), (f"b_u{l}", (HIDDEN,))]
    return s + [("W_d", (HIDDEN, OUTPUTS)), ("b_d", (OUTPUTS,))]


def initial_weights(seed=0):
    """He-normal weights, zero biases, the output layer scaled by 0.1 (the network starts near x0 = 0, s = 1)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in shapes():
        if name.startswith("b"):
            parts.append(np.zeros(int(np.prod(shp)), np.float32))
        else:
            w = rng.standard_normal(shp) * np.sqrt(2.0 / shp[0])
            parts.append((w * (0.1 if name == "W_d" else 1.0)).astype(np.float32).ravel())
    return np.concatenate(parts)


class Network:
    """The network on the device (bl_gnn): weights, Adam state, the PCG solve's workspace."""

    def __init__(self, weights=None, seed=0):
        self._lib = abi.load()
        w = initial_weights(seed) if weights is None else np.ascontiguousarray(weights, np.float32)
        self.n = self._lib.bl_gnn_param_count()
        if len(w) != self.n:
            raise abi.BeamlabError(f"{len(w)} weights for a network of {self.n}")
        h = ctypes.c_void_p()
        abi.check(self._lib.bl_gnn_create(ctypes.byref(h), abi.ptr(w), self.n), "bl_gnn_create")
        self._h = h
        abi.track(self)

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path) as z:
            return cls(z["weights"])

    def save(self, path=WEIGHTS, **meta):
        np.savez_compressed(path, weights=self.weights(), **meta)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bl_gnn_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def weights(self):
        out = np.zeros(self.n, np.float32)
        abi.check(self._lib.bl_gnn_get_params(self._h, abi.ptr(out), self.n), "bl_gnn_get_params")
        return out

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        abi.check(self._lib.bl_gnn_set_params(self._h, abi.ptr(w), self.n), "bl_gnn_set_params")

    def gradients(self, body):
        """(grads, (L1, L2)) for the shell body's current system and Cholesky solution."""
        g = np.zeros(self.n, np.float32)
        losses = np.zeros(2, np.float64)
        abi.check(self._lib.bl_gnn_gradients(self._h, body._h, abi.ptr(g), self.n, abi.ptr(losses)), "bl_gnn_gradients")
        return g, tuple(losses)

    def train_step(self, body, lr=1e-3):
        losses = np.zeros(2, np.float64)
        abi.check(self._lib.bl_gnn_train_step(self._h, body._h, lr, abi.ptr(losses)), "bl_gnn_train_step")
        return tuple(losses)

    def infer(self, body):
        """(x0, s) for the body's current system: 6 values per free node each."""
        nf = int((body.fixed == 0).sum()) if body.fixed is not None else body.nn
        x0 = np.zeros(6 * nf, np.float64)
        s = np.zeros(6 * nf, np.float64)
        abi.check(self._lib.bl_gnn_infer(self._h, body._h, abi.ptr(x0), abi.ptr(s)), "bl_gnn_infer")
        return x0, s


def pcg_step(body, net, loads=None, h=0.0005, max_iter=50, tol=1e-6, warm=abi.WARM_NETWORK,
             precond=abi.PRECOND_NETWORK, fallback=True, check_error=False, theta=0.5, dissipation=0.02,
             g=(0.0, -9.81, 0.0)):
    """One implicit step of the shell body solved by PCG from the network's warm start; the Cholesky as the fallback."""
    o = abi.PcgOpts()
    o.max_iter, o.warm, o.precond, o.fallback, o.check_error, o.tol = max_iter, warm, precond, int(fallback), \
        int(check_error), tol
    r = abi.PcgReport()
    ld = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(-1)
    gg = np.asarray(g, np.float32)
    abi.check(body._lib.bl_fem_step_pcg(body._h, net._h, abi.ptr(ld), h, theta, dissipation, abi.ptr(gg),
                                        ctypes.byref(o), ctypes.byref(r)), "bl_fem_step_pcg")
    return r.as_dict()


def plate_loads(nn, fixed, rng, count, sigma=50.0):
    """Seeded random point loads normal to the plate: `count` free nodes, N(0, sigma) N each."""
    free = np.nonzero(fixed == 0)[0]
    out = np.zeros((nn, 3), np.float32)
    idx = rng.choice(free, count, replace=False)
    out[idx, 1] = rng.normal(0.0, sigma, count).astype(np.float32)
    return out


def train(body, net, steps, lr=1e-3, loads_per_step=4, seed=0, h=0.0005, log=None):
    """Online training on the body's own steps: each step solved by the Cholesky (a triple: the system, the previous
    solution, the solution), then one Adam step of the network on it. Returns the losses per step."""
    rng = np.random.default_rng(seed)
    hist = []
    for s in range(steps):
        body.step(plate_loads(body.nn, body.fixed, rng, loads_per_step), h)
        hist.append(net.train_step(body, lr))
        if log and (s + 1) % 500 == 0:
            log(f"train step {s + 1}/{steps}: loss {hist[-1][0]:.4e} + {hist[-1][1]:.4e}")
    return np.array(hist)


def evaluate(cells, steps, net, seed=99, loads_per_step=None, max_iter=5000, tol=1e-6, warmup=3, methods=None, log=None):
    """The network's warm start and preconditioner against the method without it, on `steps` held-out steps of a plate:
    one body per method, stepped in lockstep under the same seeded loads (the same history up to the solver's
    tolerance). Per method: warm-start errors and PCG iterations (the Cholesky judging each solve)."""
    from . import fem
    methods = methods or {"network": (abi.WARM_NETWORK, abi.PRECOND_NETWORK),
                          "network_warm_bj": (abi.WARM_NETWORK, abi.PRECOND_BLOCK_JACOBI),
                          "previous_bj": (abi.WARM_PREVIOUS, abi.PRECOND_BLOCK_JACOBI),
                          "zero_bj": (abi.WARM_ZERO, abi.PRECOND_BLOCK_JACOBI)}
    bodies = {k: fem.ShellBody.plate(cells, cells) for k in methods}
    k0 = next(iter(bodies))
    count = loads_per_step or max(1, (32 * cells * cells) // (256 * 256))
    rng = np.random.default_rng(seed)
    for _ in range(warmup):
        loads = plate_loads(bodies[k0].nn, bodies[k0].fixed, rng, count)
        for b in bodies.values():
            b.step(loads, 0.0005)
    out = {k: {"warm_error": [], "iterations": [], "converged": []} for k in methods}
    for s in range(steps):
        loads = plate_loads(bodies[k0].nn, bodies[k0].fixed, rng, count)
        for k, (warm, prec) in methods.items():
            r = pcg_step(bodies[k], net, loads, max_iter=max_iter, tol=tol, warm=warm, precond=prec, check_error=True)
            out[k]["warm_error"].append(r["warm_error"])
            out[k]["iterations"].append(r["iterations"])
            out[k]["converged"].append(r["converged"])
        if log and (s + 1) % 10 == 0:
            log(f"evaluate step {s + 1}/{steps}")
    for b in bodies.values():
        b.close()
    return {k: {m: float(np.mean(v)) for m, v in d.items()} for k, d in out.items()}


def main():
    import argparse
    import sys
    import time
    from . import fem
    ap = argparse.ArgumentParser(description="train the learned warm start on a plate's own Cholesky steps")
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--eval-cells", type=int, default=32)
    ap.add_argument("--eval-steps", type=int, default=20)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    body = fem.ShellBody.plate(a.cells, a.cells)
    net = Network(initial_weights(a.seed))
    t0 = time.perf_counter()
    count = max(1, (32 * a.cells * a.cells) // (256 * 256))
    hist = train(body, net, a.steps, a.lr, loads_per_step=count, seed=a.seed + 1, log=log)
    dt = time.perf_counter() - t0
    ev = evaluate(a.eval_cells, a.eval_steps, net, log=log)
    net.save(a.out, losses=hist, cells=a.cells, steps=a.steps, lr=a.lr, train_s=dt)
    import json
    print(json.dumps({"train_steps": a.steps, "cells": a.cells, "train_s": dt, "loss_first": hist[:100].mean(0).tolist(),
                      "loss_last": hist[-100:].mean(0).tolist(), "evaluate": ev}), flush=True)


if __name__ == "__main__":
    main()
