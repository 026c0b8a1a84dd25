"""Writes the golden fixtures of the beam path from the compiled reference (oracle/_ref/libblref.so, built from
the reference's sources by oracle/Makefile): the reference's own World::step_substeps on the seeded scenes of
tests/_cases.py. Run in the container that holds the reference sources:

    python tests/golden/make_golden.py

Outputs (data only: inputs are regenerated from the seeds, outputs are the reference's states):
    c1_ref.npz          config-1 lattice: node and beam records after 1, 7, 666 and 2000 substeps, broken counts
    cases_ref.npz       every other case (the terrain cases with tests/_cases.py's terrain and posts): sha256 of the
                        final node and beam records, broken counts
    fem_ref.npz         the shell element: the reference's FemFrame::tri_stiffness (K, 18 x 18 FP64, and the rest
                        shape X) of 24 seeded triangles under three sections
    fem_step_ref.npz    the plate body's whole substep (compute_forces, solve, integrate_nodes, sync_positions) of the
                        scenarios of tests/_fem_cases.py: the update-path scenario after every step, the spun plate
                        every 100 steps (nodes' p and v, xd, w, q, forces, counters)
    fem_frame_ref.npz   frame bodies with members (tests/test_fem_oracle.py's frame cases: cantilevers on every joint,
                        portals, a stiffened plate) under seeded loads and moments: the states after 10, 20, 30 steps
    fem_plastic_ref.npz frame bodies that yield (tests/_fem_cases.py's plastic frames: members in tension, torsion and
                        at their hinges, a chain that spends the passes-again budget, plates whose shells flow) under
                        loads past their limits: the states and the plastic state after 20, 40, 60 steps

    fem_held_ref.npz    frames stepping their frame every 2 and 4 substeps (tests/_fem_cases.py's held cases): the
                        states after 12, 24, 36 substeps

`python tests/golden/make_golden.py plastic` writes fem_plastic_ref.npz and fem_held_ref.npz alone, `... cases`
c1_ref.npz and cases_ref.npz alone.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _fem_cases as fc  # noqa: E402
from _cases import CPU_CASES, TERRAIN_CASES, run_world, terrain_statics  # noqa: E402
from _contact_cases import CONTACT_CASES, golden_checkpoints  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    if not oracle.have_ref():
        oracle.build(ref=True)
    cases()
    fem()
    plastic()


def cases():
    settings, boxes, bodies, n = CPU_CASES["c1"]()
    cps = [1, 7, n // 3, n]
    st = run_world(oracle.RefWorld(settings, boxes), bodies, n, cps)
    out = {"checkpoints": np.array(cps)}
    for k, s in enumerate(st):
        out[f"nodes_{k}"], out[f"beams_{k}"] = s[0]
        out.setdefault("broken", []).append(s[-1][0])
    out["broken"] = np.array(out["broken"])
    np.savez_compressed(os.path.join(HERE, "c1_ref.npz"), **out)
    digests = {}
    for name in sorted(CPU_CASES) + sorted(CONTACT_CASES) + sorted(TERRAIN_CASES):
        settings, boxes, bodies, n = (CPU_CASES.get(name) or CONTACT_CASES.get(name) or TERRAIN_CASES[name])()
        st = run_world(oracle.RefWorld(settings, boxes), bodies, n, golden_checkpoints(name, n),
                       terrain_statics() if name in TERRAIN_CASES else None)
        h = hashlib.sha256()
        for nodes, beams in st[-1][:-1]:
            h.update(nodes.tobytes())
            h.update(beams.tobytes())
        digests[f"{name}_sha256"] = np.array(h.hexdigest())
        digests[f"{name}_broken"] = np.array(st[-1][-1])
    np.savez_compressed(os.path.join(HERE, "cases_ref.npz"), **digests)
    print("written", sorted(digests))


def fem():
    rng = np.random.default_rng(17)
    sections = np.array([oracle.SECTION_C3, (2.1e11, 0.3, 7850.0, 1e-3, 0.05, 3e-4), (5e9, 0.45, 1200.0, 4e-3, 0.2, 0.0)],
                        np.float32)
    P, K, X = [], [], []
    while len(P) < 24:
        p = rng.uniform(-1, 1, (3, 3)).astype(np.float32) * np.float32(0.05 + len(P) / 10)
        if np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0])) < 1e-3:
            continue
        k, x = oracle.ref_tri_stiffness(p, tuple(sections[len(P) % 3]))
        P.append(p), K.append(k), X.append(x)
    np.savez_compressed(os.path.join(HERE, "fem_ref.npz"), p=np.array(P), K=np.array(K), X=np.array(X),
                        section=np.array([(i % 3) for i in range(24)]), sections=sections)
    print("written fem_ref.npz")
    out = {"spun_steps": np.array(400)}
    for name, sc, every in (("events", fc.events(), 1), ("spun", fc.spinning(400), 100)):
        R = fc.make(oracle, sc, ref=True)
        for s, st in fc.apply(R, sc, ref=True):
            if s % every == every - 1:
                for k in fc.STATE_KEYS:
                    out[f"{name}_{s}_{k}"] = st[k]
                out[f"{name}_{s}_counters"] = np.array([st["failures"], st["clamps"]])
        R.close()
    np.savez_compressed(os.path.join(HERE, "fem_step_ref.npz"), **out)
    print("written fem_step_ref.npz")
    out = {}
    for name, fr in sorted(fc.frames().items()):
        R = fc.make_frame(oracle, fr, ref=True)
        for s, (ld, tq) in enumerate(fc.frame_loads(name, R.nn, 30, 3)):
            R.add_torques(tq)
            R.substep(ld, fc.H)
            if s % 10 == 9:
                st = R.read()
                for k in fc.STATE_KEYS:
                    out[f"{name}_{s}_{k}"] = st[k]
                out[f"{name}_{s}_counters"] = np.array([st["failures"], st["clamps"]])
        R.close()
    np.savez_compressed(os.path.join(HERE, "fem_frame_ref.npz"), **out)
    print("written fem_frame_ref.npz")


PLASTIC_KEYS = ("X", "th0", "r0", "tdmg", "tutil", "mstate", "qa", "qb")


def plastic():
    out = {}
    for name, fr in sorted(fc.plastic_frames().items()):
        R = fc.make_plastic(oracle, fr, ref=True)
        for s, (ld, tq) in enumerate(fc.plastic_loads(name, R.nn, 60, 5)):
            R.add_torques(tq)
            R.substep(ld, fc.H)
            if s % 20 == 19:
                st, pl = R.read(), R.plastic()
                for k in fc.STATE_KEYS:
                    out[f"{name}_{s}_{k}"] = st[k]
                for k in PLASTIC_KEYS:
                    out[f"{name}_{s}_{k}"] = pl[k]
                out[f"{name}_{s}_counters"] = np.array([st["failures"], st["clamps"], pl["repasses"], pl["passes"]])
        R.close()
    np.savez_compressed(os.path.join(HERE, "fem_plastic_ref.npz"), **out)
    print("written fem_plastic_ref.npz")
    out = {}
    for every, start in ((2, 0), (4, 1)):
        for name, make, loads in fc.held_cases():
            R = make(oracle, ref=True)
            for j, (ld, tq) in enumerate(loads(R.nn, 36)):
                R.add_torques(tq)
                R.cycle(ld, fc.H, every, j + start)
                if j % 12 == 11:
                    st = R.read()
                    for k in fc.STATE_KEYS:
                        out[f"{name}_{every}_{j}_{k}"] = st[k]
                    out[f"{name}_{every}_{j}_counters"] = np.array([st["failures"], st["clamps"]])
            R.close()
    np.savez_compressed(os.path.join(HERE, "fem_held_ref.npz"), **out)
    print("written fem_held_ref.npz")


if __name__ == "__main__":
    if sys.argv[1:] == ["plastic"]:
        plastic()
    elif sys.argv[1:] == ["cases"]:
        cases()
    else:
        main()
