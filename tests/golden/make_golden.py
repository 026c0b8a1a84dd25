"""Writes the golden fixtures of the beam path from the compiled reference (oracle/_ref/libblref.so, built from
the reference's sources by oracle/Makefile): the reference's own World::step_substeps on the seeded scenes of
tests/_cases.py. Run in the container that holds the reference sources:

    python tests/golden/make_golden.py

Outputs (data only: inputs are regenerated from the seeds, outputs are the reference's states):
    c1_ref.npz          config-1 lattice: node and beam records after 1, 7, 666 and 2000 substeps, broken counts
    cases_ref.npz       every other case: sha256 of the final node and beam records, broken counts
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from _cases import CPU_CASES, run_world  # noqa: E402
from _contact_cases import CONTACT_CASES  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    if not oracle.have_ref():
        oracle.build(ref=True)
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
    for name in sorted(CPU_CASES) + sorted(CONTACT_CASES):
        settings, boxes, bodies, n = (CPU_CASES.get(name) or CONTACT_CASES[name])()
        st = run_world(oracle.RefWorld(settings, boxes), bodies, n, sorted({1, 7, n // 3, n}))
        h = hashlib.sha256()
        for nodes, beams in st[-1][:-1]:
            h.update(nodes.tobytes())
            h.update(beams.tobytes())
        digests[f"{name}_sha256"] = np.array(h.hexdigest())
        digests[f"{name}_broken"] = np.array(st[-1][-1])
    np.savez_compressed(os.path.join(HERE, "cases_ref.npz"), **digests)
    print("written", sorted(digests))


if __name__ == "__main__":
    main()
