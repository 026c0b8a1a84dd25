"""The oracle against the compiled reference's golden fixtures (tests/golden, written by make_golden.py): runs
where the reference sources are absent (the GPU box), pinning the oracle the GPU tests compare with."""
import hashlib
import os

import numpy as np
import pytest

from _cases import CPU_CASES, TERRAIN_CASES, run_world, terrain_statics
from _contact_cases import CONTACT_CASES, golden_checkpoints

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_c1_states(oracle_mod):
    g = np.load(os.path.join(GOLD, "c1_ref.npz"))
    settings, boxes, bodies, n = CPU_CASES["c1"]()
    cps = [int(c) for c in g["checkpoints"]]
    st = run_world(oracle_mod.OracleWorld(settings, boxes), bodies, n, cps)
    for k, s in enumerate(st):
        nodes, beams = s[0]
        assert nodes.tobytes() == g[f"nodes_{k}"].tobytes(), f"nodes at {cps[k]}"
        assert beams.tobytes() == g[f"beams_{k}"].tobytes(), f"beams at {cps[k]}"
        assert s[-1][0] == int(g["broken"][k])
    assert int(g["broken"][-1]) > 0, "the fixture covers breaking"


@pytest.mark.parametrize("name", sorted(CPU_CASES) + sorted(CONTACT_CASES) + sorted(TERRAIN_CASES))
def test_case_digests(name, oracle_mod):
    g = np.load(os.path.join(GOLD, "cases_ref.npz"))
    settings, boxes, bodies, n = (CPU_CASES.get(name) or CONTACT_CASES.get(name) or TERRAIN_CASES[name])()
    st = run_world(oracle_mod.OracleWorld(settings, boxes), bodies, n, golden_checkpoints(name, n),
                   terrain_statics() if name in TERRAIN_CASES else None)
    h = hashlib.sha256()
    for nodes, beams in st[-1][:-1]:
        h.update(nodes.tobytes())
        h.update(beams.tobytes())
    assert h.hexdigest() == str(g[f"{name}_sha256"])
    assert list(st[-1][-1]) == list(g[f"{name}_broken"])
