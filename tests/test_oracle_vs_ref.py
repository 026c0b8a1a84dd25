"""The oracle (the C restatement) against the compiled reference, bit for bit, on every seeded case.

The checkpoints split the run into several World::step_substeps calls, so that the per-call post-frame work
(the top-speed and bounds check, repair_nodes) is exercised too.
"""
import numpy as np
import pytest

from _cases import CPU_CASES, assert_same, run_world
from _contact_cases import CONTACT_CASES, checkpoints


@pytest.mark.parametrize("name", sorted(CPU_CASES))
def test_oracle_matches_reference(name, ref_mod):
    settings, boxes, bodies, n = CPU_CASES[name]()
    cps = sorted({1, 7, n // 3, n})
    orc = ref_mod.OracleWorld(settings, boxes)
    ref = ref_mod.RefWorld(settings, boxes, threads=1)
    so = run_world(orc, bodies, n, cps)
    sr = run_world(ref, bodies, n, cps)
    assert_same(so, sr, f"oracle vs reference [{name}]")


@pytest.mark.parametrize("name", sorted(CONTACT_CASES))
def test_oracle_contacts_match_reference(name, ref_mod):
    settings, boxes, bodies, n = CONTACT_CASES[name]()
    cps = checkpoints(name, n)
    ow = ref_mod.OracleWorld(settings, boxes)
    so = run_world(ow, bodies, n, cps)
    sr = run_world(ref_mod.RefWorld(settings, boxes, threads=1), bodies, n, cps)
    assert_same(so, sr, f"oracle vs reference [{name}]")
    st = ow.stats()
    assert st["pair_rebuilds"] > 0, st
    # (behind a two-sided triangle the ball is out of the thin surface's reach: no contact at all)
    assert (st["contacts"] == 0) if name == "hull_through" else (st["contacts"] > 100), st
    if name == "fast_bodies":
        assert ow.fast_pairs() > 100


def test_reference_threads_bitwise(ref_mod):
    """The reference is bitwise independent of its thread count (README.md:1301): the baseline may use all cores."""
    settings, boxes, bodies, n = CPU_CASES["multi"]()
    r1 = ref_mod.RefWorld(settings, boxes, threads=1)
    s1 = run_world(r1, bodies, 200, [200])
    # a second world in the same process shares the job system: threads fixed at its first creation
    r2 = ref_mod.RefWorld(settings, boxes, threads=1)
    s2 = run_world(r2, bodies, 200, [200])
    assert_same(s1, s2, "reference run twice")
