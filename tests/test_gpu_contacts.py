"""Inter-body and self contacts on the GPU against the oracle, bit for bit (node p and v, beam state, break
counts) at several checkpoints, with the contact counters (pair rebuilds, candidate pairs, narrow-phase tests,
contacts) equal too - hull (one-sided) triangles and fast bodies among the cases; and the refusals of what the
GPU path does not step."""
import numpy as np
import pytest

from _cases import assert_same, run_world
from _contact_cases import CONTACT_CASES, checkpoints
from beamlab_amd import abi, scenes
from beamlab_amd.world import World

pytestmark = pytest.mark.gpu

TUNINGS = {
    "auto": dict(mode=abi.MODE_AUTO),
    "stepped_1": dict(mode=abi.MODE_STEPPED, block=256, use_graph=1, incidences_per_thread=1),
    "stepped_1_launches": dict(mode=abi.MODE_STEPPED, block=256, use_graph=0, incidences_per_thread=1),
    "persistent": dict(mode=abi.MODE_PERSISTENT, block=1024, substeps_per_launch=50),
}
COUNTERS = ("contacts", "pair_rebuilds", "contact_pairs", "narrow_tests", "static_contacts", "broken_beams")


def gpu_world(settings, boxes, tuning):
    w = World(settings)
    w.statics.boxes = boxes
    w.set_tuning(**tuning)
    return w


@pytest.mark.parametrize("tuning", sorted(TUNINGS))
@pytest.mark.parametrize("name", sorted(CONTACT_CASES))
def test_gpu_contacts_match_oracle(name, tuning, oracle_mod):
    print(f"contact case {name} tuning {tuning}", flush=True)
    settings, boxes, bodies, n = CONTACT_CASES[name]()
    cps = checkpoints(name, n)
    ow = oracle_mod.OracleWorld(settings, boxes)
    so = run_world(ow, bodies, n, cps)
    gw = gpu_world(settings, boxes, TUNINGS[tuning])
    sg = run_world(gw, bodies, n, cps)
    assert_same(sg, so, f"gpu[{tuning}] vs oracle [{name}]")
    st_o, st_g = ow.stats(), gw.stats()
    for key in COUNTERS:
        assert st_g[key] == st_o[key], (key, st_g[key], st_o[key])
    # (behind a two-sided triangle the ball is out of the thin surface's reach: no contact at all)
    assert (st_g["contacts"] == 0) if name == "hull_through" else (st_g["contacts"] > 100)


def test_contacts_long_frames(oracle_mod):
    """Frames longer than the pair interval (several rebuilds inside one call) and a frame of one substep."""
    settings, boxes, bodies, _ = CONTACT_CASES["lattice_pile"]()
    cps = [1, 2, 100, 333, 700]
    ow = oracle_mod.OracleWorld(settings, boxes)
    so = run_world(ow, bodies, 700, cps)
    gw = gpu_world(settings, boxes, TUNINGS["auto"])
    sg = run_world(gw, bodies, 700, cps)
    assert_same(sg, so, "gpu vs oracle [lattice_pile, long frames]")
    assert gw.stats()["pair_rebuilds"] == ow.stats()["pair_rebuilds"]


def test_contact_scene_full_size(oracle_mod):
    """The 64x32x24 lattice with 2 000 balls at full size over 660 substeps (40 pair rebuilds; the balls on the block
    rest on it from the first substep), bit for bit against the oracle. Its compactions span hundreds of workgroups (node hash,
    node-triangle search), so the look-back crosses many 64-workgroup windows; the scene's 2 001 bodies make many
    pair islands."""
    from _cases import GROUND, _boxes
    print("contact case config2 full size", flush=True)
    settings, boxes, bodies, n = abi.default_settings(), _boxes(GROUND), scenes.config2_scene(), 660
    cps = [1, 17, 330, 660]
    ow = oracle_mod.OracleWorld(settings, boxes)
    so = run_world(ow, bodies, n, cps)
    gw = gpu_world(settings, boxes, TUNINGS["stepped_1_launches"])
    sg = run_world(gw, bodies, n, cps)
    assert_same(sg, so, "gpu vs oracle [config2 full size]")
    st_o, st_g = ow.stats(), gw.stats()
    for key in COUNTERS:
        assert st_g[key] == st_o[key], (key, st_g[key], st_o[key])
    assert st_g["contacts"] > 0 and st_g["narrow_tests"] > 0


def test_jointed_capsules_refused():
    w = World(abi.default_settings())
    ball = scenes.ball_body((0.0, 1.5, 0.0), 0.1, 2.0)
    ball.capsules["joint"] = 0
    with pytest.raises(abi.BeamlabError) as e:
        w.add_body(ball)
    assert e.value.status == abi.BL_ERR_UNSUPPORTED
