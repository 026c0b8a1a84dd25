"""The static world beyond boxes: the heightfield terrain (Heightfield::sample, surface_at, max_height,
src/phys/static_world.cpp:189-231) and vertical cylinders (StaticWorld::collide_point, :364-474), with
World::collide_static's terrain test per body and frame (src/phys/world.cpp:743-756,812).

CPU: the oracle against the compiled reference, bit for bit, on lattices over hills, down a ridge's 63 degree face,
onto and into posts and through a hole in the ground, and on a car on the terrain beside a box; and what the
geometry fixes: a node comes to rest on a sloped plane at its radius from it, the terrain's surface is its nearest
vertex's (ice lets a node slide far further than concrete), a node rests on a post's top at its radius above it
and is stopped at the post's side, a node over a hole falls freely and is frozen below kill_y. GPU: the HIP path
against the oracle, bit for bit, in every launch form, counters included.
"""
import numpy as np
import pytest

from _cases import TERRAIN_CASES, assert_same, run_world, terrain_statics
from beamlab_amd import abi
from beamlab_amd.world import SoftBody

DT = np.float32(0.0005)


@pytest.mark.parametrize("name", sorted(TERRAIN_CASES))
def test_terrain_matches_reference(name, ref_mod):
    settings, boxes, bodies, n = TERRAIN_CASES[name]()
    st = terrain_statics()
    cps = [1, 7, 500, n]
    orc = ref_mod.OracleWorld(settings, boxes)
    so = run_world(orc, bodies, n, cps, st)
    sr = run_world(ref_mod.RefWorld(settings, boxes, threads=1), bodies, n, cps, st)
    assert_same(so, sr, f"oracle vs reference [{name}]")
    assert orc.stats()["static_contacts"] > 1000
    if name == "terrain":  # (the body over the hole fell through it)
        assert so[-1][4][0]["p"][:, 1].max() < -1.0


def lone_node(p, v=(0.0, 0.0, 0.0), r=0.05, m=1.0):
    b = SoftBody("node")
    b.add_node(p, m)
    b.finalize()
    b.nodes["v"][0] = v
    b.info["radius"] = np.float32(r)
    return b


def plane(slope, surfaces=None, n=21, cell=0.5, org=(-5.0, -5.0)):
    x = org[0] + cell * np.arange(n)
    h = np.tile((slope * x + 1.0).astype(np.float32), (n, 1))
    return h, cell, org, surfaces


def cpu_world(oracle_mod, terrain=None, cyl=None, settings=None):
    w = oracle_mod.OracleWorld(settings or abi.default_settings(), None)
    if terrain is not None:
        w.set_terrain(*terrain)
    if cyl is not None:
        w.set_cylinders(cyl)
    return w


def test_node_rests_on_a_sloped_plane(oracle_mod):
    """A 1 in 5 slope of concrete (grip 1.2 > 0.2): a node let down on it is held at its radius from the plane, only
    creeping (the adhesion's damping fades with the slip speed: centimetres a second, where sliding would be metres)."""
    w = cpu_world(oracle_mod, plane(0.2))
    r = np.float32(0.05)
    w.add_body(lone_node((0.3, 0.2 * 0.3 + 1.0 + 0.06, 0.2), r=r))
    w.step_substeps(2000)
    p, v = w.read_body(0)[0]["p"][0], w.read_body(0)[0]["v"][0]
    ny = 1.0 / np.sqrt(1.0 + 0.2 ** 2)
    dist = (p[1] - (0.2 * p[0] + 1.0)) * ny
    assert abs(dist - r) < 3e-3, dist
    assert np.linalg.norm(v) < 0.05 and abs(p[0] - 0.3) < 0.1


def test_surface_is_the_nearest_vertex(oracle_mod):
    """Flat terrain of ice (x < 0) and concrete (x >= 0): a node sliding at 3 m/s on the ice goes about twice as far
    as one on the concrete."""
    n = 41
    surf = np.full((n, n), abi.SURF_CONCRETE, np.uint8)
    surf[:, :20] = abi.SURF_ICE  # (x = -10 + 0.5 i: vertices left of x = 0 are ice)
    dist = []
    for x0 in (-9.0, 1.0):
        w = cpu_world(oracle_mod, (np.zeros((n, n), np.float32), 0.5, (-10.0, -10.0), surf))
        w.add_body(lone_node((x0, 0.049, 0.0), v=(3.0, 0.0, 0.0)))
        w.step_substeps(6000)
        dist.append(w.read_body(0)[0]["p"][0, 0] - x0)
    assert dist[0] > 1.6 * dist[1] > 0, dist
    assert dist[0] < 9.0  # (stayed on the ice)


def test_cylinder_top_and_side(oracle_mod):
    cyl = np.zeros(1, abi.CYLINDER)
    cyl[0]["base"], cyl[0]["radius"], cyl[0]["height"], cyl[0]["surface"] = (0.0, 0.0, 0.0), 0.5, 1.0, abi.SURF_CONCRETE
    r = np.float32(0.05)
    # on its top: at rest its radius above it
    w = cpu_world(oracle_mod, cyl=cyl)
    w.add_body(lone_node((0.1, 1.08, -0.1), r=r))
    w.step_substeps(2000)
    p, v = w.read_body(0)[0]["p"][0], w.read_body(0)[0]["v"][0]
    assert abs(p[1] - (1.0 + r)) < 3e-3 and np.linalg.norm(v) < 1e-3
    # thrown at its side (no gravity): stopped where it comes within its radius of the surface, never through
    s = abi.default_settings()
    s.gravity[:] = (0.0, 0.0, 0.0)
    w = cpu_world(oracle_mod, cyl=cyl, settings=s)
    w.add_body(lone_node((-1.0, 0.5, 0.0), v=(2.0, 0.0, 0.0), r=r))
    xs = []
    for _ in range(100):
        w.step_substeps(10)
        xs.append(w.read_body(0)[0]["p"][0, 0])
    v = w.read_body(0)[0]["v"][0]
    assert -0.5 - r - 0.01 < max(xs) < -0.5, max(xs)
    assert v[0] <= 1e-3 and xs[-1] <= max(xs)


def test_hole_lets_a_node_fall_through(oracle_mod):
    n = 21
    h = np.zeros((n, n), np.float32)
    h[9:12, 9:12] = -1e30  # (no ground round the middle vertex (x, z) = (0, 0))
    s = abi.default_settings()
    s.kill_y = np.float32(-2.0)
    w = cpu_world(oracle_mod, (h, 0.5, (-5.0, -5.0), None), settings=s)
    w.add_body(lone_node((0.1, 0.5, 0.1)))
    w.step_substeps(600)  # 0.3 s: free fall, 0.44 m
    p, v = w.read_body(0)[0]["p"][0], w.read_body(0)[0]["v"][0]
    assert abs(v[1] + 9.81 * 0.3) < 0.01 and p[1] < 0.1
    w.step_substeps(1000)
    node = w.read_body(0)[0][0]
    assert node["inv_mass"] == 0 and node["p"][1] < -2.0 and not node["v"].any()


GPU_TUNINGS = {
    "auto": dict(mode=abi.MODE_AUTO),
    "stepped_graph": dict(mode=abi.MODE_STEPPED, block=256, use_graph=1, substeps_per_launch=16, incidences_per_thread=1),
    "persistent": dict(mode=abi.MODE_PERSISTENT, block=1024, substeps_per_launch=50),
    "resident_4": dict(mode=abi.MODE_RESIDENT, block=256, incidences_per_thread=4, substeps_per_launch=33),
    "halo_4": dict(mode=abi.MODE_HALO, block=256, incidences_per_thread=4, substeps_per_launch=33),
    "stepped_launches": dict(mode=abi.MODE_STEPPED, block=256, use_graph=0, incidences_per_thread=1),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name,tuning", [("terrain", t) for t in ("auto", "stepped_graph", "persistent", "resident_4",
                                                                    "halo_4")] +
                         [("terrain_car", t) for t in ("auto", "stepped_launches")])
def test_gpu_terrain_matches_oracle(name, tuning, oracle_mod):
    from beamlab_amd.world import World
    print(f"terrain case {name} tuning {tuning}", flush=True)
    settings, boxes, bodies, n = TERRAIN_CASES[name]()
    st = terrain_statics()
    cps = [1, 7, 500, n]
    ow = oracle_mod.OracleWorld(settings, boxes)
    so = run_world(ow, bodies, n, cps, st)
    gw = World(settings)
    gw.statics.boxes = boxes
    gw.set_tuning(**GPU_TUNINGS[tuning])
    sg = run_world(gw, bodies, n, cps, st)
    assert_same(sg, so, f"gpu[{tuning}] vs oracle [{name}]")
    assert gw.stats()["static_contacts"] == ow.stats()["static_contacts"] > 1000
    gw.close()
