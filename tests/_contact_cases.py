"""Seeded contact scenes (inter-body and self contacts of slow bodies): lattices with their faces as triangles,
balls (one node and a capsule of no length each), collision groups, several islands."""
import numpy as np

from beamlab_amd import abi, scenes
from _cases import GROUND, _boxes


def case_balls_on_lattice():
    lat = scenes.lattice_body(6, 4, 3, 0.25, 2.0, (0.0, 0.3, 0.0), seed=3, faces=True, eps=0.2)
    lat.opts.self_collision = 1
    rng = np.random.default_rng(1)
    balls = [scenes.ball_body((rng.uniform(-0.6, 0.6), 1.0 + 0.12 * i, rng.uniform(-0.3, 0.3)), 0.12, 3.0)
             for i in range(6)]
    for b in balls:
        b.nodes["v"][0] = (0.0, -2.5, 0.0)
    return abi.default_settings(), _boxes(GROUND), [lat] + balls, 900


def case_lattice_pile():
    """Lattices dropped onto each other (node-triangle pairs both ways), a ball between them, self contact on."""
    bodies = []
    for i in range(3):
        b = scenes.lattice_body(4, 3, 3, 0.25, 1.0, (0.15 * i, 0.1 + 0.8 * i, 0.1 * i), seed=60 + i, faces=True,
                                eps=0.3, deform_factor=0.4)
        b.opts.self_collision = 1 if i != 1 else 0
        b.opts.contact_friction = np.float32(0.7 + 0.1 * i)
        bodies.append(b)
    bodies.append(scenes.ball_body((0.3, 2.8, 0.2), 0.1, 2.0))
    bodies[-1].nodes["v"][0] = (0.0, -2.0, 0.0)
    return abi.default_settings(), _boxes(GROUND), bodies, 900


def case_groups_islands():
    """Two far-apart groups (two islands), collision groups that exclude a pair, capsules of length."""
    a = scenes.lattice_body(4, 2, 3, 0.25, 1.0, (-5.0, 0.2, 0.0), seed=70, faces=True, eps=0.3)
    b = scenes.lattice_body(3, 2, 3, 0.25, 1.0, (-5.0, 0.9, 0.0), seed=71, faces=True, eps=0.3)
    a.opts.collision_group = 3
    b.opts.collision_group = 3  # same group: no contacts between a and b
    c = scenes.lattice_body(4, 2, 3, 0.25, 1.0, (5.0, 0.2, 0.0), seed=72, faces=True, eps=0.3)
    d = scenes.lattice_body(3, 2, 3, 0.25, 1.0, (5.1, 0.9, 0.05), seed=73, faces=True, eps=0.3)
    stick = scenes.ball_body((5.0, 1.8, 0.0), 0.05, 1.0)
    # a capsule of length: a two-node stick
    nodes = np.zeros(2, abi.NODE)
    nodes["p"] = [(4.8, 1.9, 0.0), (5.3, 2.0, 0.1)]
    nodes["mass"] = np.float32(1.0)
    nodes["inv_mass"] = np.float32(1.0)
    info = np.zeros(2, abi.NODE_INFO)
    info["flags"] = abi.NF_GROUND | abi.NF_CONTACTER
    info["friction"] = 1.0
    info["frame"] = -1
    info["ror_id"] = -1
    beams = np.zeros(1, abi.BEAM)
    beams["a"], beams["b"], beams["k"], beams["d"] = 0, 1, 1e5, 50
    beams["strength"], beams["max_pos"], beams["max_neg"], beams["support_limit"] = 1e9, 1e9, -1e9, 4
    stick.nodes, stick.info, stick.beams = nodes, info, beams
    stick.finalize()
    caps = np.zeros(1, abi.CAPSULE)
    caps["a"], caps["b"], caps["radius"], caps["joint"] = 0, 1, np.float32(0.06), -1
    stick.capsules = caps
    return abi.default_settings(), _boxes(GROUND), [a, b, c, d, stick], 800


def case_c4_small():
    """The batched scene's bodies (config4_body: random beam networks, mean degree 7) at a small size: 16 on the
    ground on a 2 m grid, 8 more dropped onto some of them from 1.5 m above."""
    rng = np.random.default_rng(44)
    sizes = np.clip(np.round(np.exp(np.log(120.0) + 0.5 * rng.standard_normal(24))), 40, 400).astype(int)
    bodies = []
    for i in range(16):
        bodies.append(scenes.config4_body(int(sizes[i]), (-3.0 + 2.0 * (i % 4), 0.05, -3.0 + 2.0 * (i // 4)), 500 + i))
    for j in range(8):
        i = int(rng.integers(0, 16))
        x0, z0 = -3.0 + 2.0 * (i % 4), -3.0 + 2.0 * (i // 4)
        bodies.append(scenes.config4_body(int(sizes[16 + j]), (x0 + rng.uniform(-0.3, 0.3), 1.5 + 1.2 * (j // 4),
                                                               z0 + rng.uniform(-0.3, 0.3)), 600 + j))
    return abi.default_settings(), _boxes(GROUND), bodies, 700


def case_crate_stacks():
    """The batched scene's layout at a reduced size: 16 stacks of two crates (700-1100 nodes each) resting on each
    other over a 40 x 40 m ground, about 10^4 contact hits a substep from the first, in 16 pair islands."""
    rng = np.random.default_rng(45)
    sizes = rng.integers(700, 1100, 32)
    return abi.default_settings(), _boxes(GROUND), scenes.stacked_crates(sizes, 40.0, 45), 150


CONTACT_CASES = {"balls_on_lattice": case_balls_on_lattice, "lattice_pile": case_lattice_pile,
                 "groups_islands": case_groups_islands, "c4_small": case_c4_small, "crate_stacks": case_crate_stacks}
