"""Seeded contact scenes (inter-body and self contacts): lattices with their faces as triangles (two-sided, or hull
triangles), balls (one node and a capsule of no length each), collision groups, several islands, fast bodies (above
8 m/s: their own pair search)."""
import numpy as np

from beamlab_amd import abi, scenes
from beamlab_amd.world import SoftBody
from _cases import GROUND, _boxes

DT = 0.0005


def ico_ball(c, r, mass, vel):
    """The reference test's make_ball (tests/test_physics.cpp:134-157): the 12 vertices of an icosahedron at radius
    r and a centre node, the 20 faces as triangles, the edges and the spokes as beams, stabilized, its velocity set,
    a 5 cm contact radius."""
    f = np.float32
    t = (f(1.0) + f(np.sqrt(f(5.0)))) * f(0.5)
    ico = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t), (t, 0, -1),
           (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
             (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    b = SoftBody("test ball")
    c = np.asarray(c, np.float32)
    for p in ico:
        p = np.asarray(p, np.float32)
        L = f(np.sqrt(f(f(p[0] * p[0]) + f(p[1] * p[1])) + f(p[2] * p[2])))
        b.add_node(c + (p * (f(1.0) / L)) * f(r), f(mass) / f(13.0))
    centre = b.add_node(c, f(mass) / f(13.0))
    edges = set()
    for fa in faces:
        for k in range(3):
            x, y = fa[k], fa[(k + 1) % 3]
            edges.add((min(x, y), max(x, y)))
        b.add_triangle(fa[0], fa[2], fa[1])
    for x, y in sorted(edges):
        b.add_beam(x, y, 2e6, 400, 1e12, 1e12)
    for i in range(12):
        b.add_beam(i, centre, 2e6, 400, 1e12, 1e12)
    b.finalize()
    b.stabilize(DT)
    b.nodes["v"][:] = np.asarray(vel, np.float32)
    b.opts.collision_radius = np.float32(0.05)
    return b


def hull_plate(o, hull):
    """The reference's hull plate (tests/test_physics.cpp:2762-2771): three fixed nodes round o, one triangle facing
    +y, one-sided (a hull triangle) or two-sided."""
    p = SoftBody("hull plate")
    o = np.asarray(o, np.float32)
    a = p.add_node(o + np.array((-2, 0, -2), np.float32), 50)
    b = p.add_node(o + np.array((2, 0, -2), np.float32), 50)
    c = p.add_node(o + np.array((0, 0, 3), np.float32), 50)
    p.add_beam(a, b, 1e6, 100, 1e12, 1e12), p.add_beam(b, c, 1e6, 100, 1e12, 1e12), p.add_beam(c, a, 1e6, 100, 1e12, 1e12)
    p.add_triangle(a, c, b, two_sided=not hull)
    p.finalize()
    p.nodes["inv_mass"] = 0
    return p


def case_hull_pushout(hull=True):
    """The reference's hull scene (tests/test_physics.cpp:2758-2802): a ball started 15 cm behind a hull triangle, no
    gravity, 60 frames of 33 substeps."""
    s = abi.default_settings()
    s.gravity[:] = (0.0, 0.0, 0.0)
    o = (0.0, 5.0, 0.0)
    return s, _boxes(), [hull_plate(o, hull), ico_ball((0.0, 5.0 - 0.15, 0.0), 0.08, 4.0, (0.0, 0.0, 0.0))], 1980


def case_hull_through():
    return case_hull_pushout(hull=False)


def case_hull_crates():
    """Lattices whose faces are hull triangles (one-sided, wound outwards) dropped onto each other: hull depths of
    their own, one touching with its faces only (a thin skin), one colliding with itself, a ball on top."""
    bodies = []
    for i in range(3):
        b = scenes.lattice_body(4, 3, 3, 0.25, 1.0, (0.1 * i, 0.1 + 0.7 * i, 0.05 * i), seed=80 + i, faces=True,
                                eps=0.3, deform_factor=0.4)
        b.tris["two_sided"] = 0
        b.opts.hull_depth = np.float32(0.12 + 0.05 * i)
        bodies.append(b)
    bodies[1].opts.faces_only = 1
    bodies[1].opts.face_skin = np.float32(0.003)
    bodies[2].opts.self_collision = 1
    bodies[0].tris["two_sided"][::5] = 1  # some of its faces two-sided
    ball = scenes.ball_body((0.2, 2.6, 0.1), 0.1, 2.0)
    ball.nodes["v"][0] = (0.0, -3.0, 0.0)
    bodies.append(ball)
    return abi.default_settings(), _boxes(GROUND), bodies, 900


def case_fast_bodies():
    """Fast bodies (over 8 m/s: their nodes out of the node hash, paired by their own grids with their partners'
    triangles and capsules): a lattice thrown at 14 m/s into a resting one, two lattices flying at each other (both
    fast), a ball at 25 m/s, a capsule stick in the way, hull faces on one of the fast lattices."""
    target = scenes.lattice_body(5, 3, 3, 0.25, 1.0, (0.0, 0.05, 0.0), seed=90, faces=True, eps=0.3)
    thrown = scenes.lattice_body(3, 3, 3, 0.25, 1.0, (-2.0, 0.3, 0.0), seed=91, faces=True, eps=0.3)
    thrown.nodes["v"][:] = (14.0, 1.0, 0.0)
    thrown.tris["two_sided"] = 0
    thrown.opts.hull_depth = np.float32(0.1)
    left = scenes.lattice_body(3, 2, 3, 0.25, 1.0, (-1.6, 0.2, 4.0), seed=92, faces=True, eps=0.3)
    right = scenes.lattice_body(3, 2, 3, 0.25, 1.0, (1.6, 0.25, 4.05), seed=93, faces=True, eps=0.3)
    left.nodes["v"][:] = (10.0, 0.5, 0.0)
    right.nodes["v"][:] = (-11.0, 0.5, 0.0)
    ball = scenes.ball_body((2.5, 0.6, 0.1), 0.1, 2.0)
    ball.nodes["v"][0] = (-25.0, 0.0, 0.0)
    stick = SoftBody("stick")
    a = stick.add_node((0.0, 1.2, -0.3), 1.0)
    b = stick.add_node((0.0, 1.25, 0.3), 1.0)
    stick.add_beam(a, b, 1e5, 50, 1e9, 1e9)
    stick.finalize()
    caps = np.zeros(1, abi.CAPSULE)
    caps["a"], caps["b"], caps["radius"], caps["joint"] = 0, 1, np.float32(0.06), -1
    stick.capsules = caps
    stick.nodes["inv_mass"] = 0  # held in place
    return abi.default_settings(), _boxes(GROUND), [target, thrown, left, right, ball, stick], 700


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
                 "groups_islands": case_groups_islands, "c4_small": case_c4_small, "crate_stacks": case_crate_stacks,
                 "hull_pushout": case_hull_pushout, "hull_through": case_hull_through, "hull_crates": case_hull_crates,
                 "fast_bodies": case_fast_bodies}


# cases stepped in frames of 33 substeps (the islands and the fast bodies' status are found per frame)
FRAMED = {"hull_pushout", "hull_through", "fast_bodies"}


def checkpoints(name, n):
    """The substep counts at which a case's states are compared (each a World::step_substeps call's end)."""
    cps = {1, 17, 40, n // 2, n}
    if name in FRAMED:
        cps |= set(range(33, n, 33))
    return sorted(cps)


def golden_checkpoints(name, n):
    """The checkpoints of a case's golden digest (tests/golden/make_golden.py)."""
    cps = {1, 7, n // 3, n}
    if name in FRAMED:
        cps |= set(range(33, n, 33))
    return sorted(cps)
