"""Seeded scenes shared by the parity tests (CPU: oracle vs compiled reference; GPU: HIP path vs oracle).

Each case returns (settings, boxes, [bodies], substeps). They cover the branches of the beam path: the plain
spring-damper, compressive and tensile yield (work hardening, weakening, the ductile switch), breaking in tension
and compression, rope and support beams, fixed nodes, the fall-off (kill_y) freeze, the 400 m/s clamp and the NaN
reset, oriented boxes, every ground model (the mud and sand fluid layers included), a yielding obstacle
(max_force), a bouncing body, several bodies of uneven size.
"""
import numpy as np

from beamlab_amd import abi, scenes
from beamlab_amd.world import make_box


def _boxes(*specs):
    out = np.zeros(len(specs), abi.BOX)
    for i, s in enumerate(specs):
        out[i] = make_box(*s)
    return out


GROUND = ((0.0, -0.5, 0.0), (100.0, 0.5, 100.0), (0.0, 0.0, 0.0, 1.0), abi.SURF_CONCRETE)


def case_c1():
    return abi.default_settings(), _boxes(GROUND), [scenes.config1_body(seed=1)], 2000


def case_plastic():
    b = scenes.lattice_body(10, 5, 4, 0.25, 2.0, (0.0, 0.6, 0.0), seed=5, eps=0.08, deform_factor=0.25)
    return abi.default_settings(), _boxes(GROUND), [b], 1500


def case_ductile_nodeform():
    b1 = scenes.lattice_body(6, 4, 3, 0.25, 2.0, (-2.0, 0.5, 0.0), seed=6, eps=0.08, deform_factor=0.25)
    b1.opts.ductile = 1
    b2 = scenes.lattice_body(6, 4, 3, 0.25, 2.0, (2.0, 0.5, 0.0), seed=7, eps=0.08, deform_factor=0.25)
    b2.opts.allow_deform = 0
    b3 = scenes.lattice_body(6, 4, 3, 0.25, 2.0, (0.0, 0.5, 3.0), seed=8, eps=0.03, deform_factor=0.25)
    b3.opts.allow_break = 0
    b3.beams["flags"][::5] |= abi.BF_NO_DEFORM
    b3.beams["flags"][1::7] |= abi.BF_NO_BREAK
    return abi.default_settings(), _boxes(GROUND), [b1, b2, b3], 1200


def case_rope_support():
    b = scenes.lattice_body(8, 4, 3, 0.25, 2.0, (0.0, 2.0, 0.0), seed=9, eps=0.2)
    rng = np.random.default_rng(9)
    t = rng.integers(0, 3, len(b.beams)).astype(np.uint8)
    b.beams["type"] = t
    b.beams["support_limit"] = np.float32(0.02)
    # a few nodes thrown apart so that ropes slacken and supports stretch past their limit
    b.nodes["v"][::5, 0] = np.float32(3.0)
    b.nodes["v"][2::5, 1] = np.float32(-4.0)
    b.beams["flags"][::11] |= abi.BF_SHOCK  # left to the shock loop: skipped
    b.beams["flags"][3::13] |= abi.BF_BROKEN
    return abi.default_settings(), _boxes(GROUND), [b], 1500


def case_fixed_kill():
    b = scenes.lattice_body(6, 6, 3, 0.25, 2.0, (0.0, 3.0, 0.0), seed=10)
    top = b.nodes["p"][:, 1] > np.float32(3.0 + 1.2)
    b.nodes["inv_mass"][top] = 0  # NF_FIXED anchors (inv_mass 0)
    b.info["flags"][top] |= abi.NF_FIXED
    b.info["flags"][::4] &= ~np.uint16(abi.NF_GROUND)
    b2 = scenes.lattice_body(4, 3, 3, 0.25, 2.0, (5.0, 1.0, 0.0), seed=11)  # falls past kill_y
    s = abi.default_settings()
    s.kill_y = np.float32(-0.6)
    boxes = _boxes(((0.0, -0.5, 0.0), (2.0, 0.5, 2.0), (0.0, 0.0, 0.0, 1.0), abi.SURF_CONCRETE))
    return s, boxes, [b, b2], 1500


def case_surfaces():
    bodies = []
    specs = []
    for k in range(abi.SURF_ICE + 1):
        x = (k - 5) * 3.0
        specs.append(((x, -0.5, 0.0), (1.4, 0.5, 1.4), (0.0, 0.0, 0.0, 1.0), k))
        b = scenes.lattice_body(4, 3, 3, 0.25, 2.0, (x, 0.3, 0.0), seed=20 + k)
        b.nodes["v"][:, 0] = np.float32(2.5)  # sliding: the Stribeck branch, then sticking
        b.info["radius"] = np.float32(0.02 * (k % 3))
        b.info["friction"] = np.float32(0.5 + 0.1 * k)
        b.opts.ground_friction = np.float32(1.0 - 0.05 * k)
        bodies.append(b)
    return abi.default_settings(), _boxes(*specs), bodies, 1500


def case_oriented_boxes():
    q1 = (0.0, 0.0, np.sin(0.2), np.cos(0.2))
    q2 = (np.sin(0.15), 0.0, 0.0, np.cos(0.15))
    q3 = (0.1, 0.3, -0.2, 0.927361849549570)
    boxes = _boxes(((0.0, -0.5, 0.0), (3.0, 0.5, 3.0), q1, abi.SURF_ASPHALT),
                   ((0.0, 0.1, 0.0), (0.4, 0.2, 0.4), q2, abi.SURF_METAL),
                   ((0.8, 0.0, 0.5), (0.3, 0.3, 0.3), q3, abi.SURF_ROCK))
    boxes[2]["max_force"] = np.float32(300.0)  # a yielding obstacle (a bush): the contact force is capped
    b = scenes.lattice_body(6, 3, 4, 0.25, 2.0, (0.2, 0.9, 0.1), seed=30, eps=0.1)
    b.opts.bounce = np.float32(0.5)
    b.opts.contact_slop = np.float32(0.002)
    b.opts.contact_push_max = np.float32(1.5)
    b.info["radius"] = np.float32(0.03)
    return abi.default_settings(), boxes, [b], 1500


def case_runaway():
    b = scenes.lattice_body(5, 3, 3, 0.25, 2.0, (0.0, 0.5, 0.0), seed=40, eps=10.0)
    b.nodes["v"][3] = (500.0, 0.0, 0.0)       # beyond 400 m/s: clamped
    b.nodes["v"][7] = (np.nan, 0.0, 0.0)      # NaN: reset to rest
    b.nodes["v"][9] = (0.0, -np.inf, 0.0)
    return abi.default_settings(), _boxes(GROUND), [b], 50


def case_multi():
    bodies = [scenes.lattice_body(nx, ny, nz, 0.25, 1.5, ((i - 2) * 4.0, 0.4 + 0.2 * i, 0.0), seed=50 + i)
              for i, (nx, ny, nz) in enumerate([(3, 2, 2), (12, 6, 5), (5, 5, 5), (2, 2, 2), (16, 4, 3)])]
    return abi.default_settings(), _boxes(GROUND), bodies, 1000


CPU_CASES = {
    "c1": case_c1, "plastic": case_plastic, "ductile_nodeform": case_ductile_nodeform,
    "rope_support": case_rope_support, "fixed_kill": case_fixed_kill, "surfaces": case_surfaces,
    "oriented_boxes": case_oriented_boxes, "runaway": case_runaway, "multi": case_multi,
}


def run_world(world, bodies, substeps, checkpoints):
    """Steps `world` and returns the states at each checkpoint (a substep count)."""
    for b in bodies:
        world.add_body(b)
    states = []
    done = 0
    for cp in checkpoints:
        world.step_substeps(cp - done)
        done = cp
        states.append([world.read_body(i) for i in range(len(bodies))] +
                      [[world.broken_beams(i) for i in range(len(bodies))]])
    return states


def canon(rec):
    """A copy with every NaN of the float fields made the same NaN: a NaN's sign and payload are not results (x86
    makes a negative NaN where the GPU makes a positive one)."""
    out = rec.copy()
    for name in rec.dtype.names:
        f = out[name]
        if f.dtype.kind == "f":
            f[np.isnan(f)] = np.nan
    return out


def assert_same(sa, sb, what):
    sa = [[(canon(n), canon(b)) for n, b in x[:-1]] + [x[-1]] for x in sa]
    sb = [[(canon(n), canon(b)) for n, b in x[:-1]] + [x[-1]] for x in sb]
    for k, (xa, xb) in enumerate(zip(sa, sb)):
        assert xa[-1] == xb[-1], f"{what}: broken-beam counts differ at checkpoint {k}: {xa[-1]} vs {xb[-1]}"
        for i, ((na, ba), (nb, bb)) in enumerate(zip(xa[:-1], xb[:-1])):
            if na.tobytes() != nb.tobytes():
                diff = np.nonzero(na.view(np.uint8).reshape(len(na), -1) != nb.view(np.uint8).reshape(len(nb), -1))[0]
                raise AssertionError(f"{what}: body {i} nodes differ at checkpoint {k}: first node {diff[0]}: "
                                     f"{na[diff[0]]} vs {nb[diff[0]]}")
            if ba.tobytes() != bb.tobytes():
                diff = np.nonzero(ba.view(np.uint8).reshape(len(ba), -1) != bb.view(np.uint8).reshape(len(bb), -1))[0]
                raise AssertionError(f"{what}: body {i} beams differ at checkpoint {k}: first beam {diff[0]}: "
                                     f"{ba[diff[0]]} vs {bb[diff[0]]}")
