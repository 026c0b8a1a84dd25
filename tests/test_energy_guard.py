"""The energy guard (SoftBody::energy_guard: motion_energy at a frame's start, guard_energy at its end;
src/phys/softbody.cpp:798-815, src/phys/world.cpp:3236,3995-3996): a body quiet for more than 15 frames that gains
more than 2 J and 5 % of its kinetic energy in a frame has its velocities scaled back to exactly that gain.

A free wheel (fixed hubs, no gravity) driven by a controller's torque gains energy every frame: the first 15 frames
freely, then each frame exactly 2 J + 5 % of its kinetic energy at the frame's start. CPU: the oracle against the
compiled reference, bit for bit; the closed form of the capped gain. GPU: the HIP path against the oracle.
"""
import numpy as np
import pytest

from beamlab_amd import abi
from test_wheels import free_wheel

FRAMES, SUB = 30, 20


def guarded_wheel(guard=1):
    b, ring = free_wheel(12, 0.3, 2.0)
    b.opts.energy_guard = guard
    return b


def drive(world, frames=FRAMES):
    """Frames of SUB substeps under a 400 N m drive torque; the node records after each."""
    world.add_body(guarded_wheel())
    world.set_wheel_controls(0, [400.0], [0.0])
    out = []
    for _ in range(frames):
        world.step_substeps(SUB)
        out.append(world.read_body(0)[0])
    return out


def kinetic(nodes):
    m = nodes["inv_mass"] > 0
    v = nodes["v"][m]
    return float(np.sum(0.5 * nodes["mass"][m].astype(np.float64) * np.sum(v.astype(np.float64) ** 2, axis=1)))


def no_gravity():
    s = abi.default_settings()
    s.gravity[:] = (0.0, 0.0, 0.0)
    return s


def test_guard_matches_reference(ref_mod):
    so = drive(ref_mod.OracleWorld(no_gravity(), None))
    sr = drive(ref_mod.RefWorld(no_gravity(), None, threads=1))
    for k, (a, b) in enumerate(zip(so, sr)):
        assert a.tobytes() == b.tobytes(), f"frame {k}"


def test_guard_caps_the_gain(oracle_mod):
    """Frames 1-15 gain what the torque gives (the same as an unguarded wheel); from frame 16 on each frame's gain is
    2 J + 5 % of the kinetic energy at its start, to float rounding."""
    so = drive(oracle_mod.OracleWorld(no_gravity(), None))
    w = oracle_mod.OracleWorld(no_gravity(), None)
    w.add_body(guarded_wheel(0))
    w.set_wheel_controls(0, [400.0], [0.0])
    free = []
    for _ in range(FRAMES):
        w.step_substeps(SUB)
        free.append(kinetic(w.read_body(0)[0]))
    ke = [kinetic(n) for n in so]
    assert ke[:15] == free[:15]
    cut = 0
    for k in range(15, FRAMES):
        gain = ke[k] - ke[k - 1]
        if abs(gain - (2.0 + 0.05 * ke[k - 1])) < 1e-4 * (2.0 + 0.05 * ke[k - 1]):
            cut += 1
        else:
            assert gain < 2.0 + 0.05 * ke[k - 1]
    assert cut >= 5 and ke[-1] < free[-1]


@pytest.mark.gpu
def test_gpu_guard_matches_oracle(oracle_mod):
    from beamlab_amd.world import World
    print("energy guard", flush=True)
    so = drive(oracle_mod.OracleWorld(no_gravity(), None))
    gw = World(no_gravity())
    sg = drive(gw)
    for k, (a, b) in enumerate(zip(sg, so)):
        assert a.tobytes() == b.tobytes(), f"frame {k}"
    gw.close()
