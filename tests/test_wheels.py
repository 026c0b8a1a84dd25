"""Wheels and tyre nodes (SoftBody::compute_wheel_forces, src/phys/softbody.cpp:689-767; World::collide_static's tyre
part and tyre_contact, src/phys/world.cpp:850-892, src/phys/static_world.cpp:53-80) with a vehicle controller
(the pre_substep hook setting each wheel's torque and brake, src/phys/world.cpp:3261).

CPU: the oracle against the compiled reference, bit for bit, the wheels' state and the body's wheel speed and spin
included, with the controller switched on, changed and off between steps; and what the physics fixes: a torque on a
free wheel gives its tread exactly that angular momentum (the opposite one mounted backward), the wheel then reads
the tread's speed; a braked car stops; a braked car stands on a 15 degree incline while
the tyres' grip holds and slides down it with a tenth of that grip. GPU: the HIP path against the oracle, bit for bit,
wheels included, with the controller and with inter-body contacts.
"""
import numpy as np
import pytest

from _cases import CPU_CASES, GROUND, _boxes, assert_same, car_body, run_world
from beamlab_amd import abi
from beamlab_amd.world import SoftBody

DT = np.float32(0.0005)


def controlled_run(world, bodies, schedule):
    """Steps `world` through `schedule`: (substeps, {body: (torque, brake) or None}); returns the states and the
    wheels after each leg."""
    for b in bodies:
        world.add_body(b)
    out = []
    for n, ctl in schedule:
        for bid, tb in ctl.items():
            world.set_wheel_controls(bid, *(tb if tb is not None else (None, None)))
        world.step_substeps(n)
        out.append(([world.read_body(i) for i in range(len(bodies))] + [[world.broken_beams(i) for i in range(len(bodies))]],
                    [world.read_wheels(i) for i in range(len(bodies))]))
    return out


SCHEDULE = [(1, {0: ([200, 200, 0, 0], [0, 0, 0, 0])}),
            (400, {0: ([250, 250, 0, -80], [0, 0, 0, 0]), 1: ([0, 0, 0, 0], [0, 0, 1500, 1500])}),
            (400, {0: ([0, 0, 0, 0], [2500, 2500, 2500, 2500])}),
            (300, {0: None, 1: None})]


def car_pair():
    b1 = car_body(91)
    b2 = car_body(92, origin=(0.0, 0.0, 4.0))
    b2.nodes["v"][:, 0] = 2.0
    return abi.default_settings(), _boxes(GROUND), [b1, b2]


def assert_same_wheels(wa, wb, what):
    for i, ((x, sx), (y, sy)) in enumerate(zip(wa, wb)):
        assert x.tobytes() == y.tobytes(), f"{what}: body {i} wheels differ: {x} vs {y}"
        assert sx.tobytes() == sy.tobytes(), f"{what}: body {i} wheel speed and spin differ: {sx} vs {sy}"


def test_wheels_match_reference(ref_mod):
    settings, boxes, bodies, n = CPU_CASES["wheels"]()
    orc, ref = ref_mod.OracleWorld(settings, boxes), ref_mod.RefWorld(settings, boxes, threads=1)
    so, sr = run_world(orc, bodies, n, [1, 500, n]), run_world(ref, bodies, n, [1, 500, n])
    assert_same(so, sr, "oracle vs reference [wheels]")
    assert_same_wheels([orc.read_wheels(i) for i in range(2)], [ref.read_wheels(i) for i in range(2)], "[wheels]")
    # the case does what it is for: the tyres touch, the braked car stopped from 3 m/s
    assert orc.stats()["static_contacts"] > 1000
    assert abs(so[-1][1][0]["v"][:8, 0].mean()) < 0.3


def test_controlled_wheels_match_reference(ref_mod):
    settings, boxes, bodies = car_pair()
    so = controlled_run(ref_mod.OracleWorld(settings, boxes), bodies, SCHEDULE)
    sr = controlled_run(ref_mod.RefWorld(settings, boxes, threads=1), bodies, SCHEDULE)
    for k, ((a, wa), (b, wb)) in enumerate(zip(so, sr)):
        assert_same([a], [b], f"oracle vs reference [controlled cars, leg {k}]")
        assert_same_wheels(wa, wb, f"[controlled cars, leg {k}]")
    # driven, the first car got going; braked, it stopped; the controller's torque is spent each substep
    v = [np.abs(s[0][0]["v"][:8, 0].mean()) for s, _ in so]
    assert v[1] > 0.3 and v[2] < 0.5 * v[1]
    assert (so[-1][1][0][0]["torque"] == 0).all()


def free_wheel(nseg=12, radius=0.3, m=2.0):
    """Two fixed hub nodes on the z axis and a free tread of 2 x nseg nodes round them (spokes to both hubs), no
    gravity, no ground."""
    b = SoftBody("wheel")
    a0 = b.add_node((0.0, 0.0, -0.1), 1.0, abi.NF_FIXED)
    a1 = b.add_node((0.0, 0.0, 0.1), 1.0, abi.NF_FIXED)
    ring = []
    for i in range(nseg):
        th = 2 * np.pi * i / nseg
        for z in (-0.1, 0.1):
            ring.append(b.add_node((radius * np.cos(th), radius * np.sin(th), z), m, 0))
    for n in ring:
        b.add_beam(n, a0, 1e5, 50, 1e9, 1e9), b.add_beam(n, a1, 1e5, 50, 1e9, 1e9)
    for i in range(0, len(ring), 2):
        b.add_beam(ring[i], ring[i + 1], 1e5, 50, 1e9, 1e9)
        b.add_beam(ring[i], ring[(i + 2) % len(ring)], 1e5, 50, 1e9, 1e9)
        b.add_beam(ring[i + 1], ring[(i + 3) % len(ring)], 1e5, 50, 1e9, 1e9)
    b.finalize()
    b.add_wheel(a0, a1, ring, radius=radius, propulsed=1)
    return b, ring


@pytest.mark.parametrize("propulsed", [1, 2])
def test_torque_gives_the_tread_its_angular_momentum(propulsed, oracle_mod):
    """One substep of a torque T on a free wheel: the tread's angular momentum about the axle (axle0 to axle1, +z) is
    T dt (each node gets T / (n r) along its tangent), -T dt for a wheel propelled backward (propulsed 2: mounted the
    other way round); the next substep the wheel reads the tread's speed along its own drive, T dt / (n r m)."""
    s = abi.default_settings()
    s.gravity[:] = (0.0, 0.0, 0.0)
    T, r, m, nseg = np.float32(40.0), 0.3, np.float32(2.0), 12
    b, ring = free_wheel(nseg, r, m)
    b.wheels["propulsed"][0] = propulsed
    w = oracle_mod.OracleWorld(s, None)
    w.add_body(b)
    w.set_wheel_controls(0, [T], [0.0])
    w.step_substeps(1)
    nodes, _ = w.read_body(0)
    p, v = nodes["p"][ring].astype(np.float64), nodes["v"][ring].astype(np.float64)
    L = np.sum(m * (p[:, 0] * v[:, 1] - p[:, 1] * v[:, 0]))
    want = T * DT if propulsed == 1 else -T * DT
    assert abs(L - want) < 1e-4 * T * DT, (L, want)
    w.set_wheel_controls(0, [0.0], [0.0])
    w.step_substeps(1)
    wh, ss = w.read_wheels(0)
    vt = T * DT / (2 * nseg * r * m)
    assert abs(wh[0]["speed"] - vt) < 2e-3 * vt, (wh[0]["speed"], vt)
    if propulsed == 1:  # (wheel_speed and wheel_spin average the forward-propelled wheels only)
        assert ss[0] == wh[0]["speed"] and abs(ss[1] - vt / r) < 2e-3 * vt / r
    else:
        assert ss[0] == 0 and ss[1] == 0
    assert wh[0]["last_torque"] == 0 and wh[0]["torque"] == 0


def chassis_v(world, bid=0):
    return world.read_body(bid)[0]["v"][:8].astype(np.float64).mean(axis=0)


def test_brakes_stop_a_rolling_car(oracle_mod):
    b = car_body(96)
    b.nodes["v"][:, 0] = 4.0
    w = oracle_mod.OracleWorld(abi.default_settings(), _boxes(GROUND))
    w.add_body(b)
    w.step_substeps(200)
    assert chassis_v(w)[0] > 3.0
    w.set_wheel_controls(0, [0.0] * 4, [2000.0] * 4)
    w.step_substeps(4000)
    assert np.linalg.norm(chassis_v(w)) < 0.05


@pytest.mark.parametrize("grip,held", [(1.0, True), (0.1, False)])
def test_braked_car_on_an_incline(grip, held, oracle_mod):
    """Gravity tilted 15 degrees (an incline of concrete, static grip 1.2): braked, the car stands while the tyre
    patches' grip holds (tan 15 = 0.27 < 1.2; it creeps a few mm a second as the brakes take the wheels' lagging
    speed) and slides down with a tenth of the grip (0.12 < 0.27)."""
    th = np.radians(15.0)
    s = abi.default_settings()
    s.gravity[:] = (9.81 * np.sin(th), -9.81 * np.cos(th), 0.0)
    s.tyre_grip = grip
    w = oracle_mod.OracleWorld(s, _boxes(GROUND))
    w.add_body(car_body(97))
    w.set_wheel_controls(0, [0.0] * 4, [5000.0] * 4)
    w.step_substeps(1000)
    x0 = w.read_body(0)[0]["p"][:8, 0].mean()
    w.step_substeps(2000)
    dx = w.read_body(0)[0]["p"][:8, 0].mean() - x0
    if held:
        assert abs(dx) < 0.03, dx
    else:
        assert dx > 0.3, dx


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", ["auto", "stepped_launches"])
def test_gpu_wheels_match_oracle(tuning, oracle_mod):
    from beamlab_amd.world import World
    print(f"wheels tuning {tuning}", flush=True)
    settings, boxes, bodies = car_pair()
    so_world = oracle_mod.OracleWorld(settings, boxes)
    so = controlled_run(so_world, bodies, SCHEDULE)
    gw = World(settings)
    gw.statics.boxes = boxes
    if tuning == "stepped_launches":
        gw.set_tuning(mode=abi.MODE_STEPPED, block=256, use_graph=0, incidences_per_thread=1)
    sg = controlled_run(gw, bodies, SCHEDULE)
    for k, ((a, wa), (b, wb)) in enumerate(zip(sg, so)):
        assert_same([a], [b], f"gpu[{tuning}] vs oracle [controlled cars, leg {k}]")
        assert_same_wheels(wa, wb, f"gpu[{tuning}] [controlled cars, leg {k}]")
    st_o, st_g = so_world.stats(), gw.stats()
    assert st_g["static_contacts"] == st_o["static_contacts"] > 0
    gw.close()


@pytest.mark.gpu
def test_gpu_wheels_with_contacts_match_oracle(oracle_mod):
    """A car driven into a lattice block (its box's faces as triangles, the block's nodes as contacters): tyres, wheels
    and inter-body contacts in one frame."""
    from beamlab_amd import scenes
    from beamlab_amd.world import World
    print("wheels with contacts", flush=True)
    car = car_body(98)
    for t in [(2, 3, 7), (2, 7, 6), (0, 4, 5), (0, 5, 1), (0, 1, 3), (0, 3, 2), (4, 6, 7), (4, 7, 5)]:
        car.add_triangle(*t)
    car._flush()
    block = scenes.lattice_body(3, 3, 3, 0.25, 2.0, (1.3, 0.05, -0.25), seed=99)
    settings, boxes = abi.default_settings(), _boxes(GROUND)
    sched = [(300, {0: ([400, 400, 0, 0], [0, 0, 0, 0])}), (600, {})]
    so = controlled_run(oracle_mod.OracleWorld(settings, boxes), [car, block], sched)
    gw = World(settings)
    gw.statics.boxes = boxes
    sg = controlled_run(gw, [car, block], sched)
    for k, ((a, wa), (b, wb)) in enumerate(zip(sg, so)):
        assert_same([a], [b], f"gpu vs oracle [car into a block, leg {k}]")
        assert_same_wheels(wa, wb, f"gpu [car into a block, leg {k}]")
    gw.close()


@pytest.mark.gpu
def test_gpu_ring_tyres_refused():
    from beamlab_amd.world import World
    w = World(abi.default_settings())
    b = car_body(70)
    b.wheels["ring"][0] = 1
    with pytest.raises(abi.BeamlabError) as e:
        w.add_body(b)
    assert e.value.status == abi.BL_ERR_UNSUPPORTED
    w.close()
