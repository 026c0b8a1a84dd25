"""Seeded scenarios of the implicit shell step shared by the oracle's tests, the golden fixtures
(tests/golden/make_golden.py) and the GPU's tests: a plate, its section, and the loads of each step.

Every scenario is a step list of loads (N, 3 per node, float32) for one substep of H, plus the plate's build
arguments. `apply(plate, scenario)` steps an oracle FemPlate or RefPlate through it and yields the state after each
step."""
import numpy as np

H = 0.0005
SECTION_C3 = (7.0e10, 0.33, 2700.0, 1.5e-3, 0.05, 3e-4)


def steel_section(t):
    """make_shell_section("Steel", t) (src/phys/frame_fem.cpp:35,140-146): E 210 GPa, G 81 GPa, rho 7850, in float."""
    E, G = np.float32(2.1e11), np.float32(8.1e10)
    nu = float(min(max(E / (np.float32(2.0) * G) - np.float32(1.0), np.float32(0.0)), np.float32(0.49)))
    return (2.1e11, nu, 7850.0, float(t), 0.05, 3e-4)


def loaded(n, steps, amp, seed):
    """A clamped n x n plate (1 m, C3's section) under gravity and random loads N(0, amp) on every node."""
    rng = np.random.default_rng(seed)
    nn = (n + 1) * (n + 1)
    loads = [rng.normal(0, amp, (nn, 3)).astype(np.float32) for _ in range(steps)]
    return dict(n=n, size=1.0, y0=0.0, section=SECTION_C3, clamped=True, g=(0.0, -9.81, 0.0), loads=loads)


def events(n=8):
    """The update's safety paths on a clamped 8 x 8 plate: a kick beyond 40 m/s (clamped, counted), a sideways kick,
    a non-finite load (every free node's solution not finite: counted, the forces applied explicitly, the NaN
    velocity reset; the held impulse carries the NaN into the next step's solve too)."""
    sc = loaded(n, 12, 10.0, 2)
    sc["loads"][3][40, 1] = 5e5
    sc["loads"][5][30, :] = (3e4, 0.0, 2e4)
    sc["loads"][7][22, 0] = np.nan
    return sc


def spinning(steps=2000):
    """The reference's spun free plate (tests/test_physics.cpp:1615-1640): 1 m, 6 x 6 cells of 2 mm steel, no
    gravity, every node turning at omega = (1.5, 5, 1) rad/s about the plate's centre."""
    return dict(n=6, size=1.0, y0=1.0, section=steel_section(0.002), clamped=False, g=(0.0, 0.0, 0.0),
                loads=[None] * steps, omega=(1.5, 5.0, 1.0), centre=(0.0, 1.0, 0.0))


def spin_state(p, nn, sc):
    """The initial v = omega x (p - c) and w = omega of the spun plate."""
    om = np.asarray(sc["omega"], np.float32)
    v = np.cross(om, p - np.asarray(sc["centre"], np.float32)).astype(np.float32)
    return v, np.tile(om, (nn, 1)).astype(np.float32)


def make(mod, sc, ref=False):
    """The scenario's plate as an oracle FemPlate (or the compiled reference's RefPlate), its initial state set."""
    cls = mod.RefPlate if ref else mod.FemPlate
    P = cls(sc["n"], sc["n"], sc["size"], sc["y0"], sc["section"], clamped=sc["clamped"])
    if "omega" in sc:
        st = P.read()
        v, w = spin_state(st["p"], len(st["p"]), sc)
        if ref:
            P.write(v=v, w=w)
        else:
            xd, _, _, q = P.state()
            P.set_state(xd, v, w, q)
    return P


def apply(P, sc, ref=False):
    """Steps P through the scenario, yielding (step, state) after each step."""
    for s, ld in enumerate(sc["loads"]):
        if ref:
            P.substep(ld, H, g=sc["g"])
        else:
            P.step(ld, H, g=sc["g"])
        yield s, P.read()


def kinetic(st, inertia):
    """0.5 m v^2 + 0.5 I w^2 summed in double, as the reference's test sums it."""
    return float((0.5 * st["mass"].astype(np.float64) * (st["v"].astype(np.float64) ** 2).sum(1)).sum()
                 + (0.5 * inertia.astype(np.float64) * (st["w"].astype(np.float64) ** 2).sum(1)).sum())


STATE_KEYS = ("p", "v", "xd", "w", "q", "force")
