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


# ---------------------------------------------------------------------------------------------- frames (members)
FJ_RIGID, FJ_BALL, FJ_HINGE_V, FJ_HINGE_H, FJ_SWIVEL, FJ_ELASTIC = range(6)


def frame_section(E, G, rho, A, Iy, Iz, J, As_y, As_z, damping=2e-5, axial=1.0, joint_k=2.0e4):
    """A member section's 12 floats (FrameSection's elastic fields, the oracle's and the reference shim's order)."""
    return np.array([E, G, rho, A, Iy, Iz, J, As_y, As_z, damping, axial, joint_k], np.float32)


def steel_tube(outer=0.04, wall=0.002, damping=2e-4, joint_k=2.0e4):
    """make_frame_section("Steel", Tube, outer, wall) (src/phys/frame_fem.cpp:95-135) without its plastic limits:
    A, I, J and the shear areas (0.53 A) in double, stored as floats."""
    D, t = outer, min(max(wall, 1e-4), outer * 0.5)
    d = D - 2 * t
    A = np.pi / 4 * (D * D - d * d)
    I = np.pi / 64 * (D ** 4 - d ** 4)
    return frame_section(2.1e11, 8.1e10, 7850.0, A, I, I, 2 * I, 0.53 * A, 0.53 * A, damping, 1.0, joint_k)


def frame_line(n, p0, p1, sec, root=FJ_RIGID):
    """The reference's frame_line (tests/test_physics.cpp:1383-1394): n members from p0 to p1, the first node
    clamped, lumped masses (a member's mass split over its ends), the root member's joint at its first node."""
    p0, p1 = np.asarray(p0, np.float32), np.asarray(p1, np.float32)
    L = np.float32(np.linalg.norm((p1 - p0).astype(np.float32)))
    m = np.float32(sec[2] * sec[3]) * L / np.float32(n)
    p = np.array([p0 + (p1 - p0) * (np.float32(i) / np.float32(n)) for i in range(n + 1)], np.float32)
    mass = np.array([m * np.float32(0.5) if i in (0, n) else m for i in range(n + 1)], np.float32)
    fixed = np.zeros(n + 1, np.uint8)
    fixed[0] = 1
    members = np.array([(i, i + 1, 0, (root if i == 0 else FJ_RIGID)) for i in range(n)], np.uint32)
    return dict(p=p, mass=mass, fixed=fixed, members=members, sections=sec[None, :])


def portal(sec, n=4, feet=FJ_RIGID):
    """The reference's portal frame (tests/test_physics.cpp:2109-2128): clamped feet at (0,0,0), (1,0,0), corners at
    height 1, three chains of n members (frame_chain), the feet's joints `feet`."""
    p = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)]
    mass = [0.0, 0.0, 0.0, 0.0]
    members = []
    fixed = [1, 1, 0, 0]
    for a, c in ((0, 2), (1, 3), (2, 3)):
        p0, p1 = np.asarray(p[a], np.float32), np.asarray(p[c], np.float32)
        m = np.float32(sec[2] * sec[3]) * np.float32(np.linalg.norm((p1 - p0).astype(np.float32))) / np.float32(n)
        mass[a] = np.float32(mass[a]) + m * np.float32(0.5)
        mass[c] = np.float32(mass[c]) + m * np.float32(0.5)
        prev = a
        for i in range(1, n + 1):
            cur = c
            if i < n:
                p.append(tuple(p0 + (p1 - p0) * (np.float32(i) / np.float32(n))))
                mass.append(m)
                fixed.append(0)
                cur = len(p) - 1
            ends = (feet if prev in (0, 1) else FJ_RIGID)
            members.append((prev, cur, 0, ends))
            prev = cur
    return dict(p=np.array(p, np.float32), mass=np.array(mass, np.float32), fixed=np.array(fixed, np.uint8),
                members=np.array(members, np.uint32), sections=sec[None, :])


def stiffened_plate(n=6, sec=None):
    """A clamped n x n plate (C3's section) with a steel tube along its middle line under the shell: members on the
    plate's own nodes (member and triangle blocks share node pairs)."""
    sec = steel_tube() if sec is None else sec
    from beamlab_amd import fem
    p, t, fx = fem.plate_mesh(n, n)
    mass = np.zeros(len(p), np.float32)
    mpa = np.float32(SECTION_C3[2]) * np.float32(SECTION_C3[3])
    for tri in t:
        a, b, c = p[tri[0]], p[tri[1]], p[tri[2]]
        area = np.float32(0.5) * np.float32(np.sqrt(np.float32(np.dot(np.cross(b - a, c - a), np.cross(b - a, c - a)))))
        for v in tri:
            mass[v] = mass[v] + (mpa * area) / np.float32(3.0)
    j = n // 2
    row = [j * (n + 1) + i for i in range(n + 1)]
    members = [(row[i], row[i + 1], 0, FJ_RIGID) for i in range(n)]
    for a, b, _, _ in members:
        m = np.float32(sec[2] * sec[3]) * np.float32(np.linalg.norm(p[b] - p[a]))
        mass[a] += m * np.float32(0.5)
        mass[b] += m * np.float32(0.5)
    return dict(p=p, mass=mass, fixed=fx, tris=t, members=np.array(members, np.uint32), sections=sec[None, :])


def make_frame(mod, fr, ref=False):
    cls = mod.RefPlate if ref else mod.FemPlate
    return cls.frame(fr["p"], fr["mass"], fr.get("fixed"), fr.get("tris"), SECTION_C3, fr.get("members"),
                     fr.get("sections"))


def frames():
    sec = steel_tube()
    return {
        "cantilever": frame_line(10, (0, 1, 0), (1, 1, 0), sec),
        "two_members": frame_line(2, (0, 1, 0), (1, 1, 0), sec),
        "portal": portal(sec),
        "portal_ball_feet": portal(sec, feet=FJ_BALL),
        "hinge_root": frame_line(10, (0, 1, 0), (1, 1, 0), sec, FJ_HINGE_V),
        "swivel_root": frame_line(10, (0, 1, 0), (1, 1, 0), sec, FJ_SWIVEL),
        "elastic_root": frame_line(10, (0, 1, 0), (1, 1, 0), steel_tube(joint_k=5.0e4), FJ_ELASTIC),
        "stiffened_plate": stiffened_plate(6),
    }


def frame_loads(name, nn, steps, seed):
    """Random loads and moments on every node; the two-member cantilever also gets a heavy tip load, bending its
    members far enough that their co-rotated frames take the slerp's arc (end frames more than 3.6 degrees apart)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(steps):
        ld = rng.normal(0, 20, (nn, 3)).astype(np.float32)
        tq = rng.normal(0, 2, (nn, 3)).astype(np.float32)
        if name == "two_members":
            ld[-1, 1] -= 3000.0
        out.append((ld, tq))
    return out


# ---------------------------------------------------------------------------------------------- plasticity
STEEL_YIELD = 3.5e8  # frame_material("Steel").yield (src/phys/frame_fem.cpp:35)


def plastic_tube(outer=0.04, wall=0.002, fy=STEEL_YIELD, hardening=0.25, damping=2e-4):
    """make_frame_section("Steel", Tube, outer, wall) (src/phys/frame_fem.cpp:95-135) with its plastic limits, in
    double and stored as floats: Np = fy A, Mp = fy Zp with Zp = (D^3 - d^3) / 6, Tp = 2 pi / 3 tau (R^3 - r^3) with
    tau = fy / sqrt(3); the hinges' hardening and half the outer size. 17 floats."""
    base = steel_tube(outer, wall, damping)
    D = float(outer)
    t = min(max(float(wall), 1e-4), D * 0.5)
    d, R = D - 2 * t, D * 0.5
    r = d * 0.5
    fy = float(np.float32(fy))
    A = np.pi / 4 * (D * D - d * d)
    Zp = (D * D * D - d * d * d) / 6
    Tp = 2 * np.pi / 3 * (fy / np.sqrt(3.0)) * (R * R * R - r * r * r)
    return np.concatenate([base, np.array([fy * A, fy * Zp, Tp, hardening, D * 0.5], np.float32)]).astype(np.float32)


def plastic_shell(fy=1.0e8):
    """C3's section (aluminium, 1.5 mm) with a yield stress: 7 floats."""
    return tuple(SECTION_C3) + (float(fy),)


def plastic_plate(n=6, fy=1.0e8):
    """A clamped n x n plate of plastic_shell(fy) as a frame body (the plate mesher's nodes, triangles and masses)."""
    from beamlab_amd import fem
    p, t, fx = fem.plate_mesh(n, n)
    mass = np.zeros(len(p), np.float32)
    mpa = np.float32(SECTION_C3[2]) * np.float32(SECTION_C3[3])
    for tri in t:
        a, b, c = p[tri[0]], p[tri[1]], p[tri[2]]
        area = np.float32(0.5) * np.float32(np.sqrt(np.float32(np.dot(np.cross(b - a, c - a), np.cross(b - a, c - a)))))
        for v in tri:
            mass[v] = mass[v] + (mpa * area) / np.float32(3.0)
    return dict(p=p, mass=mass, fixed=fx, tris=t, shell=plastic_shell(fy))


def plastic_frames():
    """Frame bodies whose members and triangles yield under plastic_loads."""
    sec = plastic_tube()
    stiff = stiffened_plate(6, plastic_tube(outer=0.02, wall=0.001))
    stiff["shell"] = plastic_shell()
    return {
        "cantilever": frame_line(10, (0, 1, 0), (1, 1, 0), sec),
        "bar": frame_line(6, (0, 1, 0), (0.6, 1, 0), sec),
        "portal": portal(sec),
        "long": frame_line(70, (0, 1, 0), (1.4, 1, 0), plastic_tube(outer=0.03, wall=0.002)),
        "plate": plastic_plate(6),
        "stiffened_plate": stiff,
    }


def make_plastic(mod, fr, ref=False):
    cls = mod.RefPlate if ref else mod.FemPlate
    return cls.frame(fr["p"], fr["mass"], fr.get("fixed"), fr.get("tris"), fr.get("shell", SECTION_C3),
                     fr.get("members"), fr.get("sections"))


def plastic_loads(name, nn, steps, seed):
    """Loads past the sections' limits, then reversed (the hinges unload: passes again): on the cantilever a tip load
    of about 1.5 Mp / L down then up and a tip torque past Tp; on the bar a pull past Np then a push; on the portal a
    sway load past its collapse load (4 Mp / H); on the long chain a tip load alternating every other step (more
    passes again than the budget holds); on the plates a load at the middle node past their collapse. Random loads and
    moments of 20 N and 2 N m on every node."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(steps):
        ld = rng.normal(0, 20, (nn, 3)).astype(np.float32)
        tq = rng.normal(0, 2, (nn, 3)).astype(np.float32)
        half = s < steps // 2
        if name == "cantilever":
            ld[-1, 1] += -1500.0 if half else 1500.0
            if steps // 6 <= s < steps // 3:
                tq[-1, 0] += 1300.0
        elif name == "bar":
            ld[-1, 0] += 1.0e5 if half else -1.2e5
        elif name == "portal":
            ld[2, 0] += 6000.0 if half else -6000.0
        elif name == "long":
            ld[-1, 1] += -1500.0 if s < steps // 3 else (3000.0 if (s // 2) % 2 == 0 else -3000.0)
        elif name == "plate":
            ld[nn // 2, 1] += -2500.0 if half else 1500.0
        elif name == "stiffened_plate":
            ld[nn // 2, 1] += -6000.0 if half else 4000.0
        out.append((ld, tq))
    return out


# ---------------------------------------------------------------------------------------------- held steps
def held_cases():
    """(name, maker(mod, ref), loads(nn, steps)) of the held-step runs: frames stepping their frame every few substeps
    (WorldSettings::frame_every), elastic and plastic members, shells, the update-path plate."""
    out = []
    for name in ("cantilever", "portal", "stiffened_plate", "two_members"):
        fr = frames()[name]
        out.append((name, lambda mod, ref=False, fr=fr: make_frame(mod, fr, ref),
                    lambda nn, steps, name=name: frame_loads(name, nn, steps, 3)))
    for name in ("cantilever", "long", "plate"):
        fr = plastic_frames()[name]
        out.append(("plastic_" + name, lambda mod, ref=False, fr=fr: make_plastic(mod, fr, ref),
                    lambda nn, steps, name=name: plastic_loads(name, nn, steps, 5)))
    sc = events()
    out.append(("events", lambda mod, ref=False, sc=sc: make(mod, sc, ref),
                lambda nn, steps, sc=sc: [(ld, np.zeros((nn, 3), np.float32)) for ld in sc["loads"][:steps]]))
    return out


def strip_mesh(L, W, nu, nv, y=1.0):
    """The reference's tri_plate (tests/test_physics.cpp:1499-1512): ShellMesher::grid from (0, y, 0) by (L / nu, 0, 0)
    and (0, 0, W / nv), the nodes at x = 0 clamped: (positions, triangles, fixed)."""
    origin = np.array([0, y, 0], np.float32)
    du = np.array([np.float32(L) / np.float32(nu), 0, 0], np.float32)
    dv = np.array([0, 0, np.float32(W) / np.float32(nv)], np.float32)
    p = np.array([(origin + du * np.float32(i)) + dv * np.float32(j) for j in range(nv + 1) for i in range(nu + 1)],
                 np.float32)
    tris = []
    for j in range(nv):
        for i in range(nu):
            n00 = j * (nu + 1) + i
            n10, n01 = n00 + 1, n00 + nu + 1
            n11 = n01 + 1
            tris += [(n00, n10, n11), (n00, n11, n01)] if (i + j) & 1 else [(n00, n10, n01), (n10, n11, n01)]
    return p, np.asarray(tris, np.uint32), (p[:, 0] < 1e-4).astype(np.uint8)


def plastic_strip(L=0.5, W=0.2, t=0.01):
    """The reference's plastic strip (tests/test_physics.cpp:1645-1686): 0.5 m x 0.2 m of 10 mm steel
    (make_shell_section("Steel", 0.01): yield 350 MPa), 10 x 4 cells, clamped at its root; the masses as ShellMesher
    puts them. Returns (frame, section, tip nodes, Pe, Pp): the tip's first-yield and collapse loads."""
    sec = tuple(steel_section(t)) + (STEEL_YIELD,)
    p, tr, fx = strip_mesh(L, W, 10, 4)
    mass = np.zeros(len(p), np.float32)
    mpa = np.float32(sec[2]) * np.float32(sec[3])
    for c in tr:
        a, b, d = p[c[0]], p[c[1]], p[c[2]]
        n = np.cross(b - a, d - a).astype(np.float32)
        area = np.float32(0.5) * np.float32(np.sqrt(np.float32(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])))
        for v in c:
            mass[v] = mass[v] + (mpa * area) / np.float32(3.0)
    tip = np.nonzero(p[:, 0] > np.float32(L) - np.float32(1e-4))[0]
    fy, tt = float(np.float32(STEEL_YIELD)), float(np.float32(t))
    Pe, Pp = fy * tt * tt * W / 6.0 / L, fy * tt * tt * W / 4.0 / L
    return dict(p=p, mass=mass, fixed=fx, tris=tr, shell=sec), tip, Pe, Pp


def run_strip(P, tip, force, steps, damp):
    """steps substeps of the strip (no gravity, theta 1, dissipation 1) with force/len(tip) down at each tip node and,
    with damp, the test's drag -40 m v on every node (its pre_substep, from the velocities before the substep)."""
    for _ in range(steps):
        st = P.read()
        ld = np.zeros((P.nn, 3), np.float32)
        ld[tip, 1] += np.float32(-force) / np.float32(len(tip))
        if damp:
            ld -= st["v"] * (st["mass"] * np.float32(40.0))[:, None]
        P.step(ld, H, theta=1.0, dissipation=1.0, g=(0.0, 0.0, 0.0))
    return P.read()
