"""Synthetic scenes of the benchmark configurations, built through the SoftBody interface.

The body recipe is the reference's soft box (build_soft_box, src/world/objects.cpp:892-966): a node lattice with
beams to neighbours, each pair once in the lattice's loop order, then finalize and stabilize(dt)
(src/phys/softbody.cpp:76-110). The beam neighbourhood is the 18-neighbourhood (every neighbour closer than
sqrt(3) spacings on the unjittered lattice). Breaking follows the reference's force rule (|force| > strength,
src/phys/softbody.cpp:458): a break strain eps means strength = eps * k * L0 with k after stabilize. Seeds drive the
reference tests' xorshift32 generator (TRng, tests/test_physics.cpp:43-55).
"""
import math

import numpy as np

from . import abi
from .world import SoftBody, World

DT = np.float32(0.0005)


class TRng:
    """xorshift32 as the reference tests' TRng (tests/test_physics.cpp:43-55)."""

    def __init__(self, seed):
        self.s = (seed * 2654435761 + 1) & 0xFFFFFFFF

    def next(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        self.s = s
        return s

    def uni(self):
        return np.float32(self.next() & 0xFFFFFF) / np.float32(0x1000000)

    def range(self, a, b):
        a, b = np.float32(a), np.float32(b)
        return a + (b - a) * self.uni()

    def uniforms(self, n):
        """n draws of uni() as a float32 array."""
        out = np.empty(n, np.float32)
        s = self.s
        for i in range(n):
            s ^= (s << 13) & 0xFFFFFFFF
            s ^= s >> 17
            s ^= (s << 5) & 0xFFFFFFFF
            out[i] = s & 0xFFFFFF
        self.s = s
        return out / np.float32(0x1000000)


# the forward neighbour offsets (b > a) within sqrt(2) spacings, in build_soft_box's dz, dy, dx loop order
_OFFSETS = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
            if (dz, dy, dx) > (0, 0, 0) and dx * dx + dy * dy + dz * dz <= 2]


def lattice_body(nx, ny, nz, spacing=0.25, node_mass=2.0, origin=(0.0, 1.0, 0.0), k=2e6, d=500.0, eps=0.05,
                 seed=1, jitter=0.01, k_range=None, weak_frac=0.0, weak_eps=0.01, deform_factor=1.0,
                 dt=DT, faces=False, name="lattice", diag_keep=1.0):
    """A node lattice body; origin is the (x, z) centre and the bottom y. diag_keep < 1 keeps every axis beam and each
    diagonal beam with that probability (a random network of lower degree)."""
    rng = TRng(seed)
    n = nx * ny * nz
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    s = np.float32(spacing)
    x0 = np.float32(origin[0]) - np.float32((nx - 1) * spacing * 0.5)
    y0 = np.float32(origin[1])
    z0 = np.float32(origin[2]) - np.float32((nz - 1) * spacing * 0.5)
    p = np.stack([x0 + ix.astype(np.float32) * s, y0 + iy.astype(np.float32) * s, z0 + iz.astype(np.float32) * s], 1)
    p = p.astype(np.float32)
    if jitter > 0:
        u = rng.uniforms(6 * n).reshape(n, 6)  # 3 position draws, then 3 (zero) velocity draws per node
        jt = np.float32(jitter)
        p = (p + (-jt + (jt - (-jt)) * u[:, :3])).astype(np.float32)
    body = SoftBody(name)
    nodes = np.zeros(n, abi.NODE)
    nodes["p"] = p
    nodes["mass"] = np.float32(node_mass)
    nodes["inv_mass"] = np.float32(1.0) / np.float32(node_mass)
    info = np.zeros(n, abi.NODE_INFO)
    info["flags"] = abi.NF_GROUND | abi.NF_CONTACTER
    info["friction"] = 1.0
    info["frame"] = -1
    info["ror_id"] = -1
    body.nodes, body.info = nodes, info

    ids = (iz * ny + iy) * nx + ix
    cols_b, cols_m = [], []
    for dx, dy, dz in _OFFSETS:
        x2, y2, z2 = ix + dx, iy + dy, iz + dz
        ok = (x2 >= 0) & (y2 >= 0) & (z2 >= 0) & (x2 < nx) & (y2 < ny) & (z2 < nz)
        cols_b.append(np.where(ok, (z2 * ny + y2) * nx + x2, -1))
        cols_m.append(ok)
    bmat = np.stack(cols_b, 1)
    mmat = np.stack(cols_m, 1)
    a_idx = np.repeat(ids, len(_OFFSETS)).reshape(n, len(_OFFSETS))
    if diag_keep < 1.0:
        diag = np.array([dx * dx + dy * dy + dz * dz > 1 for dx, dy, dz in _OFFSETS])
        u = rng.uniforms(n * len(_OFFSETS)).reshape(n, len(_OFFSETS))
        mmat = mmat & (~diag[None, :] | (u < np.float32(diag_keep)))
    sel = mmat.ravel()
    beams = np.zeros(int(sel.sum()), abi.BEAM)
    beams["a"] = a_idx.ravel()[sel]
    beams["b"] = bmat.ravel()[sel]
    nb = len(beams)
    if k_range is not None:
        u = rng.uniforms(nb).astype(np.float64)
        lo, hi = math.log(k_range[0]), math.log(k_range[1])
        beams["k"] = np.exp(lo + u * (hi - lo)).astype(np.float32)
    else:
        beams["k"] = np.float32(k)
    beams["d"] = np.float32(d)
    beams["support_limit"] = np.float32(4.0)
    body.beams = beams
    if faces:
        body.tris = box_faces(nx, ny, nz)
    body.finalize()
    body.stabilize(float(dt))
    eps_arr = np.full(nb, np.float32(eps), np.float32)
    if weak_frac > 0:
        u = rng.uniforms(nb)
        eps_arr[u < np.float32(weak_frac)] = np.float32(weak_eps)
    strength = (eps_arr * body.beams["k"]) * body.beams["L0"]
    body.beams["strength"] = strength.astype(np.float32)
    deform = (strength * np.float32(deform_factor)).astype(np.float32)
    body.beams["max_pos"] = deform
    body.beams["max_neg"] = -deform
    return body


def box_faces(nx, ny, nz):
    """The lattice's 6 faces as triangles, wound outwards (build_soft_box's faces, src/world/objects.cpp:923-958)."""
    dims = (nx, ny, nz)

    def nid(c):
        return (c[2] * ny + c[1]) * nx + c[0]

    tris = []
    for axis in range(3):
        for side in range(2):
            u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
            fixed = dims[axis] - 1 if side else 0
            for v in range(dims[v_ax] - 1):
                for u in range(dims[u_ax] - 1):
                    q = []
                    for du, dv in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        c = [0, 0, 0]
                        c[axis], c[u_ax], c[v_ax] = fixed, u + du, v + dv
                        q.append(nid(c))
                    if side:
                        tris += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
                    else:
                        tris += [(q[0], q[2], q[1]), (q[0], q[3], q[2])]
    t = np.zeros(len(tris), abi.TRIANGLE)
    arr = np.array(tris, np.uint32)
    t["a"], t["b"], t["c"] = arr[:, 0], arr[:, 1], arr[:, 2]
    t["two_sided"] = 1
    return t


def ball_body(center, radius, mass, friction=1.0, bounce=-1.0, name="ball"):
    """A ball (build_ball, src/world/objects.cpp:1213-1232): one node, a capsule of no length, the node's static
    contact radius the ball's."""
    b = SoftBody(name)
    nodes = np.zeros(1, abi.NODE)
    nodes["p"][0] = np.asarray(center, np.float32)
    nodes["mass"] = np.float32(mass)
    nodes["inv_mass"] = np.float32(1.0) / np.float32(mass)
    info = np.zeros(1, abi.NODE_INFO)
    info["flags"] = abi.NF_GROUND | abi.NF_CONTACTER
    info["friction"] = 1.0
    info["radius"] = np.float32(radius)
    info["frame"] = -1
    info["ror_id"] = -1
    b.nodes, b.info = nodes, info
    caps = np.zeros(1, abi.CAPSULE)
    caps["a"], caps["b"], caps["radius"], caps["joint"] = 0, 0, np.float32(radius), -1
    b.capsules = caps
    b.opts.collision_radius = np.float32(0.01)
    b.opts.sphere_ball = np.float32(radius)
    b.opts.ground_friction = np.float32(friction)
    b.opts.bounce = np.float32(bounce)
    return b


def ground_box(world_or_statics, half=100.0):
    """A concrete ground slab with its top at y = 0."""
    st = world_or_statics.statics if hasattr(world_or_statics, "statics") else world_or_statics
    st.add_box((0.0, -0.5, 0.0), (half, 0.5, half), (0.0, 0.0, 0.0, 1.0), abi.SURF_CONCRETE)


# ---------------------------------------------------------------------------------------------- the configurations
def config1_body(seed=1):
    """20x8x4 lattice, 0.25 m, 2 kg/node, k = 2e6, d = 500, break strain 0.05, dropped from 1 m, +-1 cm jitter."""
    return lattice_body(20, 8, 4, 0.25, 2.0, (0.0, 1.0, 0.0), k=2e6, d=500.0, eps=0.05, seed=seed, jitter=0.01,
                        name="c1_lattice")


def config2_body(seed=2, faces=False):
    """64x32x24 lattice, 0.25 m, 2 kg/node, k log-uniform in [1e5, 5e6], 2 % of beams at break strain 0.01."""
    return lattice_body(64, 32, 24, 0.25, 2.0, (0.0, 0.5, 0.0), k=None, d=500.0, eps=0.05, seed=seed, jitter=0.01,
                        k_range=(1e5, 5e6), weak_frac=0.02, weak_eps=0.01, faces=faces, name="c2_lattice")


def config2_scene(seed=2, n_balls=2000):
    """The 64x32x24 lattice with its faces as triangles and self collision on (collision radius 0.04, as
    build_soft_box sets), standing on the ground (its bottom 1 cm up, the nodes' jitter +-1 cm), and `n_balls` rigid
    balls (radius uniform in [0.1, 0.3] m, 5 kg): one layer resting on the block's top (each ball 5 mm above its
    radius from the top face, inside the triangles' contact reach: ball-lattice contacts from the first substep), 1 m
    in from its edges, the rest on the ground around it (1 mm into the ground), nearest first, the first ring 5 cm
    from its sides (0.7 m grid, +-2 cm jitter). Every body stays below 8 m/s (the slow bodies' pair schedule) over
    4000 substeps."""
    y0 = 0.01
    body = lattice_body(64, 32, 24, 0.25, 2.0, (0.0, y0, 0.0), k=None, d=500.0, eps=0.05, seed=seed, jitter=0.01,
                        k_range=(1e5, 5e6), weak_frac=0.02, weak_eps=0.01, faces=True, name="c2_lattice")
    body.opts.self_collision = 1
    body.opts.collision_radius = np.float32(0.04)
    rng = TRng(seed + 1000)
    hx, hz, top = 7.875, 2.875, y0 + 31 * 0.25
    spots = [(-6.65 + 0.7 * ix, None, -1.75 + 0.7 * iz) for iz in range(6) for ix in range(20)]
    xs = sorted({round(s * (hx + 0.35 + 0.7 * i), 4) for s in (-1, 1) for i in range(40)} |
                {round(-7.7 + 0.7 * j, 4) for j in range(23)})
    zs = sorted({round(s * (hz + 0.35 + 0.7 * i), 4) for s in (-1, 1) for i in range(40)} |
                {round(-2.8 + 0.7 * j, 4) for j in range(9)})
    ring = []
    for x in xs:
        for z in zs:
            if abs(x) < hx + 0.3 and abs(z) < hz + 0.3:
                continue
            d = max(abs(x) - hx, abs(z) - hz, 0.0)
            ring.append((round(d, 4), abs(z), abs(x), x, z))
    ring.sort()
    balls = []
    for k in range(n_balls):
        r = rng.range(0.1, 0.3)
        jx, jz = rng.range(-0.02, 0.02), rng.range(-0.02, 0.02)
        if k < len(spots):
            x, _, z = spots[k]
            y = top + float(r) + 0.005
        else:
            x, z = ring[k - len(spots)][3:]
            y = float(r) - 0.001
        balls.append(ball_body((np.float32(x) + jx, np.float32(y), np.float32(z) + jz), r, 5.0, name=f"ball{k}"))
    return [body] + balls


def make_world(bodies, ground=True, settings=None):
    w = World(settings)
    if ground:
        ground_box(w)
    for b in bodies:
        w.add_body(b)
    return w


def config4_sizes(n_bodies=512, seed=4):
    """Node counts of the batched scene: lognormal (median 2000, sigma 0.5) clipped to [500, 8000]."""
    rng = np.random.default_rng(seed)
    return np.clip(np.round(np.exp(np.log(2000.0) + 0.5 * rng.standard_normal(n_bodies))), 500, 8000).astype(int)


def crate_body(nx, ny, nz, spacing=0.25, node_mass=2.0, origin=(0.0, 1.0, 0.0), k=2e6, d=500.0, eps=0.05, seed=1,
               jitter=0.01, struts=0.5, dt=DT, name="crate"):
    """A hollow crate: the surface nodes of an nx x ny x nz lattice, its faces triangulated (box_faces, two-sided
    collision triangles), a beam on every triangle edge (surface degree about 6), and struts * n random internal
    beams between surface nodes (seeded). origin is the (x, z) centre and the bottom y."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    surf = (ix == 0) | (iy == 0) | (iz == 0) | (ix == nx - 1) | (iy == ny - 1) | (iz == nz - 1)
    remap = -np.ones(nx * ny * nz, np.int64)
    remap[surf] = np.arange(int(surf.sum()))
    n = int(surf.sum())
    rng = TRng(seed)
    s = np.float32(spacing)
    x0 = np.float32(origin[0]) - np.float32((nx - 1) * spacing * 0.5)
    y0 = np.float32(origin[1])
    z0 = np.float32(origin[2]) - np.float32((nz - 1) * spacing * 0.5)
    p = np.stack([x0 + ix[surf].astype(np.float32) * s, y0 + iy[surf].astype(np.float32) * s,
                  z0 + iz[surf].astype(np.float32) * s], 1).astype(np.float32)
    if jitter > 0:
        u = rng.uniforms(3 * n).reshape(n, 3)
        jt = np.float32(jitter)
        p = (p + (-jt + (jt - (-jt)) * u)).astype(np.float32)
    body = SoftBody(name)
    nodes = np.zeros(n, abi.NODE)
    nodes["p"] = p
    nodes["mass"] = np.float32(node_mass)
    nodes["inv_mass"] = np.float32(1.0) / np.float32(node_mass)
    info = np.zeros(n, abi.NODE_INFO)
    info["flags"] = abi.NF_GROUND | abi.NF_CONTACTER
    info["friction"] = 1.0
    info["frame"] = -1
    info["ror_id"] = -1
    body.nodes, body.info = nodes, info
    tris = box_faces(nx, ny, nz)
    for c in "abc":
        tris[c] = remap[tris[c]]
    body.tris = tris
    t = np.stack([tris["a"], tris["b"], tris["c"]], 1).astype(np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e = np.unique(np.sort(e, 1), axis=0)
    m = int(round(struts * n))
    u = rng.uniforms(2 * m).reshape(m, 2)
    sa = np.minimum((u[:, 0] * n).astype(np.int64), n - 1)
    sb = np.minimum((u[:, 1] * n).astype(np.int64), n - 1)
    st = np.sort(np.stack([sa, sb], 1), 1)
    st = st[st[:, 0] != st[:, 1]]
    allb = np.unique(np.concatenate([e, st]), axis=0)
    beams = np.zeros(len(allb), abi.BEAM)
    beams["a"], beams["b"] = allb[:, 0], allb[:, 1]
    beams["k"], beams["d"] = np.float32(k), np.float32(d)
    beams["support_limit"] = np.float32(4.0)
    body.beams = beams
    body.finalize()
    body.stabilize(float(dt))
    strength = (np.float32(eps) * body.beams["k"]) * body.beams["L0"]
    body.beams["strength"] = strength.astype(np.float32)
    body.beams["max_pos"] = strength
    body.beams["max_neg"] = -strength
    return body


def config4_body(n_nodes, origin, seed):
    """A body of the batched scene: a hollow crate (crate_body) of about n_nodes surface nodes, 0.2 m spacing,
    2 kg/node, k = 2e6, d = 500, break strain 0.05, a strut per two nodes (mean degree about 7)."""
    side = 1
    while 6 * (side + 1) ** 2 - 12 * (side + 1) + 8 <= n_nodes:
        side += 1
    nx = ny = nz = max(2, side)
    return crate_body(nx, ny, nz, 0.2, 2.0, origin, k=2e6, d=500.0, eps=0.05, seed=seed, jitter=0.005,
                      name=f"c4_body_{seed}")


def stacked_crates(sizes, arena, seed):
    """Crates (config4_body, one per size) in stacks of two on a grid over an arena x arena ground: each stack's larger
    crate stands on the ground (its bottom 1 cm up, the nodes' jitter 5 mm), the smaller one on it, its bottom 4 cm
    above the lower one's top - inside the triangles' contact reach (the bodies' 5 cm collision radius), so every
    stack has inter-body contacts from the first substep."""
    n = len(sizes)
    stacks = (n + 1) // 2
    g = int(math.ceil(math.sqrt(stacks)))
    pitch = arena / g
    rng = np.random.default_rng(seed + 1)
    bodies = []
    for k in range(stacks):
        cx = -0.5 * arena + (k % g + 0.5) * pitch + rng.uniform(-0.1, 0.1) * pitch
        cz = -0.5 * arena + (k // g + 0.5) * pitch + rng.uniform(-0.1, 0.1) * pitch
        pair = sorted(range(2 * k, min(2 * k + 2, n)), key=lambda i: -int(sizes[i]))
        y = 0.01
        for i in pair:
            b = config4_body(int(sizes[i]), (float(cx), y, float(cz)), seed=1000 + i)
            bodies.append(b)
            y = float(b.nodes["p"][:, 1].max()) + 0.04
    return bodies


def config4_scene(n_bodies=512, arena=200.0, seed=4):
    """The batched scene: n_bodies random beam networks (config4_sizes) in stacks of two (stacked_crates) over a
    200 x 200 m arena, inter-body contacts on."""
    return stacked_crates(config4_sizes(n_bodies, seed), arena, seed)
