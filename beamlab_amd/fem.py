"""Host-side mirror of the reference's shell FEM over the C ABI (bl_fem_*, include/beamlab_gpu.h).

ShellMesher.grid follows src/phys/fem_shell.cpp:34-46 (a grid of nu x nv cells, two triangles a cell, the diagonals
alternating with (i + j) & 1, each triangle's mass rho t A in thirds on its corners); ShellSection the elastic
fields of src/phys/frame_fem.h:117-149; ShellBody.step one implicit step of FemFrame::solve
(src/phys/frame_fem.cpp:2514-3027) for a body of shell triangles, computed on the GPU.
"""
import ctypes

import numpy as np

from . import abi

# the plate section of the implicit-FEM workload: aluminium, 1.5 mm (E, nu, rho, t, drill, damping)
SECTION_PLATE = (70e9, 0.33, 2700.0, 1.5e-3, 0.05, 3e-4)


def plate_mesh(nu, nv, size=1.0, y0=0.0):
    """ShellMesher::grid over a size x size square in the x-z plane centred on the origin at height y0, its border
    nodes clamped: (positions float32 [n, 3], triangles uint32 [t, 3], fixed uint8 [n])."""
    origin = np.array([-0.5 * size, y0, -0.5 * size], np.float32)
    du = np.array([np.float32(size) / np.float32(nu), 0, 0], np.float32)
    dv = np.array([0, 0, np.float32(size) / np.float32(nv)], np.float32)
    j, i = np.meshgrid(np.arange(nv + 1, dtype=np.float32), np.arange(nu + 1, dtype=np.float32), indexing="ij")
    # origin + du i + dv j, float arithmetic in the reference's order
    p = (origin[None, None, :] + du[None, None, :] * i[..., None]) + dv[None, None, :] * j[..., None]
    p = p.reshape(-1, 3).astype(np.float32)
    tris = []
    for jj in range(nv):
        for ii in range(nu):
            n00 = jj * (nu + 1) + ii
            n10, n01 = n00 + 1, n00 + nu + 1
            n11 = n01 + 1
            if (ii + jj) & 1:
                tris += [(n00, n10, n11), (n00, n11, n01)]
            else:
                tris += [(n00, n10, n01), (n10, n11, n01)]
    ii, jj = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1))
    fixed = ((ii == 0) | (jj == 0) | (ii == nu) | (jj == nv)).reshape(-1).astype(np.uint8)
    return p, np.asarray(tris, np.uint32), fixed


class ShellBody:
    """A shell body stepped implicitly on the GPU: bl_fem_create / bl_fem_step / bl_fem_get_state."""

    def __init__(self, p, tris, fixed=None, section=SECTION_PLATE, mass=None, members=None, sections=None):
        """members: (a, b, section, end_a | end_b << 8) rows or an abi.FRAME_MEMBER array; sections: rows of the 17
        abi.FrameSection fields (rows of the first 12 - no plastic limits - take FrameSection's defaults for the rest);
        section: the shell's 6 or 7 abi.ShellSection fields (a 7th, the yield stress, 0 when left out)."""
        self._lib = abi.load()
        self.p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        self.tris = np.zeros((0, 3), np.uint32) if tris is None else np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        self.nn, self.nt = len(self.p), len(self.tris)
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        self.mass = None if mass is None else np.ascontiguousarray(mass, np.float32)
        self.members = np.zeros(0, abi.FRAME_MEMBER)
        if members is not None and len(members):
            m = np.asarray(members)
            if m.dtype != abi.FRAME_MEMBER:
                m = m.astype(np.uint32).reshape(-1, 4)
                mm = np.zeros(len(m), abi.FRAME_MEMBER)
                mm["a"], mm["b"], mm["section"] = m[:, 0], m[:, 1], m[:, 2]
                mm["end_a"], mm["end_b"] = m[:, 3] & 0xff, m[:, 3] >> 8
                m = mm
            self.members = np.ascontiguousarray(m)
        secs = np.zeros((0, abi.FRAME_SECTION_FIELDS), np.float32) if sections is None else np.asarray(sections, np.float32)
        if secs.ndim == 2 and secs.shape[1] == 12:  # (no plastic limits; FrameSection's hardening and half)
            secs = np.concatenate([secs, np.tile(np.array([0, 0, 0, 0.25, 0.02], np.float32), (len(secs), 1))], axis=1)
        self.sections = np.ascontiguousarray(secs, np.float32).reshape(-1, abi.FRAME_SECTION_FIELDS)
        d = abi.FemDesc()
        d.n_nodes, d.n_tris = self.nn, self.nt
        d.p, d.tris = abi.ptr(self.p.reshape(-1)), abi.ptr(self.tris.reshape(-1))
        d.fixed = abi.ptr(self.fixed)
        d.mass = abi.ptr(self.mass)
        d.section = abi.ShellSection(*[float(x) for x in section])
        d.n_members, d.n_sections = len(self.members), len(self.sections)
        d.members = abi.ptr(self.members) if len(self.members) else None
        d.sections = abi.ptr(self.sections.reshape(-1)) if len(self.sections) else None
        h = ctypes.c_void_p()
        abi.check(self._lib.bl_fem_create(ctypes.byref(h), ctypes.byref(d)), "bl_fem_create")
        self._h = h
        self.report = abi.FemReport()
        abi.track(self)

    @classmethod
    def plate(cls, nu, nv, size=1.0, y0=0.0, section=SECTION_PLATE, clamped=True):
        p, t, fx = plate_mesh(nu, nv, size, y0)
        return cls(p, t, fx if clamped else np.zeros_like(fx), section)

    @classmethod
    def frame(cls, p, mass, fixed=None, tris=None, section=SECTION_PLATE, members=None, sections=None):
        """A frame body: nodes with their masses, shell triangles of one section, members of elastic sections."""
        return cls(p, tris, fixed, section, mass, members, sections)

    def add_torques(self, t):
        """Moments from outside on the nodes for the next step (N m, 3 per node)."""
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        abi.check(self._lib.bl_fem_add_torques(self._h, abi.ptr(t)), "bl_fem_add_torques")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bl_fem_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def step(self, loads=None, h=0.0005, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0)):
        """One implicit step; returns the report (residual, timings, factorization figures) as a dict."""
        ld = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(-1)
        gg = np.asarray(g, np.float32)
        abi.check(self._lib.bl_fem_step(self._h, abi.ptr(ld), h, theta, dissipation, abi.ptr(gg),
                                        ctypes.byref(self.report)), "bl_fem_step")
        return self.report.as_dict()

    def cycle(self, loads, dt, every, j, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0)):
        """Substep j of a frame stepping its frame every `every` substeps (WorldSettings::frame_every): at
        j % every == 0 the implicit step over every dt (bl_fem_step_span, the dissipation divided by every), else the
        held substep (bl_fem_step_held)."""
        ld = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(-1)
        gg = np.asarray(g, np.float32)
        if j % every == 0:
            abi.check(self._lib.bl_fem_step_span(self._h, abi.ptr(ld), np.float32(dt) * np.float32(every), dt, theta,
                                                 np.float32(dissipation) / np.float32(every), abi.ptr(gg),
                                                 ctypes.byref(self.report)), "bl_fem_step_span")
        else:
            abi.check(self._lib.bl_fem_step_held(self._h, abi.ptr(ld), dt, abi.ptr(gg)), "bl_fem_step_held")

    def state(self):
        x = np.zeros((self.nn, 3), np.float64)
        v = np.zeros((self.nn, 3), np.float32)
        w = np.zeros((self.nn, 3), np.float32)
        q = np.zeros((self.nn, 4), np.float32)
        abi.check(self._lib.bl_fem_get_state(self._h, abi.ptr(x.reshape(-1)), abi.ptr(v.reshape(-1)),
                                             abi.ptr(w.reshape(-1)), abi.ptr(q.reshape(-1))), "bl_fem_get_state")
        return x, v, w, q

    def read(self):
        """The body after the last substep: its nodes' p, inv_mass, v, mass (SoftBody::nodes), the frame's xd, w, q,
        the forces handed to the integrator, and the counters (solve failures, clamps, NaN resets)."""
        nodes = np.zeros((self.nn, 8), np.float32)
        force = np.zeros((self.nn, 3), np.float32)
        cnt = np.zeros(3, np.int64)
        abi.check(self._lib.bl_fem_get_nodes(self._h, abi.ptr(nodes.reshape(-1)), abi.ptr(force.reshape(-1)),
                                             abi.ptr(cnt)), "bl_fem_get_nodes")
        xd, v, w, q = self.state()
        return dict(p=nodes[:, 0:3].copy(), inv_mass=nodes[:, 3].copy(), v=nodes[:, 4:7].copy(), mass=nodes[:, 7].copy(),
                    xd=xd, w=w, q=q, force=force, failures=int(cnt[0]), clamps=int(cnt[1]), nan_resets=int(cnt[2]))

    def set_state(self, x=None, v=None, w=None, q=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dt).reshape(-1)
                for a, dt in ((x, np.float64), (v, np.float32), (w, np.float32), (q, np.float32))]
        abi.check(self._lib.bl_fem_set_state(self._h, *[abi.ptr(a) for a in arrs]), "bl_fem_set_state")

    def last_system(self, dense=True):
        nf = ctypes.c_uint32()
        abi.check(self._lib.bl_fem_last_system(self._h, None, None, None, ctypes.byref(nf)), "bl_fem_last_system")
        n6 = 6 * nf.value
        A = np.zeros((n6, n6), np.float64) if dense else None
        r = np.zeros(n6, np.float64)
        x = np.zeros(n6, np.float64)
        abi.check(self._lib.bl_fem_last_system(self._h, abi.ptr(None if A is None else A.reshape(-1)), abi.ptr(r),
                                               abi.ptr(x), ctypes.byref(nf)), "bl_fem_last_system")
        return A, r, x

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64).reshape(-1)
        x = np.zeros_like(b)
        abi.check(self._lib.bl_fem_solve(self._h, abi.ptr(b), abi.ptr(x)), "bl_fem_solve")
        return x

    def plastic(self):
        """The plastic state (bl_fem_get_plastic): the triangles' X (3 x 2), th0 (3 x 3), r0 (3 x 4), dmg, util; the
        members' up, tp, pa, pb, ba, bb, bt, damage, dmg_a, dmg_b, util, N (mstate), qa, qb and the tangent's yield,
        unload and elastic_hold bits; the counters repasses and passes."""
        tr = np.zeros((self.nt, abi.FEM_TRI_STATE), np.float32)
        me = np.zeros((len(self.members), abi.FEM_MEM_STATE), np.float32)
        cnt = np.zeros(2, np.int64)
        abi.check(self._lib.bl_fem_get_plastic(self._h, abi.ptr(tr.reshape(-1)), abi.ptr(me.reshape(-1)), abi.ptr(cnt)),
                  "bl_fem_get_plastic")
        return dict(X=tr[:, 0:6].reshape(-1, 3, 2), th0=tr[:, 6:15].reshape(-1, 3, 3), r0=tr[:, 15:27].reshape(-1, 3, 4),
                    tdmg=tr[:, 27], tutil=tr[:, 28], mstate=me[:, 0:16], qa=me[:, 16:20], qb=me[:, 20:24],
                    mbits=me[:, 24:27], repasses=int(cnt[0]), passes=int(cnt[1]))

    def tri_k(self, k):
        K = np.zeros((18, 18), np.float64)
        abi.check(self._lib.bl_fem_tri_k(self._h, k, abi.ptr(K.reshape(-1))), "bl_fem_tri_k")
        return K
