"""TEST INFRASTRUCTURE - ctypes access to the oracle (liboracle.so, the C restatement in this folder) and to the
compiled reference (oracle/_ref/libblref.so, built from the reference's sources by oracle/Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module; the product never does.
Both worlds take the same records as the C ABI (include/beamlab_gpu.h), so one input feeds the GPU, the oracle
and the reference alike.
"""
import atexit
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libblref.so")
REF_SRC = os.environ.get("BL_REFERENCE", "/root/reference")

_P = ctypes.c_void_p


def build(ref=True, jobs=8):
    """Builds the oracle (and the reference library where its sources are present)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-C", HERE, f"-j{jobs}", f"REF={REF_SRC}"] + targets, check=True,
                   stdout=subprocess.DEVNULL)


def have_ref():
    return os.path.exists(REF_LIB)


def _load(path, prefix, sigs):
    lib = ctypes.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
    return lib


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            build(ref=False)
        _oracle = _load(ORACLE_LIB, "orc_", {
            "create": (_P, [_P]),
            "destroy": (None, [_P]),
            "set_statics": (ctypes.c_int, [_P, _P, ctypes.c_int]),
            "set_cylinders": (ctypes.c_int, [_P, _P, ctypes.c_int]),
            "set_terrain": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, _P, _P, _P]),
            "add_body": (ctypes.c_int, [_P, _P]),
            "step": (ctypes.c_int, [_P, ctypes.c_int]),
            "read_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
            "write_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "stats": (ctypes.c_int, [_P, _P]),
            "read_slides": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
            "read_joints": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "read_wheels": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "set_wheel_controls": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "fast_pairs": (ctypes.c_int64, [_P]),
            "rebuild_counts": (None, [_P, _P]),
            "finalize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32]),
            "stabilize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.POINTER(ctypes.c_int)]),
            "box_from_quat": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
            "primitive_collision": (None, [_P, _P, ctypes.c_float, _P, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
            "fem_tri_local_k": (None, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
            "fem_rest_shape": (ctypes.c_int, [_P, _P]),
            "fem_plate": (_P, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _P]),
            "fem_destroy": (None, [_P]),
            "fem_nodes": (ctypes.c_uint32, [_P]),
            "fem_tris": (ctypes.c_uint32, [_P]),
            "fem_topology": (None, [_P, _P, _P, _P, _P, _P]),
            "fem_get_state": (None, [_P, _P, _P, _P, _P]),
            "fem_set_state": (None, [_P, _P, _P, _P, _P]),
            "fem_get_nodes": (None, [_P, _P, _P, _P]),
            "fem_set_fixed": (None, [_P, _P]),
            "fem_create": (_P, [ctypes.c_uint32, _P, _P, _P, ctypes.c_uint32, _P, _P, ctypes.c_uint32, _P,
                                ctypes.c_uint32, _P]),
            "fem_add_torques": (None, [_P, _P]),
            "fem_plastic": (None, [_P, _P, _P, _P]),
            "fem_step_span": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                             _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_uint32)]),
            "fem_held": (ctypes.c_int, [_P, _P, ctypes.c_float, _P]),
            "fem_step": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P, _P, _P, _P,
                                        ctypes.POINTER(ctypes.c_uint32)]),
            "fem_block_solve": (ctypes.c_int, [ctypes.c_int, _P, _P]),
        })
    return _oracle


def ref_lib():
    global _ref
    if _ref is None:
        if not have_ref():
            raise FileNotFoundError(f"{REF_LIB} is not built (reference sources at {REF_SRC} needed)")
        _ref = _load(REF_LIB, "blref_", {
            "create": (_P, [_P, ctypes.c_int]),
            "destroy": (None, [_P]),
            "threads": (ctypes.c_int, []),
            "shutdown": (None, []),
            "set_statics": (ctypes.c_int, [_P, _P, ctypes.c_int]),
            "set_cylinders": (ctypes.c_int, [_P, _P, ctypes.c_int]),
            "set_terrain": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, _P, _P, _P]),
            "box_from_quat": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
            "add_body": (ctypes.c_int, [_P, _P]),
            "step": (ctypes.c_int, [_P, ctypes.c_int]),
            "time_step": (ctypes.c_double, [_P, ctypes.c_int, ctypes.c_int]),
            "read_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
            "write_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "read_slides": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
            "read_joints": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "read_wheels": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "set_wheel_controls": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "finalize_stabilize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.c_int,
                                                  ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                  ctypes.POINTER(ctypes.c_int)]),
            "tri_stiffness": (ctypes.c_int, [_P, _P, _P, _P]),
            "fem_plate_new": (_P, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _P, ctypes.c_int]),
            "fem_plate_free": (None, [_P]),
            "fem_plate_step": (ctypes.c_int64, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
            "fem_plate_substep": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
            "fem_plate_read": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
            "fem_plate_write": (ctypes.c_int, [_P, _P, _P, _P, _P]),
            "frame_new": (_P, [ctypes.c_uint32, _P, _P, _P, ctypes.c_uint32, _P, _P, ctypes.c_uint32, _P,
                               ctypes.c_uint32, _P]),
            "frame_add_torques": (None, [_P, _P]),
            "frame_plastic": (None, [_P, _P, _P, _P]),
            "frame_cycle": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                           ctypes.c_float, _P]),
        })
        atexit.register(_ref.blref_shutdown)
    return _ref


def _ptr(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(ctypes.c_void_p)


class _CpuWorld:
    """Common front of the oracle and reference worlds (same calls as beamlab_amd.world.World)."""

    def __init__(self, lib, prefix, handle, statics=None):
        self._lib, self._pre, self._h = lib, prefix, handle
        self.bodies = []
        if statics is not None and len(statics):
            self._fn("set_statics")(self._h, _ptr(statics), len(statics))

    def _fn(self, name):
        return getattr(self._lib, self._pre + name)

    def set_cylinders(self, cyl):
        """The static world's vertical cylinders (abi.CYLINDER records)."""
        self._cyl = np.ascontiguousarray(cyl)
        self._fn("set_cylinders")(self._h, _ptr(self._cyl), len(self._cyl))

    def set_terrain(self, heights, cell, origin, surfaces=None):
        """A heightfield: heights[z, x] (-1e30: no ground), `cell` apart from origin (x, z); None removes it."""
        if heights is None:
            self._fn("set_terrain")(self._h, 0, 0, 0.0, None, None, None)
            return
        h = np.ascontiguousarray(heights, np.float32)
        o = np.asarray(origin, np.float32)
        s = None if surfaces is None else np.ascontiguousarray(surfaces, np.uint8)
        self._fn("set_terrain")(self._h, h.shape[1], h.shape[0], float(cell), _ptr(o), _ptr(h), _ptr(s))

    def add_body(self, body):
        d = body.desc()
        bid = self._fn("add_body")(self._h, ctypes.byref(d))
        self.bodies.append(body)
        return bid

    def step_substeps(self, n):
        self._fn("step")(self._h, int(n))

    def read_body(self, bid):
        from beamlab_amd import abi
        b = self.bodies[bid]
        nodes = np.zeros(len(b.nodes), abi.NODE)
        beams = np.zeros(len(b.beams), abi.BEAM)
        broken = ctypes.c_int64(0)
        self._fn("read_body")(self._h, bid, _ptr(nodes), _ptr(beams), ctypes.byref(broken))
        return nodes, beams

    def read_slides(self, bid):
        out = self.bodies[bid].slides.copy()
        self._fn("read_slides")(self._h, bid, _ptr(out))
        return out

    def read_joints(self, bid):
        fr, jt = self.bodies[bid].frames.copy(), self.bodies[bid].joints.copy()
        self._fn("read_joints")(self._h, bid, _ptr(fr), _ptr(jt))
        return fr, jt

    def read_wheels(self, bid):
        """(the wheels, [wheel_speed, wheel_spin])"""
        wh, ss = self.bodies[bid].wheels.copy(), np.zeros(2, np.float32)
        self._fn("read_wheels")(self._h, bid, _ptr(wh), _ptr(ss))
        return wh, ss

    def set_wheel_controls(self, bid, torque=None, brake=None):
        if torque is None or brake is None:
            self._fn("set_wheel_controls")(self._h, bid, None, None)
            return
        t, b = np.ascontiguousarray(torque, np.float32), np.ascontiguousarray(brake, np.float32)
        self._fn("set_wheel_controls")(self._h, bid, _ptr(t), _ptr(b))

    def broken_beams(self, bid):
        broken = ctypes.c_int64(0)
        self._fn("read_body")(self._h, bid, None, None, ctypes.byref(broken))
        return broken.value

    def write_body(self, bid, nodes=None, beams=None):
        self._fn("write_body")(self._h, bid, _ptr(nodes), _ptr(beams))

    def close(self):
        if self._h:
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleWorld(_CpuWorld):
    """The C restatement (liboracle.so)."""

    def __init__(self, settings, statics=None):
        lib = oracle_lib()
        super().__init__(lib, "orc_", lib.orc_create(ctypes.byref(settings)), statics)

    def stats(self):
        from beamlab_amd import abi
        s = abi.Stats()
        self._lib.orc_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def rebuild_counts(self):
        """The rebuilds' sizes, summed over every rebuild: hashed nodes, triangles queried, node-triangle candidates
        before and after the near filter, node-capsule and capsule-triangle candidates."""
        out = (ctypes.c_int64 * 6)()
        self._lib.orc_rebuild_counts(self._h, out)
        return dict(zip(("hashed", "tris_queried", "nt_raw", "nt", "nc", "ct"), list(out)))

    def fast_pairs(self):
        """The fast bodies' candidate pairs (World::fast_pairs), summed over the rebuilds."""
        return int(self._lib.orc_fast_pairs(self._h))


class RefWorld(_CpuWorld):
    """The compiled reference (oracle/_ref/libblref.so): World::step_substeps itself."""

    def __init__(self, settings, statics=None, threads=1):
        lib = ref_lib()
        super().__init__(lib, "blref_", lib.blref_create(ctypes.byref(settings), int(threads)), statics)

    def time_substeps(self, calls, n):
        return self._lib.blref_time_step(self._h, int(calls), int(n))


def oracle_stabilize(nodes, beams, dt, kb=0.6, db=0.5):
    b = beams.copy()
    ch = ctypes.c_int(0)
    oracle_lib().orc_stabilize(_ptr(nodes), len(nodes), _ptr(b), len(b), dt, kb, db, ctypes.byref(ch))
    return b, ch.value


def oracle_finalize(nodes, beams):
    b = beams.copy()
    oracle_lib().orc_finalize(_ptr(nodes), len(nodes), _ptr(b), len(b))
    return b


def ref_finalize_stabilize(nodes, beams, finalize, dt, kb=0.6, db=0.5):
    b = beams.copy()
    n = nodes.copy()
    ch = ctypes.c_int(0)
    ref_lib().blref_finalize_stabilize(_ptr(n), len(n), _ptr(b), len(b), int(finalize), dt, kb, db,
                                       ctypes.byref(ch))
    return b, ch.value


def box_from_quat(center, half, quat, surface, which="oracle"):
    from beamlab_amd import abi
    out = np.zeros(1, abi.BOX)
    c, h, q = (np.asarray(x, np.float32) for x in (center, half, quat))
    lib = oracle_lib() if which == "oracle" else ref_lib()
    fn = lib.orc_box_from_quat if which == "oracle" else lib.blref_box_from_quat
    fn(_ptr(c), _ptr(h), _ptr(q), surface, _ptr(out))
    return out[0]


def primitive_collision(F, vel, mass, normal, dt, surface, pen, fric, push_max=2.0, slop=0.0, restitution=-1.0):
    out = np.zeros(3, np.float32)
    F, vel, normal = (np.asarray(x, np.float32) for x in (F, vel, normal))
    oracle_lib().orc_primitive_collision(_ptr(F), _ptr(vel), mass, _ptr(normal), dt, surface, pen, fric, push_max,
                                         slop, restitution, _ptr(out))
    return out


# ---------------------------------------------------------------------------------------------- the shell FEM
SECTION_C3 = (70e9, 0.33, 2700.0, 1.5e-3, 0.05, 3e-4)  # E, nu, rho, t, drill, damping


def fem_tri_local_k(X, section=SECTION_C3):
    """The oracle's element stiffness (18 x 18, float64) of a triangle of rest shape X (3 x 2)."""
    X = np.ascontiguousarray(X, np.float32).reshape(6)
    K = np.zeros((18, 18), np.float64)
    E, nu, _, t, drill, _ = section
    oracle_lib().orc_fem_tri_local_k(_ptr(X), E, nu, t, drill, _ptr(K))
    return K


def fem_rest_shape(p):
    p = np.ascontiguousarray(p, np.float32).reshape(9)
    X = np.zeros(6, np.float32)
    if oracle_lib().orc_fem_rest_shape(_ptr(p), _ptr(X)) != 0:
        raise ValueError("degenerate triangle")
    return X.reshape(3, 2)


def ref_tri_stiffness(p, section=SECTION_C3):
    """The compiled reference's FemFrame::tri_stiffness of the triangle with corners p (3 x 3): (K, X)."""
    p = np.ascontiguousarray(p, np.float32).reshape(9)
    sec = np.asarray(section, np.float32)
    K = np.zeros((18, 18), np.float64)
    X = np.zeros(6, np.float32)
    if ref_lib().blref_tri_stiffness(_ptr(p), _ptr(sec), _ptr(K), _ptr(X)) != 0:
        raise ValueError("degenerate triangle")
    return K, X.reshape(3, 2)


def _frame_arrays(p, mass, fixed, tris, shell, members, sections):
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    nn = len(p)
    mass = np.ascontiguousarray(mass, np.float32).reshape(nn)
    fixed = np.zeros(nn, np.uint8) if fixed is None else np.ascontiguousarray(fixed, np.uint8).reshape(nn)
    tris = np.zeros((0, 3), np.uint32) if tris is None else np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    shell = np.asarray(SECTION_C3 if shell is None else shell, np.float32).reshape(-1)
    if len(shell) == 6:  # (no yield stress: elastic)
        shell = np.append(shell, np.float32(0.0)).astype(np.float32)
    members = np.zeros((0, 4), np.uint32) if members is None else np.ascontiguousarray(members, np.uint32).reshape(-1, 4)
    sections = np.zeros((0, 17), np.float32) if sections is None else np.asarray(sections, np.float32)
    if sections.ndim == 2 and sections.shape[1] == 12:  # (no plastic limits; FrameSection's hardening and half)
        pad = np.tile(np.array([0, 0, 0, 0.25, 0.02], np.float32), (len(sections), 1))
        sections = np.concatenate([sections, pad], axis=1)
    sections = np.ascontiguousarray(sections, np.float32).reshape(-1, 17)
    return p, mass, fixed, tris, shell, members, sections


TRI_STATE, MEM_STATE = 29, 27


def _plastic_record(tris, members, cnt):
    """The plastic state as arrays: triangles' X (3 x 2), th0 (3 x 3), r0 (3 x 4), dmg, util; members' up, tp, pa, pb,
    ba, bb, bt, damage, dmg_a, dmg_b, util, N, qa, qb (and the oracle's tangent bits); {repasses, passes}."""
    return dict(X=tris[:, 0:6].reshape(-1, 3, 2), th0=tris[:, 6:15].reshape(-1, 3, 3), r0=tris[:, 15:27].reshape(-1, 3, 4),
                tdmg=tris[:, 27], tutil=tris[:, 28], mstate=members[:, 0:16], qa=members[:, 16:20], qb=members[:, 20:24],
                mbits=members[:, 24:27], repasses=int(cnt[0]), passes=int(cnt[1]))


def _frame_args(arrs):
    p, mass, fixed, tris, shell, members, sections = arrs
    return (len(p), _ptr(p), _ptr(mass), _ptr(fixed), len(tris), _ptr(tris), _ptr(shell), len(members), _ptr(members),
            len(sections), _ptr(sections))


class RefPlate:
    """The compiled reference's clamped plate (ShellMesher::grid) and its implicit step as its fembench times it:
    clear_forces(g) plus the loads, FemFrame::compute_forces, FemFrame::solve (one thread)."""

    def __init__(self, nu, nv, size=1.0, y0=0.0, section=SECTION_C3, clamped=True):
        self._lib = ref_lib()
        sec = np.asarray(section, np.float32)
        self._h = self._lib.blref_fem_plate_new(int(nu), int(nv), float(size), float(y0), _ptr(sec), int(bool(clamped)))
        self.nn = (nu + 1) * (nv + 1)

    def step(self, loads, h, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0)):
        """One step; returns the frame's solve failures so far."""
        ld = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(-1)
        gg = np.asarray(g, np.float32)
        return self._lib.blref_fem_plate_step(self._h, _ptr(ld), h, theta, dissipation, _ptr(gg))

    @classmethod
    def frame(cls, p, mass, fixed=None, tris=None, shell=None, members=None, sections=None):
        """A frame body (blref_frame_new): nodes, shell triangles of one section, members (a, b, section,
        end_a | end_b << 8) of elastic sections (12 floats each)."""
        self = cls.__new__(cls)
        self._lib = ref_lib()
        self._arrs = _frame_arrays(p, mass, fixed, tris, shell, members, sections)
        self._h = self._lib.blref_frame_new(*_frame_args(self._arrs))
        self.nn = len(self._arrs[0])
        return self

    def add_torques(self, t):
        self._lib.blref_frame_add_torques(self._h, _ptr(np.ascontiguousarray(t, np.float32).reshape(-1)))

    def plastic(self):
        """The frame's plastic state (blref_frame_plastic; the tangent's bits are private there: zeros)."""
        nt, nm = len(self._arrs[3]), len(self._arrs[5])
        tr = np.zeros((nt, TRI_STATE), np.float32)
        me = np.zeros((nm, MEM_STATE), np.float32)
        cnt = np.zeros(2, np.int64)
        self._lib.blref_frame_plastic(self._h, _ptr(tr.reshape(-1)), _ptr(me.reshape(-1)), _ptr(cnt))
        return _plastic_record(tr, me, cnt)

    def cycle(self, loads, dt, every, j, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0)):
        """Substep j of a frame stepping its frame every `every` substeps (blref_frame_cycle): the implicit step over
        every dt at j % every == 0 (the dissipation divided by every), else the held step; then the integration."""
        ld = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(-1)
        gg = np.asarray(g, np.float32)
        self._lib.blref_frame_cycle(self._h, _ptr(ld), dt, int(every), int(j), theta, dissipation, _ptr(gg))

    def substep(self, loads, h, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0)):
        """The plate body's whole substep as the reference's world runs it: the step, then integrate_nodes and
        sync_positions."""
        ld = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(-1)
        gg = np.asarray(g, np.float32)
        self._lib.blref_fem_plate_substep(self._h, _ptr(ld), h, theta, dissipation, _ptr(gg))

    def read(self):
        """{nodes (bl_node records: p, inv_mass, v, mass), xd, w, q, force, counters (failures, clamps)}."""
        nodes = np.zeros((self.nn, 8), np.float32)
        xd = np.zeros((self.nn, 3), np.float64)
        w = np.zeros((self.nn, 3), np.float32)
        q = np.zeros((self.nn, 4), np.float32)
        force = np.zeros((self.nn, 3), np.float32)
        cnt = np.zeros(2, np.int64)
        if self._lib.blref_fem_plate_read(self._h, _ptr(nodes), _ptr(xd), _ptr(w), _ptr(q), _ptr(force), _ptr(cnt)) != self.nn:
            raise RuntimeError("the reference plate's frame nodes are not the body's")
        return dict(p=nodes[:, 0:3].copy(), inv_mass=nodes[:, 3].copy(), v=nodes[:, 4:7].copy(), mass=nodes[:, 7].copy(),
                    xd=xd, w=w, q=q, force=force, failures=int(cnt[0]), clamps=int(cnt[1]))

    def write(self, p=None, v=None, xd=None, w=None, q=None):
        """Overwrites parts of the state (the nodes' p and v, the frame's xd, w, q)."""
        nodes = None
        if p is not None or v is not None:
            nodes = np.zeros((self.nn, 8), np.float32)
            self._lib.blref_fem_plate_read(self._h, _ptr(nodes), None, None, None, None, None)
            if p is not None:
                nodes[:, 0:3] = p
            if v is not None:
                nodes[:, 4:7] = v
        cv = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
        self._lib.blref_fem_plate_write(self._h, _ptr(nodes), _ptr(cv(xd, np.float64)), _ptr(cv(w, np.float32)),
                                        _ptr(cv(q, np.float32)))

    def close(self):
        if self._h:
            self._lib.blref_fem_plate_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FemPlate:
    """The oracle's plate (ShellMesher::grid) and its implicit step (see fem_oracle.c)."""

    def __init__(self, nu, nv, size=1.0, y0=0.0, section=SECTION_C3, clamped=True):
        self._lib = oracle_lib()
        sec = np.asarray(section, np.float32)
        self._h = self._lib.orc_fem_plate(int(nu), int(nv), float(size), float(y0), _ptr(sec))
        self.nn = self._lib.orc_fem_nodes(self._h)
        self.nt = self._lib.orc_fem_tris(self._h)
        if not clamped:
            self._lib.orc_fem_set_fixed(self._h, _ptr(np.zeros(self.nn, np.uint8)))

    @classmethod
    def frame(cls, p, mass, fixed=None, tris=None, shell=None, members=None, sections=None):
        """A frame body (orc_fem_create): nodes, shell triangles of one section, members (a, b, section,
        end_a | end_b << 8) of elastic sections (12 floats each)."""
        self = cls.__new__(cls)
        self._lib = oracle_lib()
        self._arrs = _frame_arrays(p, mass, fixed, tris, shell, members, sections)
        self._h = self._lib.orc_fem_create(*_frame_args(self._arrs))
        self.nn = self._lib.orc_fem_nodes(self._h)
        self.nt = self._lib.orc_fem_tris(self._h)
        return self

    def add_torques(self, t):
        self._lib.orc_fem_add_torques(self._h, _ptr(np.ascontiguousarray(t, np.float32).reshape(-1)))

    def plastic(self):
        """The frame's plastic state (orc_fem_plastic)."""
        nm = len(self._arrs[5]) if hasattr(self, "_arrs") else 0
        tr = np.zeros((self.nt, TRI_STATE), np.float32)
        me = np.zeros((nm, MEM_STATE), np.float32)
        cnt = np.zeros(2, np.int64)
        self._lib.orc_fem_plastic(self._h, _ptr(tr.reshape(-1)), _ptr(me.reshape(-1)), _ptr(cnt))
        return _plastic_record(tr, me, cnt)

    def topology(self):
        tn = np.zeros((self.nt, 3), np.uint32)
        xd = np.zeros((self.nn, 3), np.float64)
        fixed = np.zeros(self.nn, np.uint8)
        mass = np.zeros(self.nn, np.float32)
        inertia = np.zeros(self.nn, np.float32)
        self._lib.orc_fem_topology(self._h, _ptr(tn), _ptr(xd), _ptr(fixed), _ptr(mass), _ptr(inertia))
        return tn, xd, fixed, mass, inertia

    def state(self):
        xd = np.zeros((self.nn, 3), np.float64)
        v = np.zeros((self.nn, 3), np.float32)
        w = np.zeros((self.nn, 3), np.float32)
        q = np.zeros((self.nn, 4), np.float32)
        self._lib.orc_fem_get_state(self._h, _ptr(xd), _ptr(v), _ptr(w), _ptr(q))
        return xd, v, w, q

    def read(self):
        """The same record as RefPlate.read: the body's nodes, the frame's xd, w, q, the forces, the counters."""
        nodes = np.zeros((self.nn, 8), np.float32)
        force = np.zeros((self.nn, 3), np.float32)
        cnt = np.zeros(3, np.int64)
        self._lib.orc_fem_get_nodes(self._h, _ptr(nodes), _ptr(force), _ptr(cnt))
        xd, v, w, q = self.state()
        return dict(p=nodes[:, 0:3].copy(), inv_mass=nodes[:, 3].copy(), v=nodes[:, 4:7].copy(), mass=nodes[:, 7].copy(),
                    xd=xd, w=w, q=q, force=force, failures=int(cnt[0]), clamps=int(cnt[1]), nan_resets=int(cnt[2]))

    def set_state(self, xd, v, w, q):
        self._lib.orc_fem_set_state(self._h, *[_ptr(np.ascontiguousarray(a, t)) for a, t in
                                              ((xd, np.float64), (v, np.float32), (w, np.float32), (q, np.float32))])

    def cycle(self, loads, dt, every, j, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0)):
        """Substep j of a frame stepping its frame every `every` substeps, as RefPlate.cycle: orc_fem_step_span over
        every dt (the dissipation divided by every) at j % every == 0, else orc_fem_held."""
        ld = np.zeros((self.nn, 3), np.float32) if loads is None else np.ascontiguousarray(loads, np.float32).reshape(self.nn, 3)
        gg = np.asarray(g, np.float32)
        if j % every == 0:
            nf = ctypes.c_uint32(0)
            self._lib.orc_fem_step_span(self._h, _ptr(ld), np.float32(dt) * np.float32(every), dt, theta,
                                        np.float32(dissipation) / np.float32(every), _ptr(gg), None, None, None, None,
                                        ctypes.byref(nf))
        else:
            self._lib.orc_fem_held(self._h, _ptr(ld), dt, _ptr(gg))

    def step(self, loads, h, theta=0.5, dissipation=0.02, g=(0.0, -9.81, 0.0), system=False):
        """One implicit step; with system=True also (A, rhs, x, free nodes) of the step's linear system."""
        if loads is None:
            loads = np.zeros((self.nn, 3), np.float32)
        loads = np.ascontiguousarray(loads, np.float32).reshape(self.nn, 3)
        gg = np.asarray(g, np.float32)
        nf = ctypes.c_uint32(0)
        if system:
            fixed = self.topology()[2]
            n = int((fixed == 0).sum())
            A = np.zeros((6 * n, 6 * n), np.float64)
            r = np.zeros(6 * n, np.float64)
            x = np.zeros(6 * n, np.float64)
            fm = np.zeros(n, np.uint32)
            rc = self._lib.orc_fem_step(self._h, _ptr(loads), h, theta, dissipation, _ptr(gg), _ptr(A), _ptr(r), _ptr(x),
                                        _ptr(fm), ctypes.byref(nf))
            if rc != 0:
                raise RuntimeError("oracle FEM step: a pivot block is not positive definite")
            return A, r, x, fm
        rc = self._lib.orc_fem_step(self._h, _ptr(loads), h, theta, dissipation, _ptr(gg), None, None, None, None,
                                    ctypes.byref(nf))
        if rc != 0:
            raise RuntimeError("oracle FEM step: a pivot block is not positive definite")

    def close(self):
        if self._h:
            self._lib.orc_fem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fem_block_solve(A, b):
    """The reference's dense block Cholesky (chol6, right_solve_lt, sub_abt; forward6, backward6) on A x = b."""
    A = np.array(A, np.float64, order="C", copy=True)
    x = np.array(b, np.float64, copy=True)
    if oracle_lib().orc_fem_block_solve(len(x) // 6, _ptr(A), _ptr(x)) != 0:
        raise RuntimeError("not positive definite")
    return x
