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
            "add_body": (ctypes.c_int, [_P, _P]),
            "step": (ctypes.c_int, [_P, ctypes.c_int]),
            "read_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
            "write_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "stats": (ctypes.c_int, [_P, _P]),
            "finalize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32]),
            "stabilize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.POINTER(ctypes.c_int)]),
            "box_from_quat": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
            "primitive_collision": (None, [_P, _P, ctypes.c_float, _P, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
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
            "box_from_quat": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
            "add_body": (ctypes.c_int, [_P, _P]),
            "step": (ctypes.c_int, [_P, ctypes.c_int]),
            "time_step": (ctypes.c_double, [_P, ctypes.c_int, ctypes.c_int]),
            "read_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
            "write_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
            "finalize_stabilize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.c_int,
                                                  ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                  ctypes.POINTER(ctypes.c_int)]),
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
