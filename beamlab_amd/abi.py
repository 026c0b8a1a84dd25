"""ctypes view of the C ABI (include/beamlab_gpu.h) and numpy dtypes of its records.

The record dtypes are layout-identical to the reference's Node / NodeInfo / Beam / Triangle / Capsule
(src/phys/softbody.h:43-196), as the C header is.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbeamlab_gpu.so")

NODE = np.dtype([("p", "<f4", 3), ("inv_mass", "<f4"), ("v", "<f4", 3), ("mass", "<f4")])
NODE_INFO = np.dtype([("flags", "<u2"), ("surface", "u1"), ("pad", "u1"), ("friction", "<f4"), ("radius", "<f4"),
                      ("frame", "<i4"), ("ror_id", "<i4")])
BEAM = np.dtype([("a", "<u4"), ("b", "<u4"), ("L", "<f4"), ("k", "<f4"), ("d", "<f4"), ("max_pos", "<f4"),
                 ("max_neg", "<f4"), ("strength", "<f4"), ("plastic", "<f4"), ("stress", "<f4"), ("L0", "<f4"),
                 ("support_limit", "<f4"), ("type", "u1"), ("flags", "u1"), ("group", "<u2")])
TRIANGLE = np.dtype([("a", "<u4"), ("b", "<u4"), ("c", "<u4"), ("surface", "u1"), ("two_sided", "u1"),
                     ("torn", "u1"), ("pad", "u1"), ("rest_edge2", "<f4")])
CAPSULE = np.dtype([("a", "<u4"), ("b", "<u4"), ("radius", "<f4"), ("joint", "<i4")])
BOX = np.dtype([("center", "<f4", 3), ("rot", "<f4", 9), ("half", "<f4", 3), ("aabb_mn", "<f4", 3),
                ("aabb_mx", "<f4", 3), ("max_force", "<f4"), ("surface", "<u4")])
SHOCK = np.dtype([("beam", "<u4"), ("type", "<u4"), ("soft_bump", "<u4")] +
                 [(n, "<f4") for n in ("spring", "damp", "spring_in", "damp_in", "prog_spring_in", "prog_damp_in",
                                       "spring_out", "damp_out", "prog_spring_out", "prog_damp_out", "damp_in_slow",
                                       "split_in", "damp_in_fast", "damp_out_slow", "split_out", "damp_out_fast",
                                       "short_bound", "long_bound", "bound_spring", "bound_damp")])
SLIDE = np.dtype([("node", "<u4"), ("rail0", "<u4"), ("nrail", "<u4"), ("seg", "<i4"), ("k", "<f4"),
                  ("threshold", "<f4"), ("break_force", "<f4"), ("broken", "<u4")])
FRAME = np.dtype([("node", "<u4"), ("q", "<f4", 4), ("w", "<f4", 3), ("inv_inertia", "<f4"), ("ref_x", "<i4"),
                  ("ref_y", "<i4")])
JOINT = np.dtype([("parent_frame", "<u4"), ("child_node", "<u4"), ("child_frame", "<i4"), ("local_offset", "<f4", 3),
                  ("rel_rot", "<f4", 4)] +
                 [(n, "<f4") for n in ("k_lin", "d_lin", "k_ang", "d_ang", "break_force", "break_torque",
                                       "yield_angle", "max_stretch", "max_bend", "plastic_lin", "plastic_ang",
                                       "stress")] + [("broken", "<u4")])
WHEEL = np.dtype([("axle0", "<u4"), ("axle1", "<u4"), ("node0", "<u4"), ("n_nodes", "<u4"), ("radius", "<f4"),
                  ("propulsed", "<i4"), ("arm", "<i4"), ("near_attach", "<i4")] +
                 [(n, "<f4") for n in ("mass", "torque", "drive_torque", "brake", "speed", "avg_speed", "last_torque",
                                       "last_retorque")] + [("detached", "<u4"), ("ring", "<u4")])
assert WHEEL.itemsize == 72
CYLINDER = np.dtype([("base", "<f4", 3), ("radius", "<f4"), ("height", "<f4"), ("surface", "<u4")])
assert CYLINDER.itemsize == 24
assert NODE.itemsize == 32 and NODE_INFO.itemsize == 20 and BEAM.itemsize == 52
assert SHOCK.itemsize == 92 and SLIDE.itemsize == 32 and FRAME.itemsize == 44 and JOINT.itemsize == 92
assert TRIANGLE.itemsize == 20 and CAPSULE.itemsize == 16 and BOX.itemsize == 92

# NodeFlag (src/phys/softbody.h:29-40)
NF_GROUND, NF_FIXED, NF_TYRE, NF_RIM, NF_CONTACTER = 1, 2, 4, 8, 16
NF_FRAME, NF_LOAD, NF_NO_MOUSE, NF_SELF_COLL = 32, 64, 128, 256
# BeamType / BeamFlag (src/phys/softbody.h:60-69)
BT_NORMAL, BT_ROPE, BT_SUPPORT = 0, 1, 2
BF_BROKEN, BF_INVISIBLE, BF_NO_DEFORM, BF_NO_BREAK, BF_WHEEL, BF_HYDRO, BF_SHOCK = 1, 2, 4, 8, 16, 32, 64
# Surface (src/phys/static_world.h:59-62)
SURF_GRASS, SURF_DIRT, SURF_ASPHALT, SURF_CONCRETE, SURF_GRAVEL, SURF_MUD, SURF_SAND = range(7)
SURF_ROCK, SURF_WOOD, SURF_METAL, SURF_ICE = range(7, 11)

BL_OK, BL_ERR_ARG, BL_ERR_HIP, BL_ERR_UNSUPPORTED, BL_ERR_STATE = 0, -1, -2, -3, -4
MODE_AUTO, MODE_STEPPED, MODE_PERSISTENT, MODE_RESIDENT, MODE_HALO = 0, 1, 2, 3, 4


class Settings(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 3), ("dt", ctypes.c_float), ("collision_hz", ctypes.c_float),
                ("frame_theta", ctypes.c_float), ("frame_dissipation", ctypes.c_float),
                ("inter_body_collisions", ctypes.c_int32), ("sleeping", ctypes.c_int32), ("kill_y", ctypes.c_float),
                ("tyre_grip", ctypes.c_float)]


class BodyOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("collision_radius", "ground_friction", "bounce", "contact_push_max",
                                               "contact_slop", "hull_depth", "sphere_ball", "contact_friction")] + \
               [(n, ctypes.c_int32) for n in ("allow_break", "allow_deform", "ductile", "self_collision",
                                               "collision_group", "faces_only")] + \
               [("face_skin", ctypes.c_float), ("air_drag", ctypes.c_float), ("aero_cda", ctypes.c_float), ("rest_damp", ctypes.c_float),
                ("energy_guard", ctypes.c_int32)]


class BodyDesc(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("info", ctypes.c_void_p), ("n_nodes", ctypes.c_uint32),
                ("beams", ctypes.c_void_p), ("n_beams", ctypes.c_uint32), ("tris", ctypes.c_void_p),
                ("n_tris", ctypes.c_uint32), ("capsules", ctypes.c_void_p), ("n_capsules", ctypes.c_uint32),
                ("opts", BodyOpts), ("shocks", ctypes.c_void_p), ("n_shocks", ctypes.c_uint32),
                ("slides", ctypes.c_void_p), ("n_slides", ctypes.c_uint32), ("rail", ctypes.c_void_p),
                ("n_rail", ctypes.c_uint32), ("frames", ctypes.c_void_p), ("n_frames", ctypes.c_uint32),
                ("joints", ctypes.c_void_p), ("n_joints", ctypes.c_uint32), ("wheels", ctypes.c_void_p),
                ("n_wheels", ctypes.c_uint32), ("wheel_nodes", ctypes.c_void_p), ("n_wheel_nodes", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("substeps", "broken_beams", "static_contacts", "vel_clamps",
                                               "nan_resets", "pair_rebuilds", "contact_pairs", "narrow_tests",
                                               "contacts", "solve_failures")] + \
               [("aabb_mn", ctypes.c_float * 3), ("aabb_mx", ctypes.c_float * 3), ("max_speed", ctypes.c_float),
                ("pad", ctypes.c_float)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_ if n not in ("aabb_mn", "aabb_mx", "pad")}
        d["aabb_mn"] = list(self.aabb_mn)
        d["aabb_mx"] = list(self.aabb_mx)
        return d


class Tuning(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("block", ctypes.c_int32), ("substeps_per_launch", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("incidences_per_thread", ctypes.c_int32)]


class ShellSection(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("E", "nu", "rho", "t", "drill", "damping", "yield_")]


class FrameSection(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("E", "G", "rho", "A", "Iy", "Iz", "J", "As_y", "As_z", "damping",
                                               "axial", "joint_k", "Np", "Mp", "Tp", "hardening", "half")]


FRAME_SECTION_FIELDS = 17
FEM_TRI_STATE, FEM_MEM_STATE = 29, 27


FRAME_MEMBER = np.dtype([("a", np.uint32), ("b", np.uint32), ("section", np.uint16), ("end_a", np.uint8),
                         ("end_b", np.uint8)])


class FemDesc(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_uint32), ("n_tris", ctypes.c_uint32), ("p", ctypes.c_void_p),
                ("tris", ctypes.c_void_p), ("fixed", ctypes.c_void_p), ("mass", ctypes.c_void_p),
                ("section", ShellSection), ("n_members", ctypes.c_uint32), ("n_sections", ctypes.c_uint32),
                ("members", ctypes.c_void_p), ("sections", ctypes.c_void_p)]


class FemReport(ctypes.Structure):
    _fields_ = [("residual", ctypes.c_double), ("factor_flops", ctypes.c_double)] + \
               [(n, ctypes.c_float) for n in ("assemble_ms", "factor_ms", "solve_ms", "update_ms")] + \
               [(n, ctypes.c_int32) for n in ("failures", "supernodes", "fronts_max")] + [("factor_nnz", ctypes.c_int64)] + \
               [(n, ctypes.c_int32) for n in ("passes", "repasses")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


WARM_ZERO, WARM_PREVIOUS, WARM_NETWORK = 0, 1, 2
PRECOND_BLOCK_JACOBI, PRECOND_NETWORK, PRECOND_FACTOR = 0, 1, 2


class PcgOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("max_iter", "warm", "precond", "fallback", "check_error", "pad")] + \
               [("tol", ctypes.c_double)]


class PcgReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("iterations", "converged", "fallback_used", "pad")] + \
               [(n, ctypes.c_double) for n in ("residual", "warm_error", "pcg_error")] + \
               [(n, ctypes.c_float) for n in ("assemble_ms", "gnn_ms", "pcg_ms", "fallback_ms", "update_ms", "total_ms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


# every entry point the header declares, with its argument types
_P = ctypes.c_void_p
SIGNATURES = {
    "bl_finalize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, _P, ctypes.c_uint32]),
    "bl_stabilize": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.POINTER(ctypes.c_int32)]),
    "bl_box_from_quat": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
    "bl_default_settings": (None, [ctypes.POINTER(Settings)]),
    "bl_default_body_opts": (None, [ctypes.POINTER(BodyOpts)]),
    "bl_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "bl_abi_version": (ctypes.c_int, []),
    "bl_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(Settings)]),
    "bl_destroy": (None, [_P]),
    "bl_set_statics": (ctypes.c_int, [_P, _P, ctypes.c_int32]),
    "bl_add_body": (ctypes.c_int, [_P, ctypes.POINTER(BodyDesc), ctypes.POINTER(ctypes.c_uint32)]),
    "bl_step": (ctypes.c_int, [_P, ctypes.c_int32]),
    "bl_step_async": (ctypes.c_int, [_P, ctypes.c_int32]),
    "bl_sync": (ctypes.c_int, [_P]),
    "bl_read_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
    "bl_write_body": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
    "bl_body_broken_beams": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int64)]),
    "bl_read_slides": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "bl_read_joints": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
    "bl_read_wheels": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
    "bl_set_cylinders": (ctypes.c_int, [_P, _P, ctypes.c_int32]),
    "bl_set_terrain": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, _P, _P, _P]),
    "bl_set_wheel_controls": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
    "bl_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "bl_reset_stats": (ctypes.c_int, [_P]),
    "bl_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "bl_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "bl_time_step": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "bl_time_kernels": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)]),
    "bl_fem_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(FemDesc)]),
    "bl_fem_destroy": (None, [_P]),
    "bl_fem_step": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P,
                                   ctypes.POINTER(FemReport)]),
    "bl_fem_get_state": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "bl_fem_set_state": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "bl_fem_get_nodes": (ctypes.c_int, [_P, _P, _P, _P]),
    "bl_fem_add_torques": (ctypes.c_int, [_P, _P]),
    "bl_fem_solve_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "bl_fem_last_system": (ctypes.c_int, [_P, _P, _P, _P, ctypes.POINTER(ctypes.c_uint32)]),
    "bl_fem_solve": (ctypes.c_int, [_P, _P, _P]),
    "bl_fem_tri_k": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "bl_fem_get_plastic": (ctypes.c_int, [_P, _P, _P, _P]),
    "bl_fem_step_span": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P,
                                        ctypes.POINTER(FemReport)]),
    "bl_fem_step_held": (ctypes.c_int, [_P, _P, ctypes.c_float, _P]),
    "bl_gnn_param_count": (ctypes.c_int, []),
    "bl_gnn_create": (ctypes.c_int, [ctypes.POINTER(_P), _P, ctypes.c_uint32]),
    "bl_gnn_destroy": (None, [_P]),
    "bl_gnn_get_params": (ctypes.c_int, [_P, _P, ctypes.c_uint32]),
    "bl_gnn_set_params": (ctypes.c_int, [_P, _P, ctypes.c_uint32]),
    "bl_gnn_gradients": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
    "bl_gnn_train_step": (ctypes.c_int, [_P, _P, ctypes.c_float, _P]),
    "bl_gnn_infer": (ctypes.c_int, [_P, _P, _P, _P]),
    "bl_fem_step_pcg": (ctypes.c_int, [_P, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P,
                                       ctypes.POINTER(PcgOpts), ctypes.POINTER(PcgReport)]),
}

_lib = None


class BeamlabError(RuntimeError):
    """A non-zero status of an entry point (status: the BL_ERR_* code, None when the library is missing)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


_live = {}  # objects holding library handles (World, ShellBody, Network), in creation order: closed at exit


def track(obj):
    """Registers an object whose close() releases library handles: the interpreter's exit closes it, newest first,
    before the HIP runtime (and a profiler attached to it) tear down."""
    _live[id(obj)] = weakref.ref(obj)


def _close_all():
    for ref in reversed(list(_live.values())):
        obj = ref()
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass
    _live.clear()


def load():
    """Loads libbeamlab_gpu.so (built by beamlab_amd.build); raises if it is missing - there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BeamlabError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    atexit.register(_close_all)
    return lib


def check(rc, what=""):
    if rc != BL_OK:
        msg = load().bl_error_string(rc).decode()
        raise BeamlabError(f"{what}: status {rc}: {msg}", rc)
    return rc


def ptr(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(ctypes.c_void_p)


def default_settings():
    s = Settings()
    s.gravity[:] = (0.0, -9.81, 0.0)
    s.dt = 0.0005
    s.collision_hz = 120.0
    s.frame_theta = 0.5
    s.frame_dissipation = 0.02
    s.inter_body_collisions = 1
    s.sleeping = 0
    s.kill_y = -200.0
    s.tyre_grip = 1.0
    return s


def default_body_opts():
    o = BodyOpts()
    o.collision_radius = 0.05
    o.ground_friction = 1.0
    o.bounce = -1.0
    o.contact_push_max = 2.0
    o.contact_slop = 0.0
    o.hull_depth = 0.3
    o.face_skin = 0.002
    o.contact_friction = 1.0
    o.allow_break = 1
    o.allow_deform = 1
    return o
