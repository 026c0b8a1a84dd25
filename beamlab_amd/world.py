"""Host-side mirror of the reference's physics interface over the C ABI.

Names and meaning follow src/phys: SoftBody's add_node / add_beam / add_triangle / finalize / stabilize
(src/phys/softbody.cpp:8-110), StaticWorld::add_box (src/phys/static_world.cpp:298-306) and World's add_body /
step_substeps (src/phys/world.h:160-169). Bodies hold their arrays as numpy records laid out like the reference's
Node / NodeInfo / Beam / Triangle / Capsule, so they go through the C ABI unchanged.
"""
import ctypes

import numpy as np

from . import abi


class SoftBody:
    """A node/beam body under construction (the reference's SoftBody builder methods)."""

    def __init__(self, name=""):
        self.name = name
        self.nodes = np.zeros(0, abi.NODE)
        self.info = np.zeros(0, abi.NODE_INFO)
        self.beams = np.zeros(0, abi.BEAM)
        self.tris = np.zeros(0, abi.TRIANGLE)
        self.capsules = np.zeros(0, abi.CAPSULE)
        self.shocks = np.zeros(0, abi.SHOCK)
        self.slides = np.zeros(0, abi.SLIDE)
        self.rail = np.zeros(0, np.uint32)
        self.frames = np.zeros(0, abi.FRAME)
        self.joints = np.zeros(0, abi.JOINT)
        self.wheels = np.zeros(0, abi.WHEEL)
        self.wheel_nodes = np.zeros(0, np.uint32)
        self.opts = abi.default_body_opts()
        self._pending_nodes, self._pending_beams, self._pending_tris = [], [], []

    # SoftBody::add_node (src/phys/softbody.cpp:8-20)
    def add_node(self, p, mass, flags=abi.NF_GROUND | abi.NF_CONTACTER):
        m = np.float32(mass)
        inv = np.float32(0.0) if (flags & abi.NF_FIXED) or m <= 0 else np.float32(1.0) / m
        self._pending_nodes.append((tuple(np.float32(x) for x in p), inv, m, flags))
        return len(self.nodes) + len(self._pending_nodes) - 1

    # SoftBody::add_beam (src/phys/softbody.cpp:22-40); L and L0 are set by finalize
    def add_beam(self, a, b, k, d, strength, deform, type_=abi.BT_NORMAL, flags=0):
        self._pending_beams.append((a, b, k, d, strength, deform, type_, flags))
        return len(self.beams) + len(self._pending_beams) - 1

    def add_triangle(self, a, b, c, two_sided=True):
        self._pending_tris.append((a, b, c, two_sided))
        return len(self.tris) + len(self._pending_tris) - 1

    # a shock on a beam (the reference's truck loader makes them: Shock, src/phys/softbody.h:88-101); the beam is
    # flagged BF_SHOCK, so that the beam loop leaves it to the shock loop. Unnamed fields keep the reference's defaults.
    def add_shock(self, beam, type_=1, **fields):
        self._flush()
        s = np.zeros(1, abi.SHOCK)
        for n, v in dict(damp_in_slow=1, split_in=1, damp_in_fast=1, damp_out_slow=1, split_out=1, damp_out_fast=1,
                         bound_spring=9e6, bound_damp=12000).items():
            s[0][n] = v
        s[0]["beam"], s[0]["type"] = beam, type_
        for n, v in fields.items():
            s[0][n] = v
        self.beams[beam]["flags"] |= abi.BF_SHOCK
        self.shocks = np.concatenate([self.shocks, s])
        return len(self.shocks) - 1

    # a slide node on a rail of nodes (SlideNode, src/phys/softbody.h:559-567)
    def add_slide(self, node, rail, k=9e6, threshold=0.0, break_force=0.0):
        s = np.zeros(1, abi.SLIDE)
        s[0]["node"], s[0]["rail0"], s[0]["nrail"], s[0]["seg"] = node, len(self.rail), len(rail), -1
        s[0]["k"], s[0]["threshold"], s[0]["break_force"] = k, threshold, break_force
        self.rail = np.concatenate([self.rail, np.asarray(rail, np.uint32)])
        self.slides = np.concatenate([self.slides, s])
        return len(self.slides) - 1

    # SoftBody::add_frame (src/phys/softbody.cpp:42-53): an oriented node (q = x, y, z, w)
    def add_frame(self, node, q=(0.0, 0.0, 0.0, 1.0), inertia=1.0):
        self._flush()
        f = np.zeros(1, abi.FRAME)
        f[0]["node"], f[0]["q"], f[0]["ref_x"], f[0]["ref_y"] = node, q, -1, -1
        f[0]["inv_inertia"] = np.float32(1.0) / np.float32(inertia) if inertia > 0 and self.nodes[node]["inv_mass"] > 0 else 0
        self.info[node]["flags"] |= abi.NF_FRAME
        self.info[node]["frame"] = len(self.frames)
        self.frames = np.concatenate([self.frames, f])
        return len(self.frames) - 1

    # SoftBody::add_joint (src/phys/softbody.cpp:55-75): the child's rest offset and orientation in the parent frame
    def add_joint(self, parent_frame, child_node, child_frame, k_lin, d_lin, k_ang, d_ang, **fields):
        self._flush()
        pf = self.frames[parent_frame]
        j = np.zeros(1, abi.JOINT)
        j[0]["parent_frame"], j[0]["child_node"], j[0]["child_frame"] = parent_frame, child_node, child_frame
        qc = _qconj(pf["q"])
        j[0]["local_offset"] = _qrotate(qc, self.nodes[child_node]["p"] - self.nodes[pf["node"]]["p"])
        j[0]["rel_rot"] = _qnormalize(_qmul(qc, self.frames[child_frame]["q"])) if child_frame >= 0 else (0, 0, 0, 1)
        j[0]["k_lin"], j[0]["d_lin"], j[0]["k_ang"], j[0]["d_ang"] = k_lin, d_lin, k_ang, d_ang
        j[0]["max_stretch"], j[0]["max_bend"] = 0.12, 0.8
        for n, v in fields.items():
            j[0][n] = v
        self.joints = np.concatenate([self.joints, j])
        return len(self.joints) - 1

    # a wheel (Wheel, src/phys/softbody.h:468-490; the vehicle builder fills it): its hub's two axle nodes and its
    # tyre nodes (even: the axle0 side); mass defaults to the tyre nodes' masses
    def add_wheel(self, axle0, axle1, nodes, radius=0.4, propulsed=0, arm=-1, near_attach=-1, mass=None):
        self._flush()
        w = np.zeros(1, abi.WHEEL)
        w[0]["axle0"], w[0]["axle1"], w[0]["node0"], w[0]["n_nodes"] = axle0, axle1, len(self.wheel_nodes), len(nodes)
        w[0]["radius"], w[0]["propulsed"], w[0]["arm"], w[0]["near_attach"] = radius, propulsed, arm, near_attach
        w[0]["mass"] = np.float32(sum(self.nodes[n]["mass"] for n in nodes)) if mass is None else mass
        self.wheel_nodes = np.concatenate([self.wheel_nodes, np.asarray(nodes, np.uint32)])
        self.wheels = np.concatenate([self.wheels, w])
        return len(self.wheels) - 1

    def _flush(self):
        if self._pending_nodes:
            n = np.zeros(len(self._pending_nodes), abi.NODE)
            inf = np.zeros(len(n), abi.NODE_INFO)
            for i, (p, inv, m, fl) in enumerate(self._pending_nodes):
                n[i]["p"] = p
                n[i]["inv_mass"] = inv
                n[i]["mass"] = m
                inf[i]["flags"] = fl
            inf["friction"] = 1.0
            inf["frame"] = -1
            inf["ror_id"] = -1
            self.nodes = np.concatenate([self.nodes, n])
            self.info = np.concatenate([self.info, inf])
            self._pending_nodes = []
        if self._pending_beams:
            b = np.zeros(len(self._pending_beams), abi.BEAM)
            for j, (a, bb, k, d, s, df, t, fl) in enumerate(self._pending_beams):
                b[j]["a"], b[j]["b"], b[j]["k"], b[j]["d"] = a, bb, k, d
                b[j]["strength"], b[j]["max_pos"], b[j]["max_neg"] = s, df, -np.float32(df)
                b[j]["type"], b[j]["flags"], b[j]["support_limit"] = t, fl, 4.0
            self.beams = np.concatenate([self.beams, b])
            self._pending_beams = []
        if self._pending_tris:
            t = np.zeros(len(self._pending_tris), abi.TRIANGLE)
            for j, (a, b, c, ts) in enumerate(self._pending_tris):
                t[j]["a"], t[j]["b"], t[j]["c"], t[j]["two_sided"] = a, b, c, int(ts)
            self.tris = np.concatenate([self.tris, t])
            self._pending_tris = []

    # SoftBody::finalize (src/phys/softbody.cpp:76-83)
    def finalize(self):
        self._flush()
        lib = abi.load()
        abi.check(lib.bl_finalize(abi.ptr(self.nodes), len(self.nodes), abi.ptr(self.beams), len(self.beams),
                                  abi.ptr(self.tris), len(self.tris)), "bl_finalize")

    # SoftBody::stabilize (src/phys/softbody.cpp:85-110); returns the number of beams scaled
    def stabilize(self, dt, k_budget=0.6, d_budget=0.5):
        self._flush()
        lib = abi.load()
        changed = ctypes.c_int32(0)
        abi.check(lib.bl_stabilize(abi.ptr(self.nodes), len(self.nodes), abi.ptr(self.beams), len(self.beams),
                                   dt, k_budget, d_budget, ctypes.byref(changed)), "bl_stabilize")
        return changed.value

    def desc(self):
        """The bl_body_desc of this body (it points into the body's arrays: keep the body alive)."""
        self._flush()
        d = abi.BodyDesc()
        d.nodes, d.info, d.n_nodes = abi.ptr(self.nodes), abi.ptr(self.info), len(self.nodes)
        d.beams, d.n_beams = abi.ptr(self.beams), len(self.beams)
        d.tris, d.n_tris = abi.ptr(self.tris), len(self.tris)
        d.capsules, d.n_capsules = abi.ptr(self.capsules), len(self.capsules)
        d.opts = self.opts
        d.shocks, d.n_shocks = abi.ptr(self.shocks), len(self.shocks)
        d.slides, d.n_slides = abi.ptr(self.slides), len(self.slides)
        d.rail, d.n_rail = abi.ptr(self.rail), len(self.rail)
        d.frames, d.n_frames = abi.ptr(self.frames), len(self.frames)
        d.joints, d.n_joints = abi.ptr(self.joints), len(self.joints)
        d.wheels, d.n_wheels = abi.ptr(self.wheels), len(self.wheels)
        d.wheel_nodes, d.n_wheel_nodes = abi.ptr(self.wheel_nodes), len(self.wheel_nodes)
        return d


# quaternions (x, y, z, w) in float32, as src/core/math.h:131-170 writes them (for the joints' rest values)
def _qmul(a, b):
    f = np.float32
    ax, ay, az, aw = (f(v) for v in a)
    bx, by, bz, bw = (f(v) for v in b)
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz], np.float32)


def _qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]], np.float32)


def _qnormalize(q):
    q = np.asarray(q, np.float32)
    n = np.float32(np.sqrt(np.float32(q @ q)))
    return q * (np.float32(1.0) / n) if n > 1e-12 else np.array([0, 0, 0, 1], np.float32)


def _qrotate(q, p):
    u = np.asarray(q[:3], np.float32)
    p = np.asarray(p, np.float32)
    t = np.cross(u, p).astype(np.float32) * np.float32(2.0)
    return (p + t * np.float32(q[3]) + np.cross(u, t)).astype(np.float32)


def make_box(center, half, quat=(0.0, 0.0, 0.0, 1.0), surface=abi.SURF_CONCRETE):
    """StaticWorld::add_box's record (src/phys/static_world.cpp:298-306)."""
    out = np.zeros(1, abi.BOX)
    c = np.asarray(center, np.float32)
    h = np.asarray(half, np.float32)
    q = np.asarray(quat, np.float32)
    abi.check(abi.load().bl_box_from_quat(abi.ptr(c), abi.ptr(h), abi.ptr(q), surface, abi.ptr(out)),
              "bl_box_from_quat")
    return out[0]


class StaticWorld:
    def __init__(self):
        self.boxes = np.zeros(0, abi.BOX)

    def add_box(self, center, half, quat=(0.0, 0.0, 0.0, 1.0), surface=abi.SURF_CONCRETE):
        self.boxes = np.concatenate([self.boxes, np.array([make_box(center, half, quat, surface)], abi.BOX)])


class World:
    """The GPU world: World::add_body / step_substeps over bl_create / bl_add_body / bl_step."""

    def __init__(self, settings=None):
        self.settings = settings or abi.default_settings()
        self.statics = StaticWorld()
        self.bodies = []
        self._ctx = None
        self._statics_sent = None
        abi.track(self)

    def _lib(self):
        return abi.load()

    def _context(self):
        if self._ctx is None:
            h = ctypes.c_void_p()
            abi.check(self._lib().bl_create(ctypes.byref(h), ctypes.byref(self.settings)), "bl_create")
            self._ctx = h
        return self._ctx

    def add_body(self, body):
        ctx = self._context()
        bid = ctypes.c_uint32(0)
        d = body.desc()
        abi.check(self._lib().bl_add_body(ctx, ctypes.byref(d), ctypes.byref(bid)), "bl_add_body")
        self.bodies.append(body)
        return bid.value

    def set_cylinders(self, cyl):
        """The static world's vertical cylinders (abi.CYLINDER records: StaticWorld::cylinders)."""
        self._cyl = np.ascontiguousarray(cyl)
        abi.check(self._lib().bl_set_cylinders(self._context(), abi.ptr(self._cyl), len(self._cyl)), "bl_set_cylinders")

    def set_terrain(self, heights, cell, origin, surfaces=None):
        """The static world's heightfield (StaticWorld::terrain): heights[z, x] (-1e30: no ground), `cell` apart from
        origin (x, z), surfaces[z, x] (None: grass); None removes it."""
        if heights is None:
            abi.check(self._lib().bl_set_terrain(self._context(), 0, 0, 0.0, None, None, None), "bl_set_terrain")
            return
        h = np.ascontiguousarray(heights, np.float32)
        o = np.asarray(origin, np.float32)
        sf = None if surfaces is None else np.ascontiguousarray(surfaces, np.uint8)
        abi.check(self._lib().bl_set_terrain(self._context(), h.shape[1], h.shape[0], float(cell), abi.ptr(o), abi.ptr(h),
                                             abi.ptr(sf)), "bl_set_terrain")

    def _sync_statics(self):
        boxes = self.statics.boxes
        key = boxes.tobytes()
        if key != self._statics_sent:
            abi.check(self._lib().bl_set_statics(self._context(), abi.ptr(boxes), len(boxes)), "bl_set_statics")
            self._statics_sent = key

    def set_tuning(self, mode=abi.MODE_AUTO, block=256, substeps_per_launch=0, use_graph=1, incidences_per_thread=4):
        t = abi.Tuning(mode, block, substeps_per_launch, use_graph, incidences_per_thread)
        abi.check(self._lib().bl_set_tuning(self._context(), ctypes.byref(t)), "bl_set_tuning")

    def get_tuning(self):
        t = abi.Tuning()
        abi.check(self._lib().bl_get_tuning(self._context(), ctypes.byref(t)), "bl_get_tuning")
        return t

    def step_substeps(self, n):
        self._sync_statics()
        abi.check(self._lib().bl_step(self._context(), int(n)), "bl_step")

    def time_substeps(self, n):
        self._sync_statics()
        ms = ctypes.c_float(0)
        abi.check(self._lib().bl_time_step(self._context(), int(n), ctypes.byref(ms)), "bl_time_step")
        return ms.value

    def time_kernels(self, n):
        """(frame ms, summed substep-kernel ms, launches) of one frame of n substeps, events around each launch."""
        self._sync_statics()
        tot, ker, nl = ctypes.c_float(0), ctypes.c_float(0), ctypes.c_int32(0)
        abi.check(self._lib().bl_time_kernels(self._context(), int(n), ctypes.byref(tot), ctypes.byref(ker),
                                              ctypes.byref(nl)), "bl_time_kernels")
        return tot.value, ker.value, nl.value

    def read_body(self, bid):
        b = self.bodies[bid]
        nodes = np.zeros(len(b.nodes), abi.NODE)
        beams = np.zeros(len(b.beams), abi.BEAM)
        abi.check(self._lib().bl_read_body(self._context(), bid, abi.ptr(nodes), abi.ptr(beams)), "bl_read_body")
        return nodes, beams

    def write_body(self, bid, nodes=None, beams=None):
        abi.check(self._lib().bl_write_body(self._context(), bid, abi.ptr(nodes), abi.ptr(beams)), "bl_write_body")

    def read_joints(self, bid):
        fr, jt = self.bodies[bid].frames.copy(), self.bodies[bid].joints.copy()
        abi.check(self._lib().bl_read_joints(self._context(), bid, abi.ptr(fr), abi.ptr(jt)), "bl_read_joints")
        return fr, jt

    def read_wheels(self, bid):
        """(the wheels as the step left them, [wheel_speed, wheel_spin])"""
        wh, ss = self.bodies[bid].wheels.copy(), np.zeros(2, np.float32)
        abi.check(self._lib().bl_read_wheels(self._context(), bid, abi.ptr(wh), abi.ptr(ss)), "bl_read_wheels")
        return wh, ss

    def set_wheel_controls(self, bid, torque=None, brake=None):
        """A controller setting each wheel's torque and brake at every substep from the next step on (None: none)."""
        if torque is None or brake is None:
            abi.check(self._lib().bl_set_wheel_controls(self._context(), bid, None, None), "bl_set_wheel_controls")
            return
        self._ctl = (np.ascontiguousarray(torque, np.float32), np.ascontiguousarray(brake, np.float32))
        abi.check(self._lib().bl_set_wheel_controls(self._context(), bid, abi.ptr(self._ctl[0]), abi.ptr(self._ctl[1])),
                  "bl_set_wheel_controls")

    def read_slides(self, bid):
        out = self.bodies[bid].slides.copy()
        abi.check(self._lib().bl_read_slides(self._context(), bid, abi.ptr(out)), "bl_read_slides")
        return out

    def broken_beams(self, bid):
        v = ctypes.c_int64(0)
        abi.check(self._lib().bl_body_broken_beams(self._context(), bid, ctypes.byref(v)), "bl_body_broken_beams")
        return v.value

    def stats(self):
        s = abi.Stats()
        abi.check(self._lib().bl_get_stats(self._context(), ctypes.byref(s)), "bl_get_stats")
        return s.as_dict()

    def close(self):
        if self._ctx is not None:
            self._lib().bl_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
