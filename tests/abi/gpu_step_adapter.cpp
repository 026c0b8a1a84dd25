// The reference-side adapter of INTEGRATION.md, as a maintainer would add it to the reference tree (e.g.
// src/phys/gpu_step.h): it hands the bodies the GPU can step to the drop-in through the C ABI and keeps the rest on
// the CPU. tests/test_abi_client.py compiles it against the reference's own headers where they lie (syntax and
// types: the records' layouts are pinned by oracle/ref_capi.cpp's static asserts).
#include "beamlab_gpu.h"
#include "phys/world.h"

#include <stdexcept>
#include <vector>

namespace bl::phys {

// Steps the eligible bodies of a World on the GPU; the others stay with World::step_substeps.
class GpuStep {
public:
    explicit GpuStep(const World& w) {
        bl_settings s;
        bl_default_settings(&s);
        s.gravity[0] = w.settings.gravity.x, s.gravity[1] = w.settings.gravity.y, s.gravity[2] = w.settings.gravity.z;
        s.dt = w.settings.dt; // WorldSettings (src/phys/world.h:24-60)
        s.frame_theta = w.settings.frame_theta;
        s.frame_dissipation = w.settings.frame_dissipation;
        s.sleeping = 0; // sleeping bodies are not stepped on the GPU
        s.tyre_grip = w.settings.tyre_grip;
        check(bl_create(&ctx_, &s));
    }
    ~GpuStep() { bl_destroy(ctx_); }
    GpuStep(const GpuStep&) = delete;
    GpuStep& operator=(const GpuStep&) = delete;

    void set_statics(const std::vector<bl_box>& boxes) { // StaticWorld boxes (src/phys/static_world.cpp:298-306)
        check(bl_set_statics(ctx_, boxes.data(), (int32_t)boxes.size()));
    }
    // StaticWorld's cylinders and terrain (a Heightfield: its vertices row by row, holes as they read)
    void set_terrain_and_cylinders(const StaticWorld& sw) {
        std::vector<bl_cylinder> cyl;
        for (const StaticCylinder& c : sw.cylinders)
            cyl.push_back(bl_cylinder{{c.base.x, c.base.y, c.base.z}, c.radius, c.height, c.surface});
        check(bl_set_cylinders(ctx_, cyl.data(), (int32_t)cyl.size()));
        if (!sw.has_terrain) {
            check(bl_set_terrain(ctx_, 0, 0, 0.0f, nullptr, nullptr, nullptr));
            return;
        }
        const Heightfield& t = sw.terrain;
        std::vector<float> h((size_t)t.nx() * t.nz());
        std::vector<uint8_t> sf(h.size());
        for (int z = 0; z < t.nz(); z++)
            for (int x = 0; x < t.nx(); x++) h[(size_t)z * t.nx() + x] = t.h(x, z), sf[(size_t)z * t.nx() + x] = t.surf(x, z);
        const float origin[2] = {t.origin().x, t.origin().y};
        check(bl_set_terrain(ctx_, t.nx(), t.nz(), t.cell(), origin, h.data(), sf.data()));
    }
    // World::add_body (src/phys/world.cpp:280-300): the body's arrays as they are, its options from its fields
    uint32_t add(SoftBody& b) {
        static_assert(sizeof(Node) == sizeof(bl_node) && sizeof(Beam) == sizeof(bl_beam) &&
                          sizeof(Triangle) == sizeof(bl_triangle) && sizeof(Capsule) == sizeof(bl_capsule),
                      "the records pass unchanged");
        bl_body_desc d{};
        d.nodes = reinterpret_cast<const bl_node*>(b.nodes.data());
        d.info = reinterpret_cast<const bl_node_info*>(b.info.data());
        d.n_nodes = (uint32_t)b.nodes.size();
        d.beams = reinterpret_cast<const bl_beam*>(b.beams.data());
        d.n_beams = (uint32_t)b.beams.size();
        d.tris = reinterpret_cast<const bl_triangle*>(b.tris.data());
        d.n_tris = (uint32_t)b.tris.size();
        d.capsules = reinterpret_cast<const bl_capsule*>(b.capsules.data());
        d.n_capsules = (uint32_t)b.capsules.size();
        bl_default_body_opts(&d.opts);
        d.opts.collision_radius = b.collision_radius, d.opts.ground_friction = b.ground_friction;
        d.opts.bounce = b.bounce, d.opts.contact_push_max = b.contact_push_max, d.opts.contact_slop = b.contact_slop;
        d.opts.hull_depth = b.hull_depth, d.opts.sphere_ball = b.sphere_ball, d.opts.contact_friction = b.contact_friction;
        d.opts.allow_break = b.allow_break, d.opts.allow_deform = b.allow_deform, d.opts.ductile = b.ductile;
        d.opts.self_collision = b.self_collision, d.opts.collision_group = b.collision_group;
        d.opts.faces_only = b.faces_only, d.opts.face_skin = b.face_skin;
        d.opts.air_drag = b.air_drag, d.opts.aero_cda = b.aero_cda, d.opts.rest_damp = b.rest_damp;
        d.opts.energy_guard = b.energy_guard;
        // its shocks and slide nodes (Shock, SlideNode: src/phys/softbody.h:88-101,559-567), field by field
        std::vector<bl_shock> shocks;
        for (const Shock& s : b.shocks) {
            bl_shock x{};
            x.beam = s.beam, x.type = s.type, x.soft_bump = s.soft_bump ? 1u : 0u;
            x.spring = s.spring, x.damp = s.damp;
            x.spring_in = s.spring_in, x.damp_in = s.damp_in, x.prog_spring_in = s.prog_spring_in;
            x.prog_damp_in = s.prog_damp_in;
            x.spring_out = s.spring_out, x.damp_out = s.damp_out, x.prog_spring_out = s.prog_spring_out;
            x.prog_damp_out = s.prog_damp_out;
            x.damp_in_slow = s.damp_in_slow, x.split_in = s.split_in, x.damp_in_fast = s.damp_in_fast;
            x.damp_out_slow = s.damp_out_slow, x.split_out = s.split_out, x.damp_out_fast = s.damp_out_fast;
            x.short_bound = s.short_bound, x.long_bound = s.long_bound;
            x.bound_spring = s.bound_spring, x.bound_damp = s.bound_damp;
            shocks.push_back(x);
        }
        std::vector<bl_slide> slides;
        std::vector<uint32_t> rail;
        for (const SlideNode& s : b.slides) {
            bl_slide x{};
            x.node = s.node, x.rail0 = (uint32_t)rail.size(), x.nrail = (uint32_t)s.rail.size(), x.seg = s.seg;
            x.k = s.k, x.threshold = s.threshold, x.break_force = s.break_force, x.broken = s.broken ? 1u : 0u;
            rail.insert(rail.end(), s.rail.begin(), s.rail.end());
            slides.push_back(x);
        }
        // its frames and joints (Frame, Joint: src/phys/softbody.h:104-136)
        std::vector<bl_frame> frames;
        for (const Frame& f : b.frames) {
            bl_frame x{};
            x.node = f.node, x.q[0] = f.q.x, x.q[1] = f.q.y, x.q[2] = f.q.z, x.q[3] = f.q.w;
            x.w[0] = f.w.x, x.w[1] = f.w.y, x.w[2] = f.w.z, x.inv_inertia = f.inv_inertia;
            x.ref_x = f.ref_x, x.ref_y = f.ref_y;
            frames.push_back(x);
        }
        std::vector<bl_joint> joints;
        for (const Joint& j : b.joints) {
            bl_joint x{};
            x.parent_frame = j.parent_frame, x.child_node = j.child_node, x.child_frame = j.child_frame;
            x.local_offset[0] = j.local_offset.x, x.local_offset[1] = j.local_offset.y, x.local_offset[2] = j.local_offset.z;
            x.rel_rot[0] = j.rel_rot.x, x.rel_rot[1] = j.rel_rot.y, x.rel_rot[2] = j.rel_rot.z, x.rel_rot[3] = j.rel_rot.w;
            x.k_lin = j.k_lin, x.d_lin = j.d_lin, x.k_ang = j.k_ang, x.d_ang = j.d_ang;
            x.break_force = j.break_force, x.break_torque = j.break_torque, x.yield_angle = j.yield_angle;
            x.max_stretch = j.max_stretch, x.max_bend = j.max_bend;
            x.plastic_lin = j.plastic_lin, x.plastic_ang = j.plastic_ang, x.stress = j.stress;
            x.broken = j.broken ? 1u : 0u;
            joints.push_back(x);
        }
        // its wheels (Wheel: src/phys/softbody.h:468-490), their tyre nodes in one list
        std::vector<bl_wheel> wheels;
        std::vector<uint32_t> wnodes;
        for (const Wheel& w : b.wheels) {
            bl_wheel x{};
            x.axle0 = w.axle0, x.axle1 = w.axle1, x.node0 = (uint32_t)wnodes.size(), x.n_nodes = (uint32_t)w.nodes.size();
            x.radius = w.radius, x.propulsed = w.propulsed, x.arm = w.arm, x.near_attach = w.near_attach, x.mass = w.mass;
            x.torque = w.torque, x.drive_torque = w.drive_torque, x.brake = w.brake;
            x.speed = w.speed, x.avg_speed = w.avg_speed, x.last_torque = w.last_torque, x.last_retorque = w.last_retorque;
            x.detached = w.detached ? 1u : 0u, x.ring = w.ring ? 1u : 0u;
            wnodes.insert(wnodes.end(), w.nodes.begin(), w.nodes.end());
            wheels.push_back(x);
        }
        d.wheels = wheels.data(), d.n_wheels = (uint32_t)wheels.size();
        d.wheel_nodes = wnodes.data(), d.n_wheel_nodes = (uint32_t)wnodes.size();
        d.frames = frames.data(), d.n_frames = (uint32_t)frames.size();
        d.joints = joints.data(), d.n_joints = (uint32_t)joints.size();
        d.shocks = shocks.data(), d.n_shocks = (uint32_t)shocks.size();
        d.slides = slides.data(), d.n_slides = (uint32_t)slides.size();
        d.rail = rail.data(), d.n_rail = (uint32_t)rail.size();
        uint32_t id;
        const int rc = bl_add_body(ctx_, &d, &id);
        if (rc == BL_ERR_UNSUPPORTED) throw Unsupported(bl_error_string(rc)); // keep this body on the CPU
        check(rc);
        bodies_.push_back(&b);
        return id;
    }
    // A vehicle's controller (its pre_substep hook's torque and brake per wheel) for the substeps of the next steps
    void set_wheel_controls(uint32_t id, const std::vector<float>& torque, const std::vector<float>& brake) {
        check(bl_set_wheel_controls(ctx_, id, torque.data(), brake.data()));
    }
    // World::step_substeps for the GPU bodies; their arrays (and slide nodes', frames', joints' and wheels' state) read back
    void step_substeps(int n) {
        check(bl_step(ctx_, n));
        for (uint32_t i = 0; i < bodies_.size(); i++) {
            SoftBody& b = *bodies_[i];
            check(bl_read_body(ctx_, i, reinterpret_cast<bl_node*>(b.nodes.data()), reinterpret_cast<bl_beam*>(b.beams.data())));
            if (!b.frames.empty() || !b.joints.empty()) {
                std::vector<bl_frame> fr(b.frames.size());
                std::vector<bl_joint> jt(b.joints.size());
                check(bl_read_joints(ctx_, i, fr.data(), jt.data()));
                for (size_t q = 0; q < fr.size(); q++) {
                    Frame& f = b.frames[q];
                    f.q = quat(fr[q].q[0], fr[q].q[1], fr[q].q[2], fr[q].q[3]), f.w = vec3(fr[q].w[0], fr[q].w[1], fr[q].w[2]);
                }
                for (size_t q = 0; q < jt.size(); q++) {
                    Joint& j = b.joints[q];
                    j.local_offset = vec3(jt[q].local_offset[0], jt[q].local_offset[1], jt[q].local_offset[2]);
                    j.rel_rot = quat(jt[q].rel_rot[0], jt[q].rel_rot[1], jt[q].rel_rot[2], jt[q].rel_rot[3]);
                    j.plastic_lin = jt[q].plastic_lin, j.plastic_ang = jt[q].plastic_ang, j.stress = jt[q].stress;
                    j.broken = jt[q].broken != 0;
                }
            }
            if (!b.wheels.empty()) {
                std::vector<bl_wheel> wh(b.wheels.size());
                float ss[2];
                check(bl_read_wheels(ctx_, i, wh.data(), ss));
                for (size_t q = 0; q < wh.size(); q++) {
                    Wheel& w = b.wheels[q];
                    w.torque = wh[q].torque, w.drive_torque = wh[q].drive_torque, w.brake = wh[q].brake;
                    w.speed = wh[q].speed, w.avg_speed = wh[q].avg_speed;
                    w.last_torque = wh[q].last_torque, w.last_retorque = wh[q].last_retorque;
                }
                b.wheel_speed = ss[0], b.wheel_spin = ss[1];
            }
            if (b.slides.empty()) continue;
            std::vector<bl_slide> sl(b.slides.size());
            check(bl_read_slides(ctx_, i, sl.data()));
            for (size_t q = 0; q < sl.size(); q++) b.slides[q].seg = sl[q].seg, b.slides[q].broken = sl[q].broken != 0;
        }
    }

    struct Unsupported : std::runtime_error {
        using std::runtime_error::runtime_error;
    };

private:
    static void check(int rc) {
        if (rc != BL_OK) throw std::runtime_error(bl_error_string(rc));
    }
    bl_ctx* ctx_ = nullptr;
    std::vector<SoftBody*> bodies_;
};

// A frame body's implicit step (FemFrame::finalize -> bl_fem_create, compute_forces + solve -> bl_fem_step): its
// triangles and members, its masses; its node records back after the step.
inline bl_fem* make_fem(const SoftBody& b, const std::vector<uint32_t>& tri_corners, const std::vector<uint8_t>& fixed,
                        const std::vector<float>& masses, const bl_shell_section& shell,
                        const std::vector<bl_frame_member>& members, const std::vector<bl_frame_section>& sections) {
    std::vector<float> p(3 * b.nodes.size());
    for (size_t i = 0; i < b.nodes.size(); i++) p[3 * i] = b.nodes[i].p.x, p[3 * i + 1] = b.nodes[i].p.y, p[3 * i + 2] = b.nodes[i].p.z;
    bl_fem_desc d{};
    d.n_nodes = (uint32_t)b.nodes.size();
    d.n_tris = (uint32_t)(tri_corners.size() / 3);
    d.p = p.data();
    d.tris = tri_corners.data();
    d.fixed = fixed.data();
    d.mass = masses.data();
    d.section = shell;
    d.n_members = (uint32_t)members.size(), d.members = members.data();
    d.n_sections = (uint32_t)sections.size(), d.sections = sections.data();
    bl_fem* f = nullptr;
    if (bl_fem_create(&f, &d) != BL_OK) return nullptr;
    return f;
}

inline int fem_substep(bl_fem* f, SoftBody& b, const World& w, const std::vector<float>& loads,
                       const std::vector<float>& moments) {
    if (!moments.empty() && bl_fem_add_torques(f, moments.data()) != BL_OK) return -1;
    const float g[3] = {w.settings.gravity.x, w.settings.gravity.y, w.settings.gravity.z};
    bl_fem_report rep;
    int rc = bl_fem_step(f, loads.empty() ? nullptr : loads.data(), w.settings.dt, w.settings.frame_theta,
                         w.settings.frame_dissipation, g, &rep);
    if (rc != BL_OK) return rc;
    std::vector<float> force(3 * b.nodes.size());
    int64_t counters[3];
    return bl_fem_get_nodes(f, reinterpret_cast<bl_node*>(b.nodes.data()), force.data(), counters);
}

} // namespace bl::phys
