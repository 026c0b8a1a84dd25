// Test infrastructure only: a C shim over the compiled reference (zenkovich/beamlab, src/phys + src/core), built
// from the reference's own sources where they lie by oracle/Makefile into oracle/_ref/libblref.so. Only the tests,
// the smoke check and bench.py's CPU-baseline leg load it, as a checker and a baseline; the product never does.
//
// It takes the same plain records as include/beamlab_gpu.h, builds the reference's SoftBody / StaticWorld from
// them (copying the records as they are, so both sides start from identical bits) and runs the reference's own
// World::step_substeps (src/phys/world.cpp:368).
#include "phys/world.h"
#include "phys/fem_shell.h"
#include "phys/softbody.h"
#include "phys/static_world.h"
#include "core/jobs.h"

#include "beamlab_gpu.h"

#include <chrono>
#include <cstddef>
#include <cstring>
#include <memory>
#include <vector>

using namespace bl;
using namespace bl::phys;

// The C records are layout-identical to the reference types: the drop-in passes the reference's arrays as they are.
static_assert(sizeof(bl_node) == sizeof(Node), "bl_node vs Node");
static_assert(offsetof(bl_node, inv_mass) == offsetof(Node, inv_mass), "Node.inv_mass");
static_assert(offsetof(bl_node, v) == offsetof(Node, v), "Node.v");
static_assert(offsetof(bl_node, mass) == offsetof(Node, mass), "Node.mass");
static_assert(sizeof(bl_node_info) == sizeof(NodeInfo), "bl_node_info vs NodeInfo");
static_assert(offsetof(bl_node_info, friction) == offsetof(NodeInfo, friction), "NodeInfo.friction");
static_assert(offsetof(bl_node_info, radius) == offsetof(NodeInfo, radius), "NodeInfo.radius");
static_assert(offsetof(bl_node_info, frame) == offsetof(NodeInfo, frame), "NodeInfo.frame");
static_assert(sizeof(bl_beam) == sizeof(Beam), "bl_beam vs Beam");
static_assert(offsetof(bl_beam, L) == offsetof(Beam, L), "Beam.L");
static_assert(offsetof(bl_beam, strength) == offsetof(Beam, strength), "Beam.strength");
static_assert(offsetof(bl_beam, support_limit) == offsetof(Beam, support_limit), "Beam.support_limit");
static_assert(offsetof(bl_beam, type) == offsetof(Beam, type), "Beam.type");
static_assert(offsetof(bl_beam, flags) == offsetof(Beam, flags), "Beam.flags");
static_assert(offsetof(bl_beam, group) == offsetof(Beam, group), "Beam.group");
static_assert(sizeof(bl_triangle) == sizeof(Triangle), "bl_triangle vs Triangle");
static_assert(offsetof(bl_triangle, two_sided) == offsetof(Triangle, two_sided), "Triangle.two_sided");
static_assert(offsetof(bl_triangle, rest_edge2) == offsetof(Triangle, rest_edge2), "Triangle.rest_edge2");
static_assert(sizeof(bl_capsule) == sizeof(Capsule), "bl_capsule vs Capsule");

namespace {

struct RefWorld {
    World w;
    std::vector<SoftBody*> bodies;
    // per body: a controller's torque and brake per wheel, set by its pre_substep hook when present
    std::vector<std::unique_ptr<std::vector<float>>> ctl;
};

int g_threads_inited = 0;

vec3 v3(const float* f) { return vec3(f[0], f[1], f[2]); }

void apply_opts(SoftBody& b, const bl_body_opts& o) {
    b.collision_radius = o.collision_radius;
    b.ground_friction = o.ground_friction;
    b.bounce = o.bounce;
    b.contact_push_max = o.contact_push_max;
    b.contact_slop = o.contact_slop;
    b.hull_depth = o.hull_depth;
    b.sphere_ball = o.sphere_ball;
    b.contact_friction = o.contact_friction;
    b.allow_break = o.allow_break != 0;
    b.allow_deform = o.allow_deform != 0;
    b.ductile = o.ductile != 0;
    b.self_collision = o.self_collision != 0;
    b.collision_group = o.collision_group;
    b.faces_only = o.faces_only != 0;
    b.face_skin = o.face_skin;
    b.air_drag = o.air_drag;
    b.aero_cda = o.aero_cda;
    b.rest_damp = o.rest_damp;
    b.energy_guard = o.energy_guard != 0;
}

} // namespace

extern "C" {

// threads: JobSystem threads (workers + caller); the first call fixes the pool for the process.
void* blref_create(const bl_settings* s, int threads) {
    if (!g_threads_inited) {
        JobSystem::get().init(threads > 1 ? threads - 1 : 0);
        g_threads_inited = threads > 1 ? threads : 1;
    }
    auto* rw = new RefWorld();
    WorldSettings& ws = rw->w.settings;
    ws.gravity = v3(s->gravity);
    ws.dt = s->dt;
    ws.collision_hz = s->collision_hz;
    ws.frame_theta = s->frame_theta;
    ws.frame_dissipation = s->frame_dissipation;
    ws.inter_body_collisions = s->inter_body_collisions != 0;
    ws.sleeping = s->sleeping != 0;
    ws.tyre_grip = s->tyre_grip;
    ws.multithreaded = threads > 1;
    rw->w.statics.kill_y = s->kill_y;
    return rw;
}

void blref_destroy(void* h) { delete static_cast<RefWorld*>(h); }

int blref_threads(void) { return JobSystem::get().num_threads(); }

// Joins the job system's workers (before the process exits: a parked worker otherwise holds the exit up).
void blref_shutdown(void) { JobSystem::get().shutdown(); }

int blref_set_statics(void* h, const bl_box* boxes, int n) {
    auto* rw = static_cast<RefWorld*>(h);
    rw->w.statics.boxes.clear();
    for (int i = 0; i < n; i++) {
        StaticBox b;
        b.center = v3(boxes[i].center);
        b.rot = mat3(v3(boxes[i].rot), v3(boxes[i].rot + 3), v3(boxes[i].rot + 6));
        b.half = v3(boxes[i].half);
        b.surface = (uint8_t)boxes[i].surface;
        b.max_force = boxes[i].max_force;
        b.update_aabb();
        rw->w.statics.boxes.push_back(b);
    }
    rw->w.statics.build_grid();
    return 0;
}

int blref_set_cylinders(void* h, const bl_cylinder* c, int n) {
    auto* rw = static_cast<RefWorld*>(h);
    rw->w.statics.cylinders.clear();
    for (int i = 0; i < n; i++) {
        StaticCylinder cy;
        cy.base = v3(c[i].base), cy.radius = c[i].radius, cy.height = c[i].height, cy.surface = (uint8_t)c[i].surface;
        rw->w.statics.cylinders.push_back(cy);
    }
    return 0;
}

// Heightfield::create, then its heights and surfaces vertex by vertex, update_bounds.
int blref_set_terrain(void* h, int nx, int nz, float cell, const float* origin, const float* heights, const uint8_t* surf) {
    auto* rw = static_cast<RefWorld*>(h);
    StaticWorld& sw = rw->w.statics;
    sw.has_terrain = nx >= 2 && nz >= 2;
    if (!sw.has_terrain) return 0;
    sw.terrain.create(nx, nz, cell, vec2(origin[0], origin[1]));
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            sw.terrain.h(x, z) = heights[(size_t)z * nx + x];
            sw.terrain.surf(x, z) = surf ? surf[(size_t)z * nx + x] : (uint8_t)0;
        }
    sw.terrain.update_bounds();
    return 0;
}

// The reference's own box (StaticWorld::add_box): to check bl_box_from_quat bit for bit.
int blref_box_from_quat(const float* c, const float* half, const float* q, uint32_t surface, bl_box* out) {
    StaticWorld sw;
    sw.add_box(v3(c), v3(half), quat(q[0], q[1], q[2], q[3]), (uint8_t)surface);
    const StaticBox& b = sw.boxes[0];
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 3; k++) {
        out->center[k] = b.center[k];
        out->half[k] = b.half[k];
        out->aabb_mn[k] = b.aabb.mn[k];
        out->aabb_mx[k] = b.aabb.mx[k];
        for (int r = 0; r < 3; r++) out->rot[3 * k + r] = b.rot.c[k][r];
    }
    out->surface = b.surface;
    out->max_force = b.max_force;
    return 0;
}

int blref_add_body(void* h, const bl_body_desc* d) {
    auto* rw = static_cast<RefWorld*>(h);
    auto b = std::make_unique<SoftBody>();
    b->nodes.resize(d->n_nodes);
    if (d->n_nodes) std::memcpy(b->nodes.data(), d->nodes, sizeof(Node) * d->n_nodes);
    b->info.resize(d->n_nodes);
    for (uint32_t i = 0; i < d->n_nodes; i++) {
        NodeInfo& in = b->info[i];
        in.flags = d->info[i].flags;
        in.surface = d->info[i].surface;
        in.friction = d->info[i].friction;
        in.radius = d->info[i].radius;
        in.frame = d->info[i].frame;
        in.ror_id = d->info[i].ror_id;
    }
    b->force.assign(d->n_nodes, vec3(0));
    b->beams.resize(d->n_beams);
    if (d->n_beams) std::memcpy(b->beams.data(), d->beams, sizeof(Beam) * d->n_beams);
    for (uint32_t i = 0; i < d->n_tris; i++) {
        Triangle t;
        t.a = d->tris[i].a, t.b = d->tris[i].b, t.c = d->tris[i].c;
        t.surface = d->tris[i].surface;
        t.two_sided = d->tris[i].two_sided != 0;
        t.torn = d->tris[i].torn != 0;
        t.rest_edge2 = d->tris[i].rest_edge2;
        b->tris.push_back(t);
    }
    for (uint32_t i = 0; i < d->n_capsules; i++) {
        Capsule c;
        c.a = d->capsules[i].a, c.b = d->capsules[i].b, c.radius = d->capsules[i].radius, c.joint = d->capsules[i].joint;
        b->capsules.push_back(c);
    }
    for (uint32_t i = 0; i < d->n_shocks; i++) {
        const bl_shock& x = d->shocks[i];
        Shock s;
        s.beam = x.beam, s.type = (uint8_t)x.type, s.soft_bump = x.soft_bump != 0;
        s.spring = x.spring, s.damp = x.damp;
        s.spring_in = x.spring_in, s.damp_in = x.damp_in, s.prog_spring_in = x.prog_spring_in, s.prog_damp_in = x.prog_damp_in;
        s.spring_out = x.spring_out, s.damp_out = x.damp_out, s.prog_spring_out = x.prog_spring_out;
        s.prog_damp_out = x.prog_damp_out;
        s.damp_in_slow = x.damp_in_slow, s.split_in = x.split_in, s.damp_in_fast = x.damp_in_fast;
        s.damp_out_slow = x.damp_out_slow, s.split_out = x.split_out, s.damp_out_fast = x.damp_out_fast;
        s.short_bound = x.short_bound, s.long_bound = x.long_bound;
        s.bound_spring = x.bound_spring, s.bound_damp = x.bound_damp;
        b->shocks.push_back(s);
    }
    for (uint32_t i = 0; i < d->n_slides; i++) {
        const bl_slide& x = d->slides[i];
        SlideNode s;
        s.node = x.node;
        s.rail.assign(d->rail + x.rail0, d->rail + x.rail0 + x.nrail);
        s.seg = x.seg, s.k = x.k, s.threshold = x.threshold, s.break_force = x.break_force, s.broken = x.broken != 0;
        b->slides.push_back(s);
    }
    for (uint32_t i = 0; i < d->n_frames; i++) {
        const bl_frame& x = d->frames[i];
        Frame f;
        f.node = x.node, f.q = quat(x.q[0], x.q[1], x.q[2], x.q[3]), f.w = v3(x.w), f.inv_inertia = x.inv_inertia;
        f.torque = vec3(0), f.ref_x = x.ref_x, f.ref_y = x.ref_y;
        b->frames.push_back(f);
    }
    for (uint32_t i = 0; i < d->n_joints; i++) {
        const bl_joint& x = d->joints[i];
        Joint j{};
        j.parent_frame = x.parent_frame, j.child_node = x.child_node, j.child_frame = x.child_frame;
        j.local_offset = v3(x.local_offset), j.rel_rot = quat(x.rel_rot[0], x.rel_rot[1], x.rel_rot[2], x.rel_rot[3]);
        j.k_lin = x.k_lin, j.d_lin = x.d_lin, j.k_ang = x.k_ang, j.d_ang = x.d_ang;
        j.break_force = x.break_force, j.break_torque = x.break_torque, j.yield_angle = x.yield_angle;
        j.max_stretch = x.max_stretch, j.max_bend = x.max_bend;
        j.plastic_lin = x.plastic_lin, j.plastic_ang = x.plastic_ang, j.stress = x.stress, j.broken = x.broken != 0;
        j.radius = 0.05f;
        b->joints.push_back(j);
    }
    for (uint32_t i = 0; i < d->n_wheels; i++) {
        const bl_wheel& x = d->wheels[i];
        Wheel wh;
        wh.axle0 = x.axle0, wh.axle1 = x.axle1;
        wh.nodes.assign(d->wheel_nodes + x.node0, d->wheel_nodes + x.node0 + x.n_nodes);
        wh.radius = x.radius, wh.propulsed = x.propulsed, wh.arm = x.arm, wh.near_attach = x.near_attach;
        wh.mass = x.mass, wh.torque = x.torque, wh.drive_torque = x.drive_torque, wh.brake = x.brake;
        wh.speed = x.speed, wh.avg_speed = x.avg_speed, wh.last_torque = x.last_torque, wh.last_retorque = x.last_retorque;
        wh.detached = x.detached != 0, wh.ring = x.ring != 0;
        b->wheels.push_back(wh);
    }
    apply_opts(*b, d->opts);
    SoftBody* p = rw->w.add_body(std::move(b));
    rw->bodies.push_back(p);
    rw->ctl.push_back(std::make_unique<std::vector<float>>());
    return (int)rw->bodies.size() - 1;
}

int blref_read_slides(void* h, uint32_t id, bl_slide* out) {
    auto* rw = static_cast<RefWorld*>(h);
    if (id >= rw->bodies.size()) return -1;
    const SoftBody& b = *rw->bodies[id];
    for (size_t i = 0; i < b.slides.size(); i++) out[i].seg = b.slides[i].seg, out[i].broken = b.slides[i].broken ? 1u : 0u;
    return 0;
}

int blref_read_wheels(void* h, uint32_t id, bl_wheel* wheels, float* speed_spin) {
    auto* rw = static_cast<RefWorld*>(h);
    if (id >= rw->bodies.size()) return -1;
    const SoftBody& b = *rw->bodies[id];
    for (size_t i = 0; wheels && i < b.wheels.size(); i++) {
        const Wheel& w = b.wheels[i];
        bl_wheel& x = wheels[i];
        x.axle0 = w.axle0, x.axle1 = w.axle1, x.n_nodes = (uint32_t)w.nodes.size();
        x.radius = w.radius, x.propulsed = w.propulsed, x.arm = w.arm, x.near_attach = w.near_attach, x.mass = w.mass;
        x.torque = w.torque, x.drive_torque = w.drive_torque, x.brake = w.brake;
        x.speed = w.speed, x.avg_speed = w.avg_speed, x.last_torque = w.last_torque, x.last_retorque = w.last_retorque;
        x.detached = w.detached ? 1u : 0u, x.ring = w.ring ? 1u : 0u;
    }
    if (speed_spin) speed_spin[0] = b.wheel_speed, speed_spin[1] = b.wheel_spin;
    return 0;
}

// A vehicle controller in the reference's own form: a pre_substep hook setting each wheel's torque and brake.
int blref_set_wheel_controls(void* h, uint32_t id, const float* torque, const float* brake) {
    auto* rw = static_cast<RefWorld*>(h);
    if (id >= rw->bodies.size()) return -1;
    SoftBody& b = *rw->bodies[id];
    if (!torque || !brake) {
        b.pre_substep = nullptr;
        return 0;
    }
    std::vector<float>* c = rw->ctl[id].get();
    c->assign(2 * b.wheels.size(), 0.0f);
    for (size_t k = 0; k < b.wheels.size(); k++) (*c)[2 * k] = torque[k], (*c)[2 * k + 1] = brake[k];
    b.pre_substep = [c](SoftBody& body, float) {
        for (size_t k = 0; k < body.wheels.size(); k++) body.wheels[k].torque = (*c)[2 * k], body.wheels[k].brake = (*c)[2 * k + 1];
    };
    return 0;
}

int blref_read_joints(void* h, uint32_t id, bl_frame* frames, bl_joint* joints) {
    auto* rw = static_cast<RefWorld*>(h);
    if (id >= rw->bodies.size()) return -1;
    const SoftBody& b = *rw->bodies[id];
    for (size_t i = 0; frames && i < b.frames.size(); i++) {
        const Frame& f = b.frames[i];
        bl_frame& x = frames[i];
        x.node = f.node, x.q[0] = f.q.x, x.q[1] = f.q.y, x.q[2] = f.q.z, x.q[3] = f.q.w;
        x.w[0] = f.w.x, x.w[1] = f.w.y, x.w[2] = f.w.z, x.inv_inertia = f.inv_inertia, x.ref_x = f.ref_x, x.ref_y = f.ref_y;
    }
    for (size_t i = 0; joints && i < b.joints.size(); i++) {
        const Joint& j = b.joints[i];
        bl_joint& x = joints[i];
        x.parent_frame = j.parent_frame, x.child_node = j.child_node, x.child_frame = j.child_frame;
        x.local_offset[0] = j.local_offset.x, x.local_offset[1] = j.local_offset.y, x.local_offset[2] = j.local_offset.z;
        x.rel_rot[0] = j.rel_rot.x, x.rel_rot[1] = j.rel_rot.y, x.rel_rot[2] = j.rel_rot.z, x.rel_rot[3] = j.rel_rot.w;
        x.k_lin = j.k_lin, x.d_lin = j.d_lin, x.k_ang = j.k_ang, x.d_ang = j.d_ang;
        x.break_force = j.break_force, x.break_torque = j.break_torque, x.yield_angle = j.yield_angle;
        x.max_stretch = j.max_stretch, x.max_bend = j.max_bend;
        x.plastic_lin = j.plastic_lin, x.plastic_ang = j.plastic_ang, x.stress = j.stress, x.broken = j.broken ? 1u : 0u;
    }
    return 0;
}

int blref_step(void* h, int n) {
    static_cast<RefWorld*>(h)->w.step_substeps(n);
    return 0;
}

// Steps `calls` times `n` substeps; returns the wall time in seconds.
double blref_time_step(void* h, int calls, int n) {
    auto* rw = static_cast<RefWorld*>(h);
    auto t0 = std::chrono::steady_clock::now();
    for (int c = 0; c < calls; c++) rw->w.step_substeps(n);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int blref_read_body(void* h, uint32_t id, bl_node* nodes, bl_beam* beams, int64_t* broken) {
    auto* rw = static_cast<RefWorld*>(h);
    if (id >= rw->bodies.size()) return -1;
    const SoftBody& b = *rw->bodies[id];
    if (nodes && !b.nodes.empty()) std::memcpy(nodes, b.nodes.data(), sizeof(Node) * b.nodes.size());
    if (beams && !b.beams.empty()) std::memcpy(beams, b.beams.data(), sizeof(Beam) * b.beams.size());
    if (broken) *broken = b.stats.broken_beams;
    return 0;
}

int blref_write_body(void* h, uint32_t id, const bl_node* nodes, const bl_beam* beams) {
    auto* rw = static_cast<RefWorld*>(h);
    if (id >= rw->bodies.size()) return -1;
    SoftBody& b = *rw->bodies[id];
    if (nodes && !b.nodes.empty()) std::memcpy(b.nodes.data(), nodes, sizeof(Node) * b.nodes.size());
    if (beams && !b.beams.empty()) std::memcpy(b.beams.data(), beams, sizeof(Beam) * b.beams.size());
    return 0;
}

// The reference's SoftBody::finalize + stabilize on a copy of the arrays (src/phys/softbody.cpp:76-110).
int blref_finalize_stabilize(bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb, int do_finalize, float dt,
                             float kb, float db, int* changed) {
    SoftBody b;
    b.nodes.resize(n);
    std::memcpy(b.nodes.data(), nodes, sizeof(Node) * n);
    b.info.resize(n);
    b.beams.resize(nb);
    std::memcpy(b.beams.data(), beams, sizeof(Beam) * nb);
    if (do_finalize) b.finalize();
    int c = dt > 0 ? b.stabilize(dt, kb, db) : 0;
    if (changed) *changed = c;
    std::memcpy(beams, b.beams.data(), sizeof(Beam) * nb);
    return 0;
}

// One shell triangle through the reference's own FemFrame: its rest shape (FemFrame::finalize) and its element
// stiffness (FemFrame::tri_stiffness -> tri_local_k, src/phys/frame_fem.cpp:1524-1579, 1649-1651), 18 x 18 row-major.
// sec: E, nu, rho, t, drill, damping.
int blref_tri_stiffness(const float p[9], const float sec[6], double* K, float* X) {
    SoftBody b;
    for (int i = 0; i < 3; i++) b.add_node(vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), 1.0f);
    ShellSection s;
    s.E = sec[0], s.nu = sec[1], s.rho = sec[2], s.t = sec[3], s.drill = sec[4], s.damping = sec[5];
    const uint16_t si = b.fem.add_shell_section(s);
    b.fem.add_tri(0, 1, 2, si);
    b.fem.finalize(b);
    if (b.fem.tris.empty() || b.fem.tris[0].broken) return -1;
    b.fem.tri_stiffness(0, reinterpret_cast<double(*)[18]>(K));
    for (int i = 0; i < 3; i++) X[2 * i] = b.fem.tris[0].X[i][0], X[2 * i + 1] = b.fem.tris[0].X[i][1];
    return 0;
}

// The clamped plate of the implicit-FEM workload (clamped = 0: its border free) built the reference's way (ShellMesher::grid, src/phys/fem_shell.cpp:
// 34-46; border nodes NF_FIXED; ShellMesher::finish; FemFrame::finalize) and its implicit step as the reference's
// fembench times it (tests/test_physics.cpp:3486-3494): clear_forces(g) plus the loads, FemFrame::compute_forces,
// FemFrame::solve. Positions are not advanced (the world's integration does that in the reference).
void* blref_fem_plate_new(int nu, int nv, float size, float y0, const float sec[6], int clamped) {
    auto* b = new SoftBody();
    ShellSection s;
    s.E = sec[0], s.nu = sec[1], s.rho = sec[2], s.t = sec[3], s.drill = sec[4], s.damping = sec[5];
    const uint16_t si = b->fem.add_shell_section(s);
    ShellMesher m(*b);
    const vec3 origin(-0.5f * size, y0, -0.5f * size), du(size / (float)nu, 0, 0), dv(0, 0, size / (float)nv);
    const std::vector<uint32_t> id = m.grid(origin, du, dv, nu, nv, si, false);
    for (int j = 0; j <= nv; j++)
        for (int i = 0; i <= nu; i++)
            if (clamped && (i == 0 || j == 0 || i == nu || j == nv)) b->info[id[(size_t)j * (nu + 1) + i]].flags |= NF_FIXED;
    m.finish();
    b->finalize();
    b->fem.finalize(*b);
    return b;
}

void blref_fem_plate_free(void* h) { delete static_cast<SoftBody*>(h); }

// one step; returns the frame's solve failures so far
int64_t blref_fem_plate_step(void* h, const float* loads, float dt, float theta, float dissipation, const float g[3]) {
    SoftBody& b = *static_cast<SoftBody*>(h);
    b.clear_forces(vec3(g[0], g[1], g[2]));
    if (loads)
        for (size_t i = 0; i < b.nodes.size(); i++) b.force[i] += vec3(loads[3 * i], loads[3 * i + 1], loads[3 * i + 2]);
    b.fem.compute_forces(b);
    b.fem.solve(b, dt, dt, theta, dissipation);
    return (int64_t)b.fem.solve_failures;
}

// A frame body built the reference's way: nodes (SoftBody::add_node with their masses, NF_FIXED where pinned), shell
// triangles of one section (FemFrame::add_shell_section, add_tri; 7 floats: E, nu, rho, t, drill, damping, yield),
// members (FemFrame::add_section, add_element: a, b, section, end_a | end_b << 8) of sections of 17 floats each (E, G,
// rho, A, Iy, Iz, J, As_y, As_z, damping, axial, joint_k, Np, Mp, Tp, hardening, half), then SoftBody::finalize and
// FemFrame::finalize. The body may deform and does not break (SoftBody::allow_break off: no tears, splits or letting
// go).
void* blref_frame_new(uint32_t nn, const float* p, const float* mass, const uint8_t* fixed, uint32_t nt,
                      const uint32_t* tris, const float shell[7], uint32_t nm, const uint32_t* members, uint32_t nsec,
                      const float* sections) {
    auto* b = new SoftBody();
    b->allow_break = false;
    for (uint32_t i = 0; i < nn; i++)
        b->add_node(vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), mass[i], fixed && fixed[i] ? NF_FIXED : NF_NONE);
    if (nt) {
        ShellSection s;
        s.E = shell[0], s.nu = shell[1], s.rho = shell[2], s.t = shell[3], s.drill = shell[4], s.damping = shell[5];
        s.yield = shell[6];
        const uint16_t si = b->fem.add_shell_section(s);
        for (uint32_t k = 0; k < nt; k++) b->fem.add_tri(tris[3 * k], tris[3 * k + 1], tris[3 * k + 2], si);
    }
    for (uint32_t k = 0; k < nsec; k++) {
        const float* x = &sections[17 * k];
        FrameSection s;
        s.E = x[0], s.G = x[1], s.rho = x[2], s.A = x[3], s.Iy = x[4], s.Iz = x[5], s.J = x[6], s.As_y = x[7], s.As_z = x[8];
        s.damping = x[9], s.axial = x[10], s.joint_k = x[11];
        s.Np = x[12], s.Mp = x[13], s.Tp = x[14], s.hardening = x[15], s.half = x[16];
        b->fem.add_section(s);
    }
    for (uint32_t k = 0; k < nm; k++)
        b->fem.add_element(members[4 * k], members[4 * k + 1], (uint16_t)members[4 * k + 2], (uint8_t)(members[4 * k + 3] & 0xff),
                           (uint8_t)(members[4 * k + 3] >> 8));
    b->finalize();
    b->fem.finalize(*b);
    return b;
}

// moments from outside on the frame's nodes for the next step (FemFrame::torque, added; 3 per body node)
void blref_frame_add_torques(void* h, const float* t) {
    SoftBody& b = *static_cast<SoftBody*>(h);
    for (size_t i = 0; i < b.nodes.size(); i++) {
        const int sl = b.fem.slot((uint32_t)i);
        if (sl >= 0) b.fem.torque[sl] += vec3(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    }
}

// The plastic state: per triangle 29 floats (FrameTri: X[6], th0[9], r0[12] x y z w, dmg, util), per member 27
// (FrameElement: up, tp, pa, pb, ba, bb, bt, damage, dmg_a, dmg_b, util, N, qa, qb; three zeros where the oracle puts
// the private tangent's bits); counters {FemFrame::repasses, passes_}. Any pointer may be null.
void blref_frame_plastic(void* h, float* tris, float* members, int64_t* counters) {
    const FemFrame& f = static_cast<SoftBody*>(h)->fem;
    if (tris)
        for (size_t k = 0; k < f.tris.size(); k++) {
            const FrameTri& t = f.tris[k];
            float* o = &tris[29 * k];
            for (int i = 0; i < 3; i++) {
                o[2 * i] = t.X[i][0], o[2 * i + 1] = t.X[i][1];
                o[6 + 3 * i] = t.th0[i].x, o[7 + 3 * i] = t.th0[i].y, o[8 + 3 * i] = t.th0[i].z;
                o[15 + 4 * i] = t.r0[i].x, o[16 + 4 * i] = t.r0[i].y, o[17 + 4 * i] = t.r0[i].z, o[18 + 4 * i] = t.r0[i].w;
            }
            o[27] = t.dmg, o[28] = t.util;
        }
    if (members)
        for (size_t k = 0; k < f.elems.size(); k++) {
            const FrameElement& e = f.elems[k];
            float* o = &members[27 * k];
            const float v[27] = {e.up, e.tp, e.pa.x, e.pa.y, e.pb.x, e.pb.y, e.ba.x, e.ba.y, e.bb.x, e.bb.y, e.bt, e.damage,
                                 e.dmg_a, e.dmg_b, e.util, e.N, e.qa.x, e.qa.y, e.qa.z, e.qa.w, e.qb.x, e.qb.y, e.qb.z,
                                 e.qb.w, 0, 0, 0};
            std::memcpy(o, v, sizeof(v));
        }
    if (counters) counters[0] = f.repasses, counters[1] = f.passes_;
}

// The plate body's whole substep as World::simulate_island runs a frame body stepping every substep
// (src/phys/world.cpp:3265-3269, 3573-3580, 3812-3822, 3876-3886): clear_forces(g) plus the loads,
// FemFrame::compute_forces, FemFrame::solve (its update writes the new velocities as forces, :3000-3027), then
// SoftBody::integrate_nodes (src/phys/softbody.cpp:769-795) and FemFrame::sync_positions (:3407-3422).
int blref_fem_plate_substep(void* h, const float* loads, float dt, float theta, float dissipation, const float g[3]) {
    SoftBody& b = *static_cast<SoftBody*>(h);
    b.clear_forces(vec3(g[0], g[1], g[2]));
    if (loads)
        for (size_t i = 0; i < b.nodes.size(); i++) b.force[i] += vec3(loads[3 * i], loads[3 * i + 1], loads[3 * i + 2]);
    b.fem.compute_forces(b);
    b.fem.solve(b, dt, dt, theta, dissipation);
    vec3 mn(1e30f), mx(-1e30f);
    float max_v2 = 0;
    b.integrate_nodes(0, b.nodes.size(), dt, mn, mx, max_v2);
    b.fem.sync_positions(b, dt);
    return 0;
}

// A substep of a frame body stepping its frame every `every` substeps (WorldSettings::frame_every), as
// World::simulate_island runs it (src/phys/world.cpp:3244-3252, 3573-3580, 3612-3642, 3665, 3812-3822): clear_forces(g)
// plus the loads; at substep j % every == 0 FemFrame::compute_forces and FemFrame::solve over h = every dt (the
// dissipation per step divided by every), else the held step (held_begin, held_component for each component,
// held_end; FemFrame::hold without a factorization); then integrate_nodes and sync_positions over dt.
int blref_frame_cycle(void* h, const float* loads, float dt, int every, int j, float theta, float dissipation,
                      const float g[3]) {
    SoftBody& b = *static_cast<SoftBody*>(h);
    b.clear_forces(vec3(g[0], g[1], g[2]));
    if (loads)
        for (size_t i = 0; i < b.nodes.size(); i++) b.force[i] += vec3(loads[3 * i], loads[3 * i + 1], loads[3 * i + 2]);
    if (j % every == 0) {
        b.fem.compute_forces(b);
        b.fem.solve(b, dt * (float)every, dt, theta, dissipation / (float)every);
    } else {
        const int nc = b.fem.held_begin(b, dt);
        if (nc == 0) {
            b.fem.hold(b, dt);
        } else {
            for (int c = 0; c < nc; c++) b.fem.held_component(b, c);
            b.fem.held_end(b);
        }
    }
    vec3 mn(1e30f), mx(-1e30f);
    float max_v2 = 0;
    b.integrate_nodes(0, b.nodes.size(), dt, mn, mx, max_v2);
    b.fem.sync_positions(b, dt);
    return 0;
}

// The plate's state: the body's nodes (p, inv_mass, v, mass), the frame's double positions, spins and orientations
// (scattered to the body's node order: the frame numbers its nodes as the triangles first name them), the forces the
// update left (3 per node), and the counters {solve_failures, clamps}. Any pointer may be null. Returns the frame's
// node count.
int blref_fem_plate_read(void* h, bl_node* nodes, double* xd, float* w, float* q, float* force, int64_t* counters) {
    const SoftBody& b = *static_cast<SoftBody*>(h);
    const FemFrame& f = b.fem;
    if (nodes) std::memcpy(nodes, b.nodes.data(), sizeof(Node) * b.nodes.size());
    for (size_t i = 0; i < f.node.size(); i++) {
        const size_t n = f.node[i];
        if (xd) xd[3 * n] = f.xd[3 * i], xd[3 * n + 1] = f.xd[3 * i + 1], xd[3 * n + 2] = f.xd[3 * i + 2];
        if (w) w[3 * n] = f.w[i].x, w[3 * n + 1] = f.w[i].y, w[3 * n + 2] = f.w[i].z;
        if (q) q[4 * n] = f.q[i].x, q[4 * n + 1] = f.q[i].y, q[4 * n + 2] = f.q[i].z, q[4 * n + 3] = f.q[i].w;
    }
    if (force)
        for (size_t i = 0; i < b.force.size(); i++) force[3 * i] = b.force[i].x, force[3 * i + 1] = b.force[i].y, force[3 * i + 2] = b.force[i].z;
    if (counters) counters[0] = f.solve_failures, counters[1] = f.clamps;
    return (int)f.node.size();
}

// The plate's state written (nodes as bl_node; the frame's xd, w, q in the body's node order); null pointers leave a
// part as it is.
int blref_fem_plate_write(void* h, const bl_node* nodes, const double* xd, const float* w, const float* q) {
    SoftBody& b = *static_cast<SoftBody*>(h);
    FemFrame& f = b.fem;
    if (nodes) std::memcpy(b.nodes.data(), nodes, sizeof(Node) * b.nodes.size());
    for (size_t i = 0; i < f.node.size(); i++) {
        const size_t n = f.node[i];
        if (xd) f.xd[3 * i] = xd[3 * n], f.xd[3 * i + 1] = xd[3 * n + 1], f.xd[3 * i + 2] = xd[3 * n + 2];
        if (w) f.w[i] = vec3(w[3 * n], w[3 * n + 1], w[3 * n + 2]);
        if (q) f.q[i] = quat(q[4 * n], q[4 * n + 1], q[4 * n + 2], q[4 * n + 3]);
    }
    return 0;
}

} // extern "C"
