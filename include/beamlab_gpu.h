/*
 * beamlab_gpu.h - C ABI of the MI355X (gfx950) soft-body step.
 *
 * A drop-in for the CPU step of BeamLab's node/beam/shell physics (reference: zenkovich/beamlab, src/phys).
 * The host keeps its own SoftBody/World objects and hands their arrays over through plain pointers and sizes;
 * the records below are layout-identical to the reference's types so that a caller can pass
 * `body.nodes.data()`, `body.beams.data()` ... unchanged (the layout is asserted against the reference headers
 * in oracle/ref_capi.cpp).
 *
 * Conventions
 *   - Every entry point returns an int status: BL_OK (0) or a negative BL_ERR_* code. No exceptions cross the ABI.
 *     Numerical events (broken beams, clamps, failed pivots) are counters in bl_stats, not status codes, as the
 *     reference reports them (src/phys/softbody.h:588-593, src/phys/frame_fem.h:342-344).
 *   - Ownership: the caller owns every host array; bl_* copies in and out. Device memory belongs to the context.
 *   - Threading: one host thread per context (the reference has a single caller of World::step_frame); a context
 *     owns one HIP stream.
 */
#ifndef BEAMLAB_GPU_H
#define BEAMLAB_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BL_ABI_VERSION 6

enum {
    BL_OK = 0,
    BL_ERR_ARG = -1,         /* a null pointer, an index out of range, a size that does not fit */
    BL_ERR_HIP = -2,         /* a HIP runtime call failed (no device, out of device memory, ...) */
    BL_ERR_UNSUPPORTED = -3, /* a feature of the body that this build does not step on the GPU */
    BL_ERR_STATE = -4,       /* a call out of order (e.g. stepping before any body was added) */
};

/* ---------------------------------------------------------------------------------------------------------------
 * Records, layout-identical to the reference's
 * ------------------------------------------------------------------------------------------------------------- */

/* == bl::phys::Node (src/phys/softbody.h:43-48): 32 B, p, inv_mass, v, mass */
typedef struct bl_node {
    float p[3];
    float inv_mass;
    float v[3];
    float mass;
} bl_node;

/* == bl::phys::NodeInfo (src/phys/softbody.h:50-58): 20 B */
typedef struct bl_node_info {
    uint16_t flags; /* BL_NF_* == NodeFlag (src/phys/softbody.h:29-40) */
    uint8_t surface;
    uint8_t pad;
    float friction;
    float radius;
    int32_t frame;
    int32_t ror_id;
} bl_node_info;

enum {
    BL_NF_GROUND = 1 << 0,
    BL_NF_FIXED = 1 << 1,
    BL_NF_TYRE = 1 << 2,
    BL_NF_RIM = 1 << 3,
    BL_NF_CONTACTER = 1 << 4,
    BL_NF_FRAME = 1 << 5,
    BL_NF_LOAD = 1 << 6,
    BL_NF_NO_MOUSE = 1 << 7,
    BL_NF_SELF_COLL = 1 << 8,
};

/* == bl::phys::Beam (src/phys/softbody.h:71-85): 52 B */
typedef struct bl_beam {
    uint32_t a, b;
    float L;             /* current rest length (plastic) */
    float k, d;          /* spring N/m, damping N s/m */
    float max_pos;       /* compressive yield force (> 0) */
    float max_neg;       /* tensile yield force (< 0) */
    float strength;      /* break force */
    float plastic;
    float stress;        /* last force (positive = compression) */
    float L0;
    float support_limit;
    uint8_t type;        /* BL_BT_* == BeamType (src/phys/softbody.h:60) */
    uint8_t flags;       /* BL_BF_* == BeamFlag (src/phys/softbody.h:61-69) */
    uint16_t group;
} bl_beam;

enum { BL_BT_NORMAL = 0, BL_BT_ROPE = 1, BL_BT_SUPPORT = 2 };
enum {
    BL_BF_BROKEN = 1 << 0,
    BL_BF_INVISIBLE = 1 << 1,
    BL_BF_NO_DEFORM = 1 << 2,
    BL_BF_NO_BREAK = 1 << 3,
    BL_BF_WHEEL = 1 << 4,
    BL_BF_HYDRO = 1 << 5,
    BL_BF_SHOCK = 1 << 6,
};

/* == bl::phys::Triangle (src/phys/softbody.h:126-137): 20 B */
typedef struct bl_triangle {
    uint32_t a, b, c;
    uint8_t surface;
    uint8_t two_sided;
    uint8_t torn;
    uint8_t pad;
    float rest_edge2;
} bl_triangle;

/* == bl::phys::Capsule (src/phys/softbody.h:192-196): 16 B */
typedef struct bl_capsule {
    uint32_t a, b;
    float radius;
    int32_t joint;
} bl_capsule;

/* A shock absorber (== bl::phys::Shock, src/phys/softbody.h:88-101, without the frame seats): the beam it drives is
 * flagged BL_BF_SHOCK, so that the beam loop leaves it to the shock loop (SoftBody::compute_shock_forces,
 * src/phys/softbody.cpp:472-572), which runs after the beams. */
typedef struct bl_shock {
    uint32_t beam;           /* index into the body's beams */
    uint32_t type;           /* 1 shocks (bounded wheel beams), 2 shocks2, 3 shocks3 */
    uint32_t soft_bump;      /* shocks2 's' */
    float spring, damp;      /* type 1, in bounds */
    float spring_in, damp_in, prog_spring_in, prog_damp_in;
    float spring_out, damp_out, prog_spring_out, prog_damp_out;
    float damp_in_slow, split_in, damp_in_fast, damp_out_slow, split_out, damp_out_fast;
    float short_bound, long_bound; /* fractions of L */
    float bound_spring, bound_damp; /* stiffness at the bump stops */
} bl_shock;

/* A slide node (== bl::phys::SlideNode, src/phys/softbody.h:559-567; SoftBody::compute_slide_forces,
 * src/phys/softbody.cpp:574-623): a spring from `node` to the nearest point of its rail, the rail's nodes being
 * rail[rail0 .. rail0 + nrail) of the body description; `seg` and `broken` are the state the step updates. */
typedef struct bl_slide {
    uint32_t node;
    uint32_t rail0, nrail;
    int32_t seg;             /* current segment (its first rail node's index in the rail); -1: search the whole rail */
    float k;                 /* default 9e6 */
    float threshold;         /* free play (m) */
    float break_force;       /* 0: unbreakable */
    uint32_t broken;
} bl_slide;

/* An oriented node (== bl::phys::Frame, src/phys/softbody.h:104-113): orientation and angular velocity, integrated
 * after the nodes (SoftBody::integrate_frames, src/phys/softbody.cpp:997-1021) from the joints' torques; a kinematic
 * frame (ref_x, ref_y: nodes of the body, the truck format's triads) takes its triad's orientation instead. */
typedef struct bl_frame {
    uint32_t node;
    float q[4];              /* x, y, z, w */
    float w[3];              /* angular velocity (world) */
    float inv_inertia;       /* isotropic; 0: not integrated */
    int32_t ref_x, ref_y;    /* the triad's nodes (x towards ref_x, z normal to the plane of ref_y); -1: dynamic */
} bl_frame;

/* An orientation-preserving joint (== bl::phys::Joint, src/phys/softbody.h:116-136; SoftBody::compute_joint_forces,
 * src/phys/softbody.cpp:625-679): a 6-DOF spring from a parent frame to a child node (and its frame, if any), with
 * plastic yield and breaking; local_offset, rel_rot, the plastic strains, stress and broken are the step's state. */
typedef struct bl_joint {
    uint32_t parent_frame, child_node;
    int32_t child_frame;     /* -1: the child has no orientation */
    float local_offset[3];   /* the child's position in the parent frame (rest) */
    float rel_rot[4];        /* the child's orientation relative to the parent (rest): x, y, z, w */
    float k_lin, d_lin, k_ang, d_ang;
    float break_force, break_torque, yield_angle;
    float max_stretch, max_bend;
    float plastic_lin, plastic_ang, stress;
    uint32_t broken;
} bl_joint;

/* A wheel (== bl::phys::Wheel, src/phys/softbody.h:468-490, its node-tyre part; SoftBody::compute_wheel_forces,
 * src/phys/softbody.cpp:689-767): the drive torque spread over its tyre nodes as tangential tread forces, the
 * brake's stop torque, the tread's speed, the reaction torque on its arm. Its tyre nodes are
 * wheel_nodes[node0 .. node0 + n_nodes) of the body description (even: the axle0 side, odd: the axle1 side). The
 * tyre nodes touching the ground share their wheel's grip (World::collide_static, src/phys/world.cpp:852-892).
 * torque, drive_torque, brake and the state fields are the step's. Ring tyres (Wheel::ring) are not stepped on the
 * GPU. */
typedef struct bl_wheel {
    uint32_t axle0, axle1;
    uint32_t node0, n_nodes;
    float radius;
    int32_t propulsed;        /* 0 none, 1 forward, 2 backward */
    int32_t arm, near_attach; /* the reaction torque's arm node and the axle node nearest it; -1: none */
    float mass;               /* the tyre nodes' mass */
    float torque, drive_torque, brake;
    float speed, avg_speed, last_torque, last_retorque;
    uint32_t detached;
    uint32_t ring;            /* must be 0 */
} bl_wheel;

/* A static oriented box (== StaticBox, src/phys/static_world.h:115-126, without hull planes). `rot` holds the
 * columns of the rotation (the box's local axes in world space), as bl::mat3 does; aabb_mn/aabb_mx as
 * StaticBox::update_aabb (src/phys/static_world.cpp:242-246) computes them (bl_box_from_quat fills both). */
typedef struct bl_box {
    float center[3];
    float rot[9];
    float half[3];
    float aabb_mn[3];
    float aabb_mx[3];
    float max_force;  /* > 0: a yielding obstacle, the per-node contact force is capped */
    uint32_t surface; /* BL_SURF_* == Surface (src/phys/static_world.h:59-62) */
} bl_box;

enum {
    BL_SURF_GRASS = 0, BL_SURF_DIRT, BL_SURF_ASPHALT, BL_SURF_CONCRETE, BL_SURF_GRAVEL, BL_SURF_MUD, BL_SURF_SAND,
    BL_SURF_ROCK, BL_SURF_WOOD, BL_SURF_METAL, BL_SURF_ICE, BL_SURF_COUNT
};

/* The world settings the path reads (a subset of WorldSettings, src/phys/world.h:24-77). */
typedef struct bl_settings {
    float gravity[3];
    float dt;                  /* substep, default 0.0005 (2 kHz) */
    float collision_hz;        /* pair-search rate, default 120 */
    float frame_theta;         /* FEM implicit step theta, default 0.5 */
    float frame_dissipation;   /* FEM numerical dissipation, default 0.02 */
    int32_t inter_body_collisions;
    int32_t sleeping;          /* must be 0: sleeping is not stepped on the GPU */
    float kill_y;              /* StaticWorld::kill_y, default -200 */
    float tyre_grip;           /* WorldSettings::tyre_grip: the tyre nodes' friction multiplier, default 1 */
} bl_settings;

/* A vertical cylinder of the static world (== bl::phys::StaticCylinder, src/phys/static_world.h:128-132). */
typedef struct bl_cylinder {
    float base[3];
    float radius, height;
    uint32_t surface;
} bl_cylinder;

/* The per-body settings the path reads (SoftBody, src/phys/softbody.h:649-716). */
typedef struct bl_body_opts {
    float collision_radius;    /* default 0.05 */
    float ground_friction;     /* default 1 */
    float bounce;              /* default -1 (RoR: 80 % of the approach taken) */
    float contact_push_max;    /* default 2 */
    float contact_slop;        /* default 0 */
    float hull_depth;          /* default 0.3 */
    float sphere_ball;         /* default 0 */
    float contact_friction;    /* default 1 */
    int32_t allow_break;       /* default 1 */
    int32_t allow_deform;      /* default 1 */
    int32_t ductile;           /* default 0 */
    int32_t self_collision;    /* default 0 */
    int32_t collision_group;   /* default 0 */
    int32_t faces_only;        /* default 0 */
    float face_skin;           /* default 0.002: a faces-only hull triangle's reach (m) */
    /* features the GPU path does not step; non-zero values are rejected with BL_ERR_UNSUPPORTED */
    float air_drag;
    float aero_cda;
    float rest_damp;
    /* SoftBody::energy_guard (src/phys/softbody.h:790-799): after 15 quiet frames, a frame's energy gain beyond 2 J
     * and 5 % of its kinetic energy at the frame's start is scaled back (checked on the host after each frame) */
    int32_t energy_guard;
} bl_body_opts;

typedef struct bl_body_desc {
    const bl_node* nodes;
    const bl_node_info* info;
    uint32_t n_nodes;
    const bl_beam* beams;
    uint32_t n_beams;
    const bl_triangle* tris;
    uint32_t n_tris;
    const bl_capsule* capsules;
    uint32_t n_capsules;
    bl_body_opts opts;
    /* shocks and slide nodes, stepped after the beams in this order (World::simulate_island's base_forces,
     * src/phys/world.cpp:3255-3273); null with a zero count for none */
    const bl_shock* shocks;
    uint32_t n_shocks;
    const bl_slide* slides;
    uint32_t n_slides;
    const uint32_t* rail;    /* the slides' rail nodes */
    uint32_t n_rail;
    /* oriented nodes and the joints between them, stepped between the shocks and the slides */
    const bl_frame* frames;
    uint32_t n_frames;
    const bl_joint* joints;
    uint32_t n_joints;
    /* wheels (after the joints, before the slides) and their tyre nodes */
    const bl_wheel* wheels;
    uint32_t n_wheels;
    const uint32_t* wheel_nodes;
    uint32_t n_wheel_nodes;
} bl_body_desc;

/* Counters (WorldStats / BodyStats / FemFrame counters: src/phys/world.h:79-132, src/phys/softbody.h:588-593). */
typedef struct bl_stats {
    int64_t substeps;
    int64_t broken_beams;
    int64_t static_contacts;   /* node-vs-static contacts, summed over substeps */
    int64_t vel_clamps;        /* nodes clamped to 400 m/s */
    int64_t nan_resets;        /* non-finite velocities reset to 0 */
    int64_t pair_rebuilds;
    int64_t contact_pairs;     /* candidate pairs at the last rebuild */
    int64_t narrow_tests;      /* candidate pairs tested, summed over substeps */
    int64_t contacts;          /* inter-body contacts, summed over substeps */
    int64_t solve_failures;
    float aabb_mn[3];          /* bounds and top speed of the last substep, all bodies */
    float aabb_mx[3];
    float max_speed;
    float pad;
} bl_stats;

/* Execution choices (written by the autotuner, or by hand). Zero fields take the defaults. */
typedef struct bl_tuning {
    int32_t mode;                  /* BL_MODE_* */
    int32_t block;                 /* threads per workgroup: 256, 512 or 1024 */
    int32_t substeps_per_launch;   /* persistent forms: substeps per launch (0: a whole bl_step) */
    int32_t use_graph;             /* stepped form: capture the launches in a hipGraph */
    int32_t incidences_per_thread; /* beam ends per thread: 1, 2, 4 or 8 (a chunk holds block x this many) */
} bl_tuning;

enum {
    BL_MODE_AUTO = 0,
    BL_MODE_STEPPED = 1,    /* one launch per substep, one workgroup per chunk */
    BL_MODE_PERSISTENT = 2, /* one workgroup per body, many substeps per launch */
    BL_MODE_RESIDENT = 3,   /* one workgroup per chunk, state in registers, many substeps per launch (cooperative) */
    BL_MODE_HALO = 4,       /* as RESIDENT, each chunk's neighbour nodes staged in LDS, neighbour-only waits */
};

/* ---------------------------------------------------------------------------------------------------------------
 * Host-side setup helpers (no device needed)
 * ------------------------------------------------------------------------------------------------------------- */

/* SoftBody::finalize's beam part (src/phys/softbody.cpp:76-83): L = L0 = max(1e-3, |pa - pb|); and each
 * triangle's rest_edge2. */
int bl_finalize(const bl_node* nodes, uint32_t n_nodes, bl_beam* beams, uint32_t n_beams, bl_triangle* tris,
                uint32_t n_tris);

/* SoftBody::stabilize's beam part (src/phys/softbody.cpp:85-110): per node sum(k) dt^2 / m <= k_budget and
 * sum(d) dt / m <= d_budget, each beam scaled by the smaller factor of its two nodes. *changed = beams scaled. */
int bl_stabilize(const bl_node* nodes, uint32_t n_nodes, bl_beam* beams, uint32_t n_beams, float dt, float k_budget,
                 float d_budget, int32_t* changed);

/* A box from its centre, half extents and rotation quaternion (x, y, z, w), as StaticWorld::add_box does
 * (src/phys/static_world.cpp:298-306, to_mat3 src/core/math.h:163-170). */
int bl_box_from_quat(const float center[3], const float half[3], const float quat_xyzw[4], uint32_t surface,
                     bl_box* out);

void bl_default_settings(bl_settings* s);
void bl_default_body_opts(bl_body_opts* o);
const char* bl_error_string(int status);
int bl_abi_version(void);

/* ---------------------------------------------------------------------------------------------------------------
 * The step
 * ------------------------------------------------------------------------------------------------------------- */

typedef struct bl_ctx bl_ctx;

/* A context on the current HIP device: owns the device copies of the world and one stream. */
int bl_create(bl_ctx** out, const bl_settings* settings);
void bl_destroy(bl_ctx* ctx);

/* The static world (replaces the previous boxes). World::statics (src/phys/world.h:158). */
int bl_set_statics(bl_ctx* ctx, const bl_box* boxes, int32_t n_boxes);

/* The static world's vertical cylinders (replaces the previous ones): StaticWorld::cylinders. */
int bl_set_cylinders(bl_ctx* ctx, const bl_cylinder* cylinders, int32_t n_cylinders);

/* The static world's terrain (StaticWorld::terrain, a Heightfield: src/phys/static_world.h:62-112,
 * src/phys/static_world.cpp:117-230), replacing the previous one: nx x nz vertices `cell` apart from `origin` (x, z),
 * heights[z * nx + x] (-1e30: no ground there), surfaces[z * nx + x] (null: all 0, grass). nx = 0 removes it. The
 * terrain's road and paved layers are not stepped on the GPU. */
int bl_set_terrain(bl_ctx* ctx, int32_t nx, int32_t nz, float cell, const float origin[2], const float* heights,
                   const uint8_t* surfaces);

/* Adds a body (World::add_body, src/phys/world.cpp:283-289). Bodies are stepped together. *id = its index. */
int bl_add_body(bl_ctx* ctx, const bl_body_desc* desc, uint32_t* id);

/* Runs `substeps` substeps of every body (World::step_substeps, src/phys/world.cpp:368) and waits for them. */
int bl_step(bl_ctx* ctx, int32_t substeps);

/* Enqueues `substeps` substeps on the context's stream without waiting (bl_sync waits). */
int bl_step_async(bl_ctx* ctx, int32_t substeps);
int bl_sync(bl_ctx* ctx);

/* Copies a body's state out (either pointer may be null). Beams come back with L, max_pos, max_neg, strength,
 * stress and flags as the step left them; the other fields as they were added. */
int bl_read_body(bl_ctx* ctx, uint32_t id, bl_node* nodes_out, bl_beam* beams_out);

/* Overwrites a body's node and beam state (a snapshot restore; sizes as added; either pointer may be null). */
int bl_write_body(bl_ctx* ctx, uint32_t id, const bl_node* nodes, const bl_beam* beams);

/* A body's broken beams since it was added (BodyStats::broken_beams, src/phys/softbody.h:589), shocks that broke
 * included. */
int bl_body_broken_beams(bl_ctx* ctx, uint32_t id, int64_t* out);

/* A body's slide nodes as the step left them (SlideNode::seg, SlideNode::broken; the other fields as added). */
int bl_read_slides(bl_ctx* ctx, uint32_t id, bl_slide* slides_out);

/* A body's frames and joints as the step left them (either pointer may be null). */
int bl_read_joints(bl_ctx* ctx, uint32_t id, bl_frame* frames_out, bl_joint* joints_out);

/* A body's wheels as the step left them and its SoftBody::wheel_speed, wheel_spin (either pointer may be null). */
int bl_read_wheels(bl_ctx* ctx, uint32_t id, bl_wheel* wheels_out, float speed_spin[2]);

/* A controller of the body's wheels (a vehicle's pre_substep hook, src/phys/softbody.h:807, run by base_forces once a
 * substep before the forces, src/phys/world.cpp:3260, setting Wheel::torque and Wheel::brake): from the next bl_step
 * on, every substep starts with wheel k's torque = torque[k] and brake = brake[k] (n_wheels values each). Null
 * pointers remove it: the wheels keep what the step leaves (a torque is spent in one substep, a brake stays). */
int bl_set_wheel_controls(bl_ctx* ctx, uint32_t id, const float* torque, const float* brake);

/* Counters since bl_create (or the last bl_reset_stats). */
int bl_get_stats(bl_ctx* ctx, bl_stats* out);
int bl_reset_stats(bl_ctx* ctx);

/* Execution choices; bl_get_tuning returns the ones in force. */
int bl_set_tuning(bl_ctx* ctx, const bl_tuning* t);
int bl_get_tuning(bl_ctx* ctx, bl_tuning* t);

/* Times `substeps` substeps (already resident inputs) with HIP events on the context's stream. *ms = elapsed. */
int bl_time_step(bl_ctx* ctx, int32_t substeps, float* ms);

/* Times one frame of `substeps` with an event pair around every kernel launch of the substeps (no graph capture):
 * *kernel_ms = summed duration of the substep kernels, *total_ms = the frame on the stream, *launches = launches.
 * A frame with inter-body contacts times the force-phase kernel (beams and gather) of each substep. */
int bl_time_kernels(bl_ctx* ctx, int32_t substeps, float* total_ms, float* kernel_ms, int32_t* launches);

/* ---------------------------------------------------------------------------------------------------------------
 * FEM frames: the implicit co-rotational step of a frame body - shell triangles (FemFrame's triangles:
 * src/phys/frame_fem.h:154-181) and Timoshenko members (FemFrame::elems: src/phys/frame_fem.h:197-224, eval_forces
 * src/phys/frame_fem.cpp:1896-2070) - FemFrame::solve (src/phys/frame_fem.cpp:2514-3027). Replaces SoftBody::fem for a
 * body that may deform and does not break (SoftBody::allow_break off: members and triangles yield - the hinges, bars
 * and twists of members, the membranes and bending of shells - but are never split, torn or let go), with no mounts
 * or contacts; its implicit step may span several substeps, the held substeps between answered through its
 * factorization. The system - 6x6 blocks per node pair, FP64 - is factored each step by a supernodal
 * (multifrontal) Cholesky over a nested-dissection order, its dense fronts on FP64 matrix cores; a step that unloads
 * a yielding member is assembled, factored and solved again with its elastic tangent (FemFrame::solve_impl's passes,
 * src/phys/frame_fem.cpp:2959-2996, on their budget, :1863-1864).
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct bl_fem bl_fem;

/* == ShellSection (src/phys/frame_fem.h:117-149), the fields the step reads */
typedef struct bl_shell_section {
    float E, nu, rho, t;
    float drill;   /* drilling spring share of G t A per corner (default 0.05) */
    float damping; /* stiffness-proportional damping (s, default 3e-4) */
    float yield;   /* yield stress (Pa; 0: stays elastic) */
} bl_shell_section;

/* == FrameSection (src/phys/frame_fem.h:90-112), the fields the step reads */
typedef struct bl_frame_section {
    float E, G, rho;
    float A, Iy, Iz, J;     /* area, second moments about the member's y and z, torsion constant */
    float As_y, As_z;       /* shear areas (Timoshenko; 0: no shear deformation) */
    float damping;          /* stiffness-proportional damping (s) */
    float axial;            /* factor on the axial stiffness */
    float joint_k;          /* BL_FJ_ELASTIC ends: rotational stiffness (N m/rad) */
    float Np, Mp, Tp;       /* plastic axial force (N), moment and torque (N m); Np 0: stays elastic */
    float hardening;        /* the plastic moment's growth over a hinge's first radian (default 0.25) */
    float half;             /* half the outer size (m, default 0.02): the torsional damage's lever */
} bl_frame_section;

/* == FrameJoint (src/phys/frame_fem.h:185-193): what a member's end frees at its node */
enum { BL_FJ_RIGID = 0, BL_FJ_BALL = 1, BL_FJ_HINGE_V = 2, BL_FJ_HINGE_H = 3, BL_FJ_SWIVEL = 4, BL_FJ_ELASTIC = 5 };

/* == FrameElement (src/phys/frame_fem.h:197-224), as FemFrame::add_element takes it */
typedef struct bl_frame_member {
    uint32_t a, b;          /* its end nodes */
    uint16_t section;       /* index into bl_fem_desc::sections */
    uint8_t end_a, end_b;   /* BL_FJ_* of each end */
} bl_frame_member;

typedef struct bl_fem_desc {
    uint32_t n_nodes, n_tris;
    const float* p;         /* 3 per node: rest positions (the body's Node::p) */
    const uint32_t* tris;   /* 3 per triangle (FrameTri::n), counter-clockwise about the rest normal */
    const uint8_t* fixed;   /* per node: 1 = pinned (inv_mass 0) */
    const float* mass;      /* per node (the body's Node::mass), or null: the triangles' thirds of their section's rho t
                               per area (ShellMesher::tri; a body with members passes its masses) */
    bl_shell_section section;
    uint32_t n_members, n_sections;
    const bl_frame_member* members;   /* co-rotational Timoshenko members (FemFrame::elems), may be null */
    const bl_frame_section* sections;
} bl_fem_desc;

typedef struct bl_fem_report {
    double residual;        /* || A x - b || / || b || of the step's solve, FP64 on the device */
    double factor_flops;    /* floating-point operations of the numeric factorization */
    float assemble_ms, factor_ms, solve_ms, update_ms; /* device time of the phases (HIP events) */
    int32_t failures;       /* a pivot block not positive definite (the step then leaves the state as it was) */
    int32_t supernodes, fronts_max; /* the symbolic factorization: supernodes, largest front (scalar rows) */
    int64_t factor_nnz;     /* nonzeros of L (scalars) */
    int32_t passes;         /* systems assembled, factored and solved this step (1 + the passes again) */
    int32_t repasses;       /* passes again of components of 60 nodes or more (FemFrame::repasses, this step) */
} bl_fem_report;

/* FemFrame::finalize for the triangles (src/phys/frame_fem.cpp:259-276, 3075-3086) and the symbolic analysis. */
int bl_fem_create(bl_fem** out, const bl_fem_desc* d);
void bl_fem_destroy(bl_fem* f);
/* One substep of h of the shell body (theta, numerical dissipation, gravity g) as World::simulate_island runs a frame
 * body stepping every substep (src/phys/world.cpp:3265-3269, 3573-3580, 3812-3822, 3876-3886): element pass,
 * assembly, factorization, substitution, the update (the new velocities as forces, src/phys/frame_fem.cpp:3000-3027),
 * SoftBody::integrate_nodes and FemFrame::sync_positions. loads: 3 per node (N, host memory) or null. A factorization
 * that fails is counted (rep->failures, bl_fem_get_nodes) and the gathered forces act explicitly, as the reference
 * does; BL_ERR_STATE only for the device's own trouble (a substitution's wait timed out). */
int bl_fem_step(bl_fem* f, const float* loads, float h, float theta, float dissipation, const float g[3],
                bl_fem_report* rep);
/* One implicit step spanning h of which the first substep, of `step`, is integrated: a frame stepping every h / step
 * substeps (WorldSettings::frame_every; FemFrame::solve(b, h, step, ...), src/phys/frame_fem.cpp:2514-2545). The
 * forces are taken for the whole of h; the held impulse and the prediction of the smooth forces for the substeps
 * between are kept (:1026-1042), the update and integration are over `step`. The world passes the dissipation per
 * step divided by the substeps the step spans (src/phys/world.cpp:3247-3252). bl_fem_step is this with step = h. */
int bl_fem_step_span(bl_fem* f, const float* loads, float h, float step, float theta, float dissipation,
                     const float g[3], bl_fem_report* rep);
/* A held substep of `step` between two implicit steps (FemFrame::held_begin / held_component / held_end,
 * src/phys/frame_fem.cpp:2342-2512): the forces (gravity g, the loads) and moments from outside beyond the last step's
 * prediction through that step's factorization, the velocity changes as forces, then integrate_nodes and
 * sync_positions; without a factorization (none yet, or the last one failed) FemFrame::hold (:2321-2340). */
int bl_fem_step_held(bl_fem* f, const float* loads, float step, const float g[3]);
/* The state: positions (double, 3 per node), v, w (3 per node), q (4 per node, x y z w); any pointer may be null. */
int bl_fem_get_state(bl_fem* f, double* x, float* v, float* w, float* q);
int bl_fem_set_state(bl_fem* f, const double* x, const float* v, const float* w, const float* q);
/* Moments from outside on the nodes for the next step (FemFrame::torque: N m, 3 per node, host memory), added to any
 * given before; the step uses them up. */
int bl_fem_add_torques(bl_fem* f, const float* torques);
/* The body's node records after the last step (SoftBody::nodes: p, inv_mass, v, mass), the forces the step handed the
 * integrator (SoftBody::force: dv m / step of a free node, the gathered forces where the solve was skipped; 3 per
 * node) and the counters {FemFrame::solve_failures, FemFrame::clamps, NaN velocity resets}
 * (src/phys/frame_fem.h:342-343, frame_fem.cpp:3000-3027, softbody.cpp:769-795). Any pointer may be null. */
int bl_fem_get_nodes(bl_fem* f, bl_node* nodes, float* force, int64_t counters[3]);
/* The last step's linear system over the free nodes in node order: dense (6 nf)^2 row-major A (lower triangle and
 * its mirror), rhs and solution (6 nf each); *nf the free nodes (testing; small meshes). */
int bl_fem_last_system(bl_fem* f, double* A, double* rhs, double* x, uint32_t* nf);
/* Solves A x = b with the last step's factorization (b, x: 6 per free node, node order, host memory). */
int bl_fem_solve(bl_fem* f, const double* b, double* x);
/* Solves A x = b with the last step's factorization, allocating nothing: b and x are DEVICE pointers (6 per free
 * node, node order, FP64), and the substitutions' scratch (per-front vectors, block flags, y) is a file-scope device
 * array of 32 MiB per GPU, held until the process ends (fetched once per device with hipGetSymbolAddress at
 * bl_fem_create). No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Runs on the engine's stream and waits for
 * it. Every call on one GPU shares that GPU's array: calls that overlap on one GPU race. A system whose scratch
 * exceeds 32 MiB is rejected (BL_ERR_ARG); *scratch_bytes (may be null) receives what it needs, and with b or x
 * null the call only reports that. Results are those of bl_fem_solve, bit for bit. */
int bl_fem_solve_noalloc(bl_fem* f, const double* b, double* x, uint64_t* scratch_bytes);
/* The element stiffness of triangle `tri` as the device built it (18 x 18 row-major, FP64). */
int bl_fem_tri_k(bl_fem* f, uint32_t tri, double* K);
/* The plastic state (FrameTri, src/phys/frame_fem.h:154-181: per triangle BL_FEM_TRI_STATE floats - the rest shape
 * X[6], the corners' plastic rest rotations th0[9], the rest frames r0[12] (x y z w), dmg, util; FrameElement,
 * :197-224: per member BL_FEM_MEM_STATE floats - up, tp, pa[2], pb[2], ba[2], bb[2], bt, damage, dmg_a, dmg_b, util, N,
 * qa[4], qb[4], then the tangent's yield, unload and elastic_hold bits, FemFrame::Tangent) and the counters
 * {FemFrame::repasses, passes_} so far. Any pointer may be null. */
#define BL_FEM_TRI_STATE 29
#define BL_FEM_MEM_STATE 27
int bl_fem_get_plastic(bl_fem* f, float* tris, float* members, int64_t counters[2]);

/* ---------------------------------------------------------------------------------------------------------------
 * The learned warm start and preconditioner of the implicit step's solve (no counterpart in the reference: the new
 * part of the FEM path). A 4-layer message-passing network over the free nodes (hidden 64, bfloat16 products with
 * FP32 sums on the matrix cores) predicts the step's solution and a per-dof scaling s of the block-Jacobi
 * preconditioner (S D^-1 S). PCG starts from the prediction; the supernodal Cholesky is the fallback when PCG misses
 * the tolerance within max_iter iterations, and the yardstick of the prediction's error. The network is trained on the
 * engine's own (system, solution) pairs: a Cholesky step (bl_fem_step), then bl_gnn_train_step.
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct bl_gnn bl_gnn;

enum { BL_WARM_ZERO = 0, BL_WARM_PREVIOUS = 1, BL_WARM_NETWORK = 2 };
/* BL_PRECOND_FACTOR: the engine's last Cholesky factor, of an earlier step's system (factor reuse between solves, as
 * the reference's held steps solve with the last factorization, src/phys/frame_fem.cpp:2342-2440). The fallback
 * refreshes it when PCG misses tol within max_iter; with no factor yet, the step is a Cholesky step. */
enum { BL_PRECOND_BLOCK_JACOBI = 0, BL_PRECOND_NETWORK = 1, BL_PRECOND_FACTOR = 2 };

typedef struct bl_pcg_opts {
    int32_t max_iter;    /* K: PCG iterations before the fallback */
    int32_t warm;        /* BL_WARM_* */
    int32_t precond;     /* BL_PRECOND_* */
    int32_t fallback;    /* 1: the Cholesky when PCG misses tol within max_iter */
    int32_t check_error; /* 1: also the Cholesky solution, to report the warm start's and PCG's errors (not timed use) */
    int32_t pad;
    double tol;          /* relative residual || b - A x || / || b || */
} bl_pcg_opts;

typedef struct bl_pcg_report {
    int32_t iterations, converged, fallback_used, pad;
    double residual;       /* of the solution the update applied (FP64, on the device) */
    double warm_error;     /* || x0 - x* || / || x* || (check_error) */
    double pcg_error;      /* || x_pcg - x* || / || x* || (check_error) */
    float assemble_ms, gnn_ms, pcg_ms, fallback_ms, update_ms, total_ms;
} bl_pcg_report;

int bl_gnn_param_count(void);
/* The network with the given weights (bl_gnn_param_count floats, in the order of the layers). */
int bl_gnn_create(bl_gnn** out, const float* params, uint32_t n);
void bl_gnn_destroy(bl_gnn* g);
int bl_gnn_get_params(bl_gnn* g, float* out, uint32_t n);
int bl_gnn_set_params(bl_gnn* g, const float* params, uint32_t n); /* (resets the optimizer) */
/* The loss of the engine's current system against its current solution, and its gradients (n floats, or null). */
int bl_gnn_gradients(bl_gnn* g, bl_fem* f, float* grads, uint32_t n, double losses[2]);
/* One Adam step on the engine's current (system, solution). */
int bl_gnn_train_step(bl_gnn* g, bl_fem* f, float lr, double losses[2]);
/* The warm start x0 and the scales s (6 per free node each, host memory, either may be null) for the current system. */
int bl_gnn_infer(bl_gnn* g, bl_fem* f, double* x0, double* s);
/* One implicit step solved by PCG (the network's warm start / preconditioner as opts choose), the Cholesky as the
 * fallback; g holds the solve's workspace and is required. */
int bl_fem_step_pcg(bl_fem* f, bl_gnn* g, const float* loads, float h, float theta, float dissipation, const float gr[3],
                    const bl_pcg_opts* opts, bl_pcg_report* rep);

#ifdef __cplusplus
}
#endif

#endif /* BEAMLAB_GPU_H */
