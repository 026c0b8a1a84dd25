/* Test infrastructure only: the FP64 restatement of the reference's shell FEM step used as the checker of the HIP
 * FEM path (see fem_oracle.c). Never linked into the product. */
#ifndef BL_FEM_ORACLE_H
#define BL_FEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_fem orc_fem;

/* A member section's fields (FrameSection, src/phys/frame_fem.h:90-112, the ones the step reads): their order in a
 * section. Np, Mp, Tp 0: the member stays elastic. */
enum {
    ORC_SEC_E, ORC_SEC_G, ORC_SEC_RHO, ORC_SEC_A, ORC_SEC_IY, ORC_SEC_IZ, ORC_SEC_J, ORC_SEC_ASY, ORC_SEC_ASZ,
    ORC_SEC_DAMPING, ORC_SEC_AXIAL, ORC_SEC_JOINT_K, ORC_SEC_NP, ORC_SEC_MP, ORC_SEC_TP, ORC_SEC_HARDENING,
    ORC_SEC_HALF, ORC_FRAME_SECTION_FIELDS
};
/* A shell section's fields (ShellSection, src/phys/frame_fem.h:117-149): E, nu, rho, t, drill, damping, yield (Pa; 0:
 * elastic) */
enum { ORC_SHELL_FIELDS = 7 };
/* The plastic state read back per triangle (FrameTri: X[6], th0[9], r0[12] (x y z w), dmg, util) and per member
 * (FrameElement: up, tp, pa[2], pb[2], ba[2], bb[2], bt, damage, dmg_a, dmg_b, util, N, qa[4], qb[4]; then the
 * tangent's yield, unload and elastic_hold bits, FemFrame::Tangent) */
enum { ORC_TRI_STATE = 29, ORC_MEM_STATE = 27 };

/* the element stiffness of one triangle (tri_local_k, src/phys/frame_fem.cpp:1524-1579), 18 x 18 row-major */
void orc_fem_tri_local_k(const float X[6], float E, float nu, float t, float drill, double* K);
/* the rest shape of a triangle from its corners (FemFrame::finalize, src/phys/frame_fem.cpp:259-276): X[6] */
int orc_fem_rest_shape(const float p[9], float X[6]);

/* A plate of nu x nv cells (ShellMesher::grid, src/phys/fem_shell.cpp:34-46) of `size` metres in the x-z plane,
 * centred on the origin at height y0, its border nodes clamped. Section: E, nu, rho, t, drill, damping (elastic). */
orc_fem* orc_fem_plate(int nu, int nv, float size, float y0, const float section[6]);
/* A frame body: nodes (p, mass as the body has it, pinned flags), shell triangles of one section (ORC_SHELL_FIELDS
 * floats), members (4 words each: a, b, section, end_a | end_b << 8, the ends' FrameJoint) of sections of
 * ORC_FRAME_SECTION_FIELDS floats. FemFrame::finalize and node_inertia. The body may deform (SoftBody::allow_deform)
 * and does not break (allow_break off: a yielding member or triangle is never split, torn or let go). */
orc_fem* orc_fem_create(uint32_t nn, const float* p, const float* mass, const uint8_t* fixed, uint32_t nt,
                        const uint32_t* tris, const float shell[ORC_SHELL_FIELDS], uint32_t nm,
                        const uint32_t* members, uint32_t nsec, const float* sections);
/* the plastic state (ORC_TRI_STATE floats per triangle, ORC_MEM_STATE per member; either may be null) and the
 * counters {FemFrame::repasses, passes_} */
void orc_fem_plastic(const orc_fem* f, float* tris, float* members, int64_t counters[2]);
/* moments from outside on the nodes for the next step (FemFrame::torque, 3 per node), added to any set before */
void orc_fem_add_torques(orc_fem* f, const float* torques);
void orc_fem_destroy(orc_fem* f);
uint32_t orc_fem_nodes(const orc_fem* f);
uint32_t orc_fem_tris(const orc_fem* f);
/* node arrays: tri corners (3 per triangle), positions (3 per node, double), fixed flags, mass, inertia */
void orc_fem_topology(const orc_fem* f, uint32_t* tri_nodes, double* xd, uint8_t* fixed, float* mass, float* inertia);
/* the state: v, w (3 per node), q (4 per node: x y z w) */
void orc_fem_get_state(const orc_fem* f, double* xd, float* v, float* w, float* q);
void orc_fem_set_state(orc_fem* f, const double* xd, const float* v, const float* w, const float* q);
/* the nodes pinned (1) or free (0), their inverse masses again (ShellMesher::finish) */
void orc_fem_set_fixed(orc_fem* f, const uint8_t* fixed);
/* the body's node records {p, inv_mass, v, mass} (8 floats each, as bl_node), the forces the last step handed the
 * integrator (3 per node), counters {solve failures, clamps, NaN resets}; any pointer may be null */
void orc_fem_get_nodes(const orc_fem* f, float* nodes, float* force, int64_t counters[3]);

/* One substep of h of the plate body: element pass, gather, assembly, block Cholesky (chol6 / right_solve_lt /
 * sub_abt in the reference's right-looking order over the nodes' natural order, dense), substitution, the update
 * (velocities as forces), the body's integration and the frame's position sync. loads: 3 per node
 * (N), added to gravity m g. If `A` is non-null it receives the assembled system of the free nodes in the natural
 * order: dense (nf*6)^2 row-major, `rhs` nf*6, `x` the solution; `free_map` the free nodes. Returns 0, or -1 when a
 * pivot block is not positive definite. */
int orc_fem_step(orc_fem* f, const float* loads, float h, float theta, float dissipation, const float g[3], double* A,
                 double* rhs, double* x, uint32_t* free_map, uint32_t* nfree);

/* One implicit step spanning h (a frame stepping every h / step substeps: WorldSettings::frame_every) of which the
 * first substep, of `step`, is integrated: the forces taken for the whole of h, the held impulse and the prediction
 * for the held substeps kept (frame_fem.cpp:1026-1042), the update over `step`. orc_fem_step is this with step = h.
 * The caller passes the dissipation per step as the world does (frame_dissipation / the steps it spans). */
int orc_fem_step_span(orc_fem* f, const float* loads, float h, float step, float theta, float dissipation,
                      const float g[3], double* A, double* rhs, double* x, uint32_t* free_map, uint32_t* nfree);
/* A held substep of `step` between two implicit steps (FemFrame::held_*, or FemFrame::hold without a factorization) */
int orc_fem_held(orc_fem* f, const float* loads, float step, const float g[3]);

/* the reference's dense block Cholesky on an n x n block SPD matrix (row-major (6n)^2): solves A x = b in place */
int orc_fem_block_solve(int n, double* A, double* b);

#ifdef __cplusplus
}
#endif
#endif
