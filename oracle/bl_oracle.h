/*
 * TEST INFRASTRUCTURE - the oracle. A plain-C CPU restatement of BeamLab's soft-body step (zenkovich/beamlab,
 * src/phys), written from the reference's semantics line by line (each function cites the lines it follows).
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, as the checker; the product
 * (beamlab_amd/) never links or calls it.
 *
 * Parity pinning of each part (see tests/):
 *   - beam force / yield / break, integration, ground contact, stabilize, finalize, box setup: bit-exact against
 *     the compiled reference (oracle/_ref) on seeded lattices and against closed forms (2-node spring, ballistic
 *     flight, the 400 m/s clamp, NaN reset, the stability budget).
 */
#ifndef BL_ORACLE_H
#define BL_ORACLE_H

#include "beamlab_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

orc_world* orc_create(const bl_settings* s);
void orc_destroy(orc_world* w);
int orc_set_statics(orc_world* w, const bl_box* boxes, int n);
int orc_set_cylinders(orc_world* w, const bl_cylinder* c, int n);
int orc_set_terrain(orc_world* w, int nx, int nz, float cell, const float* origin, const float* heights,
                    const uint8_t* surfaces);
int orc_add_body(orc_world* w, const bl_body_desc* d);
int orc_step(orc_world* w, int substeps);
int orc_read_body(orc_world* w, uint32_t id, bl_node* nodes, bl_beam* beams, int64_t* broken);
int orc_write_body(orc_world* w, uint32_t id, const bl_node* nodes, const bl_beam* beams);
int orc_stats(orc_world* w, bl_stats* out);
int orc_read_slides(orc_world* w, uint32_t id, bl_slide* out);
int orc_read_joints(orc_world* w, uint32_t id, bl_frame* frames, bl_joint* joints);
int orc_read_wheels(orc_world* w, uint32_t id, bl_wheel* wheels, float* speed_spin);
/* a controller setting each wheel's torque and brake at every substep (null: none) */
int orc_set_wheel_controls(orc_world* w, uint32_t id, const float* torque, const float* brake);
/* the fast bodies' candidate pairs (World::fast_pairs), summed over the rebuilds */
int64_t orc_fast_pairs(const orc_world* w);
/* summed over every rebuild (every island): hashed nodes, triangles queried, NT before and after the near filter,
 * NC, CT */
void orc_rebuild_counts(const orc_world* w, int64_t* out);

int orc_finalize(const bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb);
int orc_stabilize(const bl_node* nodes, uint32_t n, bl_beam* beams, uint32_t nb, float dt, float kb, float db,
                  int* changed);
int orc_box_from_quat(const float* c, const float* half, const float* q, uint32_t surface, bl_box* out);

/* primitive_collision on its own (src/phys/static_world.cpp:8-51): out = force */
void orc_primitive_collision(const float F[3], const float vel[3], float mass, const float normal[3], float dt,
                             int surface, float pen, float friction_coef, float push_max, float slop,
                             float restitution, float out[3]);

#ifdef __cplusplus
}
#endif

#endif
