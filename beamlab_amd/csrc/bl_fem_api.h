// The FEM engine's phases and device arrays, for the learned warm start and the PCG solve (bl_gnn.hip). Internal to
// the library: the C ABI is include/beamlab_gpu.h.
#pragma once
#include "beamlab_gpu.h"
#include <hip/hip_runtime.h>
#include <cstdint>

namespace blf {

struct FemView {
    uint32_t nf;             // free nodes (the system's block rows, in node order)
    const double* D;         // 36 per free node: the diagonal block (row-major; its lower triangle is the system's)
    const double* O;         // 36 per pair: the lower block (row node's rows, column node's columns)
    const uint4* pairs;      // per pair: row free index, column free index, ..
    uint32_t npairs;
    const uint32_t* adj_off; // per free node: its pairs (adj_off[i] .. adj_off[i + 1])
    const uint2* adj;        // (pair, other free index | 1 << 31 when this node is the pair's row node)
    const double* rhs;       // 6 per free node
    double* x;               // 6 per free node: the step's solution (what the update applies)
    const double* xprev;     // 6 per free node: the previous step's solution
    const float* p0;         // 3 per free node: rest position
    hipStream_t st;
    uint64_t serial;         // unique per engine in the process
};

FemView fem_view(bl_fem* f);
// the element pass and the assembly of a step (the previous solution kept as xprev)
int fem_assemble(bl_fem* f, const float* loads, float h, float theta, float dissipation, const float g[3],
                 float* ms, float step = 0.0f);
// x := A^-1 rhs by the supernodal Cholesky; *fail: bit 1 a pivot not positive, bit 2 a wait timed out
int fem_cholesky(bl_fem* f, float* ms_factor, float* ms_solve, int* fail);
// whether the factor buffer holds a successful factorization (of this step's system or an earlier one)
bool fem_has_factor(bl_fem* f);
// a copy of the factor, put back by fem_restore_factor (around a factorization that must leave it as it was)
int fem_keep_factor(bl_fem* f);
int fem_restore_factor(bl_fem* f);
// z := (L L^T)^-1 r with that factor, enqueued on the engine's stream; every launch returns at once while *skip is
// set on the device (r and z 6 per free node)
int fem_precond(bl_fem* f, const double* r, double* z, const int* skip);
// the substitutions' error word (nonzero: a wait timed out), after the stream drains
int fem_solve_error(bl_fem* f, int* err);
// || A x - rhs || / || rhs || (waits for the stream)
int fem_residual(bl_fem* f, const double* x, double* rel);
// y := A x (enqueued on the engine's stream)
void fem_spmv(bl_fem* f, const double* x, double* y);
// the end of the substep from f's solution x: the update (velocities as forces, clamps), the integration, the position
// sync; explicit_step: the factorization failed, the gathered forces act instead
int fem_update(bl_fem* f, float* ms, bool explicit_step = false);
// the factorization's figures for a report
void fem_report_static(bl_fem* f, bl_fem_report* rep);
// whether a section of the body yields (a shell yield stress or a member's plastic limits)
bool fem_plastic(const bl_fem* f);

} // namespace blf
