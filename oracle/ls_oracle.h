/*
 * ls_oracle.h — CPU FP64 oracle of the LeachingSim per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke entry point and the
 * cpu_baseline leg of bench.py may load this library, and only as the values
 * the GPU path is compared against, or as the timed CPU baseline.  The
 * product path (liblsgpu.so) never links or calls it.
 *
 * The oracle restates src/Simulation.cpp of javalosp/LeachingSim for a single
 * domain (no MPI ranks) with these semantic decisions: one-sided differences
 * only at the global faces, saturation kept at inactive voxels, solid_density
 * 1.0, optional virtual Air planes on the i faces (LS_BC_ZFACES), floored
 * pressure rows (|diag| < 1e-16) handled as lagged Dirichlet nodes and
 * post-solved from their own row, leaching in exact sequential order.
 * It uses the same parameter, field, phase and report types as the GPU
 * library (include/lsgpu.h) so the two can be driven by one test sequence.
 */
#ifndef LS_ORACLE_H_
#define LS_ORACLE_H_

#include "../include/lsgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lso_ctx lso_ctx;

int lso_create(lso_ctx** out, int ni, int nj, int nk, const ls_params* params);
void lso_destroy(lso_ctx* ctx);
const char* lso_last_error(const lso_ctx* ctx);
int lso_set_params(lso_ctx* ctx, const ls_params* params);
int lso_upload(lso_ctx* ctx, int field, const void* host, size_t elem_bytes);
int lso_download(lso_ctx* ctx, int field, void* host, size_t elem_bytes);
int lso_init_state(lso_ctx* ctx, long long* initial_sulphide_count);
int lso_step(lso_ctx* ctx, double t, double dt, ls_step_report* report);
int lso_phase(lso_ctx* ctx, int phase, double t, double dt, ls_step_report* report);
int lso_snapshot(lso_ctx* ctx);
int lso_rollback(lso_ctx* ctx);
void lso_set_threads(int nthreads);
/* Linear solvers: 0 (the default) Jacobi-PCG on the reduced pressure
 * system to press_rtol on ||b - A x|| / ||b|| and BiCGStab for transport;
 * 1 the reference's KSP for pressure and transport: restarted GMRES(30)
 * with left Jacobi (pressure on the full reference-form system), stopping
 * on the preconditioned residual norm <= rtol ||D^-1 b||, each solve issued
 * twice (src/Simulation.cpp:146-157, 804-855, 1247-1275). */
int lso_set_ksp(lso_ctx* ctx, int which);

/* The reference's full pressure system (src/Simulation.cpp:658-798) for the
 * current state, as assembled by MatSetValue/VecSetValue (ADD_VALUES), in
 * COO form: rows[nnz], cols[nnz], vals[nnz], rhs[N].  Call with rows == NULL
 * to get nnz only.  Under the z-face boundary mode the virtual-plane links are
 * included (diagonal and right-hand side). */
long long lso_pressure_system(lso_ctx* ctx, double t, int* rows, int* cols,
                              double* vals, double* rhs);
/* The reference's transport system (src/Simulation.cpp:966-1237) for the
 * given field (LS_F_CONC or LS_F_ACID), COO form as above.  Like the
 * reference, this zeroes the flux at Rock and Precipitate voxels. */
/* the implicit saturation system of the current state (COO + rhs), as
 * assembled by src/Simulation.cpp:1858-1964 */
long long lso_saturation_system(lso_ctx* ctx, double dt, int* rows, int* cols, double* vals,
                                double* rhs);
long long lso_transport_system(lso_ctx* ctx, double dt, int field, int* rows,
                               int* cols, double* vals, double* rhs);
/* y = A x with the reduced pressure operator of the last pressure phase
 * (rows outside the solved set return 0). */
int lso_apply_pressure(lso_ctx* ctx, const double* x, double* y);
/* Per-voxel class of the last pressure phase: 0 Dirichlet, 1 solved (R),
 * 2 floored (F). */
int lso_pressure_classes(lso_ctx* ctx, int8_t* out);

#ifdef __cplusplus
}
#endif

#endif /* LS_ORACLE_H_ */
