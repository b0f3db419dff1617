// ls_context.h — device-resident state of one simulation context and the
// internal entry points the C ABI (ls_api.hip) dispatches to.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "ls_common.h"

// Device scalars of one CG solve (kept on the device so iterations run
// without host round trips; the host polls `done` once per batch).
struct CgScalars {
    double rr;        // r^T D^-1 r  (scaled residual dot)
    double pAp;
    double alpha;
    double beta;
    double wrr;       // ||b - A x||^2 of the unscaled system
    double bnorm2;    // ||b||^2
    double tol2;      // (rtol * ||b||)^2, or the inner target for FP32 solves
    double relres;
    double rr_prev;   // r.r before the last update (for beta after a residual replacement)
    int iter;
    int maxit;
    int done;         // 1 = stop (converged or iteration cap)
    int converged;
    int xlive;        // 1 = this iteration's x += alpha p is still owed (the
                      // direction kernel applies it)
};

// Device scalars of the two-RHS BiCGStab (conc, acid).
struct BicgScalars {
    double rho[2], alpha[2], omega[2], beta[2];
    double r0v[2], ss[2], tt[2], ts[2], rr[2], r0r[2];
    double bnorm2[2];
    double tol2[2];
    int s_conv[2];
    int done[2];
    int converged[2];
    int iter[2];
    int maxit;
    int all_done;
};

// Device scalars of the flexible CG with the learned preconditioner.
struct FcgScalars {
    double beta, alpha, gamma, scale_c;
    double pq, pq_prev, pr, rr, rr_prev;
    double tol2;
    int iter, maxit, done, converged, broke, first;
};

struct StepScalars {
    double sat_max;           // max |div q| + evap over Pore voxels
    double sat_dt_sub_over_phi;
    int sat_nsub;
    double leach_max_rate;
    int leach_nsub;
    double leach_dt_sub;
    int leach_needed;         // some reactive voxel can change this step
    int leached;
    double sums[9];           // pmin pmax cmin cmax amin amax  sum c, sum a, sum s
    double n_floored;
    double nsul;
    double sat_lo, sat_hi;    // raw range of an implicit saturation solve
};

struct ls_ctx {
    int ni = 0, nj = 0, nk = 0;
    size_t N = 0;
    int label_bytes = 1;
    ls_params P{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    double log_k_rock = 0.0, log_k_pore = 0.0;   // std::log10 on the host

    // --- state (src/Simulation.h:143-158) ---
    uint8_t* img = nullptr;
    uint8_t* img2 = nullptr;     // double buffer for label updates
    float* frac = nullptr;
    double *inv = nullptr, *s = nullptr, *p = nullptr, *conc = nullptr, *acid = nullptr;
    double *qx = nullptr, *qy = nullptr, *qz = nullptr;
    // --- snapshot (src/Simulation.h:71-84) ---
    uint8_t* b_img = nullptr;
    float* b_frac = nullptr;
    double *b_inv = nullptr, *b_s = nullptr, *b_p = nullptr, *b_conc = nullptr,
           *b_acid = nullptr, *b_qx = nullptr, *b_qy = nullptr, *b_qz = nullptr;
    // --- properties ---
    double *perm = nullptr, *pc = nullptr;
    // --- phase-scoped pool ---
    // One allocation holds either the pressure phase's buffers (system, CG
    // vectors, scratch pw) or the transport phase's (rows, right-hand sides,
    // BiCGStab vectors), laid over the same bytes: neither phase reads what
    // the other leaves, and every phase that uses the pool starts by writing
    // what it reads.  A transport or implicit-saturation solve therefore
    // ends press_assembled.
    void* pool = nullptr;
    size_t pool_bytes = 0;
    size_t pool_press_bytes = 0, pool_trans_bytes = 0;
    int pool_key = -1;          // layout the pool was cut for
    int n_pw = 0;               // pressure scratch vectors (N doubles each)
    double* pw[2 * LS_WARM_MAX + 2] = {};
    // transport form of the layout: stored rows (nk even), red-black Schur
    bool t_stored = false, t_schur = false;
    // the pool holds the rows and vectors of the last transport solve (no
    // pressure-phase use since)
    bool trans_valid = false;
    // --- pressure system ---
    int8_t* cls = nullptr;                           // 0 Dirichlet, 1 solved, 2 floored
    double* pD = nullptr;                            // Dirichlet values
    double* sd = nullptr;                            // sqrt(d) on solved rows, else 0
    double* bhat = nullptr;                          // D^-1/2 b
    double* c64[3] = {nullptr, nullptr, nullptr};    // D^-1/2 C D^-1/2 faces, FP64
    float* c32[3] = {nullptr, nullptr, nullptr};     // same, FP32
    // CG vectors (scaled variables)
    double *x64 = nullptr, *r64 = nullptr, *p64 = nullptr, *ap64 = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *ap32 = nullptr;
    // --- transport (two right-hand sides) ---
    // r, r0, p, v, s, t, ph, sh for conc and acid: N/2 each in the Schur
    // form (black unknowns; no ph and sh there: null), else N
    double* tv[2][8] = {};
    double* tb[2] = {};      // right-hand sides (N)
    double* tinvd = nullptr; // Jacobi inverse diagonal (N; not in the Schur form)
    // Schur form: the scaled black unknowns y = D_b x_b, red values of the
    // last red pass
    double* txs[2] = {};
    double* twr[2] = {};
    // mixed-precision Schur form (conc alone, acid known zero): FP32 copies
    // of the off-diagonal rows over acid's vectors, and conc's FP32 cycle
    // vectors r, rh, p, v, s, t, e, w laid two to an FP64 vector over its
    // rh, p, v, s (tf[1] stays null)
    bool t_mixed = false;
    float* trows32 = nullptr;
    float* tf[2][8] = {};
    BicgScalars* bicg32 = nullptr;
    BicgScalars* h_bicg32 = nullptr;
    int t_cycles = 0, t_fallback = 0;   // of the last mixed-precision solve
    bool t_rows32_valid = false;        // the FP32 rows belong to the last solve
    double* taux[2] = {};    // implicit saturation: its source field and idle system (N)
    double* arows = nullptr; // stored transport rows (nk even)
    // the black half of conc at the last transport solves, newest first (N/2
    // each, allocated on first use), the time from each to the next newer
    // one (thist_dt[0]: to the current field), and how many belong to the
    // current state's history
    double* thist[3] = {nullptr, nullptr, nullptr};
    double thist_dt[3] = {0.0, 0.0, 0.0};
    int tprev_n = 0;
    // --- reductions ---
    double* partials = nullptr;     // MAX_PARTIALS * 8 values
    CgScalars* cg = nullptr;
    BicgScalars* bicg = nullptr;
    StepScalars* ss = nullptr;
    FcgScalars* fcg = nullptr;
    FcgScalars* h_fcg = nullptr;
    // learned preconditioner: U-Net on the 2x restricted grid
    struct UNet* net = nullptr;
    int net_dims[3] = {0, 0, 0};
    float* rc = nullptr;            // restricted residual (coarse, FP32)
    std::vector<float> net_weights;  // host copy, set by ls_set_network
    int net_failures = 0;            // consecutive probes where the network lost
    int net_skip = 0;                // solves left that go straight to mixed PCG
    // the bar of the learned solver: the residual curve (ms of wall time,
    // decades of reduction) of the last converged mixed PCG solve
    std::vector<std::pair<double, double>> race, race_tmp;
    bool have_rate = false;
    // host pinned mirrors
    CgScalars* h_cg = nullptr;
    BicgScalars* h_bicg = nullptr;
    StepScalars* h_ss = nullptr;

    // Facts about the state that let phases skip work exactly:
    // acid_zero: the acid field is 0 everywhere (set by the initial state,
    //   kept by transport, whose acid system then has b = 0 and x = 0, and by
    //   leaching, which only moves acid where it is non-zero; an upload sets it
    //   from the data).  The acid source of the reference is commented out
    //   (src/Simulation.cpp:997-1128), so this holds unless acid is injected.
    // frac_clean: no voxel holds 0 < frac <= 1e-4 (the initial state has frac
    //   0 or 1; any upload of frac or a general leaching sweep clears it).
    // With both, the leaching sweep cannot change anything.
    bool acid_zero = false, frac_clean = false;
    bool b_acid_zero = false, b_frac_clean = false;   // at the last snapshot
    // The step's snapshot copies neither p nor (while acid is known zero and
    // frac clean, so that leaching cannot change it) frac: the backup is
    // stale and the field itself still holds the snapshot's values.  The
    // pressure phase then writes the new p into the backup's buffer and swaps
    // the two; the leaching sweep copies frac before it writes; any other
    // write of the state first copies the field (materialize_backups).
    bool bp_stale = false, bfrac_stale = false;
    double p_top_prev = 0.0;
    bool have_prev = false;
    bool sat_valid = true;     // physics check of the last implicit saturation solve
    // warm-start history: the last n_hist pressure solutions (physical p,
    // FP64), newest at hist[(hist_head + hist_ring - 1) % hist_ring]
    double* hist[LS_WARM_MAX] = {};
    // block partials of a residual the projected start computed (0: none)
    int resid_ready = 0;
    double hist_ptop[LS_WARM_MAX] = {};   // effective top pressure of each
    int n_hist = 0;
    int hist_head = 0;
    int hist_ring = 1;   // ring length: the configured warm_history (1..LS_WARM_MAX)
    double* wsum = nullptr;        // projection dot products (device)
    double* h_wsum = nullptr;      // pinned mirror
    bool press_assembled = false;

    hipEvent_t ev[18];   // step phase marks (ls_step uses 0-10), sampled stencil and copy
                         // timing (12-15), the transport's sampled red pass (16-17)
    // sampled stencil timing (profiling on): events 12/13 bracket one launch
    int profiling = 0;
    // prof_pending: 64 / 32 (FP64 / FP32 stencil) or 16 (Schur apply) when an
    // event pair awaits reading; slots 0, 1, 2 accumulate them
    int prof_pending = 0;
    double prof_ms[4] = {0.0, 0.0, 0.0, 0.0};   // slot 3: the transport's red Schur pass
    int prof_n[4] = {0, 0, 0, 0};
    double prof_tred_bytes = 0.0;  // algorithmic bytes of the sampled red pass
    double prof_schur_bytes = 0.0; // algorithmic bytes of the sampled Schur kernel
    std::vector<void*> allocations;
    // device bytes held: arrays that scale with the grid, and the rest
    size_t grid_bytes = 0, other_bytes = 0;
};

// ---- internal entry points (implemented in the kernel files) ----
namespace ls {
// ls_api.hip: (re)cut the phase-scoped pool for the current parameters
int pool_setup(ls_ctx* c);
// k_transport.hip: the transport form the parameters and grid select, and
// the doubles its stored rows take
void transport_form(const ls_ctx* c, bool* stored, bool* schur, bool* mixed = nullptr);
size_t transport_rows_doubles(const ls_ctx* c);
int copy_d2d(ls_ctx* c, void* dst, const void* src, size_t bytes);
int materialize_backups(ls_ctx* c);
size_t transport_rows32_floats(const ls_ctx* c);
// k_physics.hip
int init_state(ls_ctx* c, long long* nsul);
int phase_properties(ls_ctx* c);
int assemble_pressure(ls_ctx* c, double t);
int finalize_pressure(ls_ctx* c, double ptop);
int phase_flux(ls_ctx* c);
int phase_saturation(ls_ctx* c, double dt, ls_step_report* rep);
// k_transport.hip
int phase_saturation_implicit(ls_ctx* c, double dt, ls_step_report* rep);
int phase_surface(ls_ctx* c);
int phase_precipitation(ls_ctx* c);
int summarize(ls_ctx* c, ls_step_report* rep);
int derived_field(ls_ctx* c, int field, double* out_dev);
int zero_flux_rock_precip(ls_ctx* c);
// k_pressure.hip
int solve_pressure(ls_ctx* c, double t, ls_step_report* rep);
int apply_pressure_unscaled(ls_ctx* c, const double* x_dev, double* y_dev);
int time_stencil(ls_ctx* c, int precision, int reps, double* ms, double* bytes);
int debug_schur_apply(ls_ctx* c, int form, const float* p, float* q, double* dot);
int learned_setup(ls_ctx* c);
void learned_free(ls_ctx* c);
int learned_pack_debug(ls_ctx* c, const float* r_dev, float* out_dev);
int learned_galerkin_debug(ls_ctx* c, float* ac_dev);
int learned_sample_debug(ls_ctx* c, const double* pstar_dev, float* x_dev, float* rc_dev,
                         double* eae);
// k_transport.hip
int phase_transport(ls_ctx* c, double dt, ls_step_report* rep);
int time_transport_pass(ls_ctx* c, int which, int reps, double* ms, double* bytes);
// k_leach.hip
int phase_leaching(ls_ctx* c, double dt, ls_step_report* rep);
}  // namespace ls

#define LS_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return -10;                                                          \
        }                                                                        \
    } while (0)

#define LS_LAUNCH_CHECK()                                                        \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess) {                                                  \
            c->err = std::string("kernel launch: ") + hipGetErrorString(e_);     \
            return -11;                                                          \
        }                                                                        \
    } while (0)
