/*
 * lsgpu.h — C ABI of the MI355X (gfx950) per-timestep hot path of LeachingSim.
 *
 * The reference has no plugin or FFI layer: its hot path is the body of the
 * time loop in src/main.cpp:293-453, made of public methods of
 * `class Simulation` (src/Simulation.h:29-186).  Each entry point below
 * replaces one of those methods (or a group that the loop always calls
 * together) and is cited with the reference lines it stands in for.
 *
 * Conventions
 *   - Plain C types only: pointers, sizes, ints, doubles.  No torch or HIP
 *     types cross this boundary.
 *   - Grid (ni, nj, nk) = the reference's internal (i, j, k) = CLI
 *     (zext, yext, xext) (src/main.cpp:94).  Storage is k-fastest:
 *     idx = k + nk*(j + nj*i) (src/Domain.cpp:106-117), the same byte order
 *     as the .raw input (src/MPIRawLoader.h:70).
 *   - Labels: Air 1, Pore 2, Rock 3, Sulphide 4, Precipitate 5
 *     (src/compiler_opts.h:69-82).  uint8 or uint16 on upload; any other
 *     value is rejected (an unknown label would leave an empty pressure row,
 *     src/Simulation.cpp:683-792).
 *   - Status: 0 = OK; > 0 = a solve did not converge or the state is
 *     unphysical (the reference's `bool false`, which triggers the retry
 *     manager, src/main.cpp:437-452); < 0 = fatal error (the reference's
 *     thrown std::runtime_error / MPI_Abort).  ls_last_error() explains.
 *   - Ownership: the context owns all device memory.  Host buffers are owned
 *     by the caller and copied at the boundary.
 *   - Threading: one HIP stream per context; calls on one context must be
 *     serialised by the caller.
 */
#ifndef LSGPU_H_
#define LSGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSGPU_ABI_VERSION 10

/* Labels (src/compiler_opts.h:69-82). */
enum ls_label {
    LS_AIR = 1,
    LS_PORE = 2,
    LS_ROCK = 3,
    LS_SULPHIDE = 4,
    LS_PRECIPITATE = 5
};

/* Pressure solver choice.  The reference always uses PETSc GMRES + Jacobi
 * (src/Simulation.cpp:131-159); the solvers here stop on the same criterion,
 * ||b - A x||_2 / ||b||_2 <= press_rtol on the reduced SPD pressure system. */
enum ls_solver {
    LS_SOLVER_PLAIN64 = 0,  /* FP64 Jacobi-preconditioned CG                     */
    LS_SOLVER_MIXED = 1,    /* FP32 CG (red-black Schur complement, mixed_schur)
                               inside FP64 reliable updates                     */
    LS_SOLVER_LEARNED = 2   /* flexible CG, U-Net preconditioner + warm start,
                               FP32 inner / FP64 outer, plain PCG fallback       */
};

/* Pressure boundary mode. */
enum ls_bc_mode {
    LS_BC_REFERENCE = 0,    /* Dirichlet data from Air voxels only
                               (src/Simulation.cpp:683-721)                      */
    LS_BC_ZFACES = 1        /* additionally: virtual Air planes at i = -1
                               (p = ramped top pressure) and i = ni (p = 0),
                               linked to Pore rows by the reference's Pore-Air
                               conductance (K_air = 1e-9)                        */
};

/* Warm start of the pressure solve. */
enum ls_warm_start {
    LS_WARM_PREVIOUS = 0,   /* previous pressure (src/Simulation.cpp:157,808)    */
    LS_WARM_PROJECTED = 1,  /* Galerkin projection onto the last warm_history
                               solutions: the A-norm-optimal combination under
                               the new operator (never worse than the previous
                               pressure; exact for a fixed operator under the
                               linear top-pressure ramp)                        */
    LS_WARM_LEARNED = 2     /* the projection basis plus the U-Net correction of
                               the projected residual                           */
};
#define LS_WARM_MAX 8       /* largest projection basis                          */

/* All physics parameters of src/main.cpp:39,69-91 plus solver settings.
 * ls_default_params() fills the reference's defaults. */
typedef struct ls_params {
    double D;                /* --D            (1.0)                 */
    double dx;               /* --voxel_size   (1.0)                 */
    double Dpore_fac;        /* --Dpore_fac    (1.0)                 */
    double kext;             /* --kext         (1.0, inert)          */
    double kreac;            /* --kreac        (1.0)                 */
    double theta;            /* --theta        (1.0)                 */
    double c_sat;            /* --csat         (0.95)                */
    double evap_flux;        /* --evap_flux    (0.0)                 */
    double top_pressure;     /* --top_pressure (1e4)                 */
    double max_cap_grad;     /* --max_cap_grad (5e4)                 */
    double vg_alpha;         /* 1.0  (src/main.cpp:90)               */
    double vg_n;             /* 2.0  (src/main.cpp:89)               */
    double s_res;            /* 0.05 (src/main.cpp:91)               */
    double porosity;         /* --porosity     (0.1)                 */
    double solid_density;    /* never assigned in the reference (src/Simulation.h:135); default 1.0 */
    double time_scalar;      /* --time_scalar  (1.0)                 */
    double alpha_transfer;   /* --alpha_transfer (1e-5)              */
    int instant_precip;      /* --instant_precip (0)                 */
    int bc_mode;             /* enum ls_bc_mode                      */
    int solver;              /* enum ls_solver                       */
    int warm_start;          /* enum ls_warm_start                   */
    double press_rtol;       /* 1e-8                                 */
    int press_maxit;         /* 10000 (src/Simulation.cpp:156)       */
    double adr_rtol;         /* 1e-8                                 */
    int adr_maxit;           /* 10000                                */
    int warm_history;        /* solutions kept for LS_WARM_PROJECTED / LEARNED (1:
                                the newest solution, optimally scaled), 1 ..
                                LS_WARM_MAX (LEARNED uses one slot for the
                                network correction).  More vectors make the
                                start exact when the right-hand side moves in
                                their span under a fixed operator (a ramp plus a
                                fixed evaporation source); on the 512^3
                                fractured workload, whose operator drifts every
                                step, 2 and 4 vectors left the start's residual
                                no lower and made a step 2-13 ms slower         */
    double mixed_inner_rtol; /* mixed solver: the FP64 residual replaces the FP32 one
                                whenever the latter has fallen by this factor (0.1) */
    double learned_race;    /* learned solver: a race against the residual curve
                                of the last mixed PCG solve; after the same wall
                                time flexible CG must have reduced the residual by
                                at least this multiple of the decades mixed PCG
                                had, else mixed PCG takes over from the saved
                                start (1.0; 0 = no race)                         */
    int implicit_sat;        /* --implicit_sat (0): implicit upwind saturation solve
                                inside 3 Picard passes (src/main.cpp:312-370,
                                src/Simulation.cpp:1858-2053), solved to adr_rtol */
    int mixed_schur;         /* mixed solver (1): FP32 inner CG on the red-black
                                Schur complement of the scaled operator (about
                                half the iterations of Jacobi-PCG), on vectors
                                compacted to the black voxels with one fused
                                pass per operator apply when nk % 8 == 0,
                                nk <= 2560 and a row of nk / 2 values fills at
                                least 3/8 of the lanes of its waves (256 values
                                each; so 192 <= nk <= 2560), else as 2; 3: the
                                fused form wherever nk allows it; 2: the same
                                on full-length vectors in two passes (any nk,
                                vectorised when nk % 4 == 0); 0: inner CG on
                                the operator                                    */
    int adr_schur;           /* transport and implicit saturation (1): BiCGStab on
                                the red-black Schur complement of the stored rows
                                (black voxels, about half the iterations), the
                                red values recovered after convergence; needs
                                nk even; 0: BiCGStab on the full system         */
    int adr_extrapolate;     /* transport on the Schur complement (1): conc starts
                                from its extrapolation in time through the fields
                                of the last solves (cubic through the last four,
                                lower orders after fewer solves) instead of
                                the old field (the reference's start,
                                src/Simulation.cpp:1247); 0: the old field       */
    int adr_mixed;           /* transport and implicit saturation on the Schur
                                complement with nk % 8 == 0 (1): BiCGStab in FP32
                                on FP32 copies of the rows, for a correction of
                                the FP64 iterate, inside FP64 refinement cycles
                                whose FP64 residual is the stop test; 0: BiCGStab
                                in FP64                                          */
    double adr_mixed_inner_rtol; /* adr_mixed: a cycle's FP32 solve stops at this
                                reduction of its own residual, or earlier at half
                                the FP64 tolerance (1e-6)                         */
} ls_params;

/* Fields that can be uploaded / downloaded.  Element types are fixed:
 * IMG is uint8 or uint16 (label_bytes of the context), FRAC is float,
 * everything else double.  KREL, PC, PERM, MU are derived fields, computed
 * from the state on download (src/Simulation.cpp:187-248,1596-1664). */
enum ls_field {
    LS_F_IMG = 0,
    LS_F_FRAC = 1,
    LS_F_INVENTORY = 2,   /* precipitate_inventory */
    LS_F_SATURATION = 3,
    LS_F_PRESSURE = 4,
    LS_F_CONC = 5,
    LS_F_ACID = 6,        /* conc_acid */
    LS_F_QX = 7,          /* flux_x (i direction) */
    LS_F_QY = 8,
    LS_F_QZ = 9,
    LS_F_KREL = 10,       /* rel_permeability (derived)  */
    LS_F_PC = 11,         /* cap_pressure (derived)      */
    LS_F_PERM = 12,       /* permeability (derived)      */
    LS_F_MU = 13,         /* viscosity (derived)         */
    LS_F_COUNT = 14
};

/* Phases of one coupled step, in the order of src/main.cpp:312-432. */
enum ls_phase_id {
    LS_PH_PROPERTIES = 0,   /* updateRelativePermeability + updateCapillaryPressure
                               + initProperties  (src/Simulation.cpp:1634,1596,187) */
    LS_PH_PRESSURE = 1,     /* setupPressureEqns(t) + solvePressure   (:658, :804)   */
    LS_PH_FLUX = 2,         /* calculateFlux                           (:866)         */
    LS_PH_SATURATION = 3,   /* updateSaturation(dt)                    (:1672)        */
    LS_PH_SURFACE = 4,      /* handleSurfaceEffects                    (:2063)        */
    LS_PH_TRANSPORT = 5,    /* setupConcentrationEqns + solveConc + solveAcid
                                                                       (:966,:1243,:2167) */
    LS_PH_PRECIPITATION = 6,/* handlePrecipitation                     (:2114)        */
    LS_PH_LEACHING = 7,     /* updateFrac(dt)                          (:315)         */
    LS_PH_COUNT = 8
};

/* What one step (or phase) reports back: enough to print the reference's
 * log lines (src/Simulation.cpp:457,838,1772; src/main.cpp:426-430). */
typedef struct ls_step_report {
    int status;
    int press_iters;          /* total CG iterations (inner iterations for mixed) */
    int press_outer;          /* refinement / outer iterations                    */
    int press_fallback;       /* learned path fell back to plain PCG              */
    int press_unet_applies;   /* U-Net preconditioner applications               */
    double press_relres;      /* final ||b-Ax||/||b|| (FP64, reduced system)      */
    double press_bnorm;
    int adr_iters;
    double adr_relres;        /* max over the two right-hand sides                */
    int sat_substeps;
    double sat_max_divq;
    int leach_substeps;
    double leach_max_rate;
    int leached;              /* voxels fully leached this step                   */
    int leach_general_path;   /* 1 if the order-exact wavefront path ran          */
    long long n_floored;      /* pore rows with |diagonal| < 1e-16               */
    double p_min, p_max, conc_min, conc_max, acid_min, acid_max;
    double mass_conc, mass_acid, mass_water;   /* FP64 sums over the grid        */
    double top_pressure_eff;
    double t_props_ms, t_pressure_ms, t_flux_ms, t_sat_ms, t_surface_ms,
           t_transport_ms, t_precip_ms, t_leach_ms, t_snapshot_ms, t_total_ms;
    /* With profiling on (ls_set_profiling): HIP-event durations of sampled
     * pressure-stencil launches (the first launch of every CG batch) during
     * this step, summed, and their counts, for FP64 and FP32 stencils. */
    double prof_stencil64_ms, prof_stencil32_ms;
    int prof_stencil64_n, prof_stencil32_n;
    double press_relres0;     /* ||b-Ax0||/||b|| of the warm start               */
    int warm_basis;           /* projection vectors kept (0: previous pressure)  */
    int warm_net;             /* 1 if the network correction was in the basis   */
    int sat_iters;            /* implicit saturation: solver iterations (last pass) */
    int sat_valid;            /* implicit saturation: raw solution within
                                 [-0.05, 1.05] (the reference's physics_valid)  */
    double sat_raw_min, sat_raw_max;   /* implicit saturation: before the clamp */
    /* With profiling on: the same for the FP32 Schur complement apply of
     * the mixed solver (the fused kernel, or the gather pass of the two-pass
     * form) and its algorithmic HBM bytes per launch. */
    double prof_schur_ms;
    int prof_schur_n;
    double prof_schur_bytes;
    /* Iterations of each transport system (conc, acid), for the reference's
     * per-solve log lines (src/Simulation.cpp:1278, 2208). */
    int adr_iters_conc, adr_iters_acid;
    /* With profiling on: HIP-event durations of sampled launches of the
     * transport's red Schur pass (w = D_r^-1 A_rb p, the first launch of every
     * BiCGStab batch), their count and its algorithmic HBM bytes per launch. */
    double prof_tred_ms;
    int prof_tred_n;
    double prof_tred_bytes;
    /* Mixed-precision transport (adr_mixed): FP64 refinement cycles of the
     * last transport solve, and 1 if it handed over to FP64 BiCGStab (a
     * cycle that did not cut the FP64 residual by 16x). */
    int adr_cycles;
    int adr_fallback;
} ls_step_report;

typedef struct ls_ctx ls_ctx;

/* Fill the reference's defaults (src/main.cpp:39,89-91). */
void ls_default_params(ls_params* p);
/* ABI check for foreign bindings: returns LSGPU_ABI_VERSION and the sizes of
 * the structs above as this library was compiled. */
int ls_abi_info(size_t* sizeof_params, size_t* sizeof_report);

/* Replaces Simulation::setupDomain + setupPETSc (src/Simulation.cpp:81-159):
 * allocates all device state for an ni x nj x nk grid on the current HIP
 * device (each extent >= 2, ni nj nk < 2^32; else -2).  label_bytes is 1
 * (uint8 build) or 2 (uint16 build). */
int ls_create(ls_ctx** out, int ni, int nj, int nk, int label_bytes,
              const ls_params* params);
void ls_destroy(ls_ctx* ctx);
const char* ls_last_error(const ls_ctx* ctx);
int ls_get_params(const ls_ctx* ctx, ls_params* out);
int ls_set_params(ls_ctx* ctx, const ls_params* params);

/* Replaces readRAW + the init* functions' writes (src/Simulation.cpp:68,
 * 165-310): copy one unpadded host field (ni*nj*nk elements) to the device. */
int ls_upload(ls_ctx* ctx, int field, const void* host, size_t elem_bytes);
/* Copy one field back (for VTK output, checkpoints and tests). */
int ls_download(ls_ctx* ctx, int field, void* host, size_t elem_bytes);

/* Reference initial state (initFields/initFrac/initSaturation/initAcid,
 * src/Simulation.cpp:165-180,254-310,1573-1590) on the device, from the
 * uploaded labels: p = conc = acid = q = inventory = 0, frac = 1 on Sulphide,
 * s = 0.1 on Pore and Sulphide.  Returns the initial sulphide count. */
int ls_init_state(ls_ctx* ctx, long long* initial_sulphide_count);

/* One coupled step: the body of the retry loop, src/main.cpp:303-432
 * (saveState, then properties, pressure, flux, saturation, surface effects,
 * conc + acid transport, precipitation, leaching).  On status > 0 the state
 * is left as the failed attempt made it; call ls_rollback() and retry with
 * a smaller dt, as src/main.cpp:437-452 does. */
int ls_step(ls_ctx* ctx, double t, double dt, ls_step_report* report);

/* One phase of the step on the current state (for tests and the per-method
 * host mirror).  Fields the report does not cover are left zero. */
int ls_phase(ls_ctx* ctx, int phase, double t, double dt, ls_step_report* report);

/* Replaces saveState / restoreState (src/Simulation.cpp:2312-2353). */
int ls_snapshot(ls_ctx* ctx);
int ls_rollback(ls_ctx* ctx);

/* Test / measurement hooks (no reference counterpart). */
/* Turn sampled kernel timing inside ls_step/ls_phase on (1) or off (0). */
int ls_set_profiling(ls_ctx* ctx, int on);
/* y = A x for the reduced pressure operator assembled by the last pressure
 * phase (rows outside the solved set return 0).  Host buffers, N doubles. */
int ls_debug_apply_pressure(ls_ctx* ctx, const double* x, double* y);
/* The scaling of the last assembled pressure system: sqrt(d) (0 off the
 * solved set), the scaled right-hand side D^-1/2 b and the row classes
 * (0 Dirichlet, 1 solved, 2 floored).  N elements each. */
int ls_debug_pressure_system(ls_ctx* ctx, double* sqrt_d, double* bhat, int8_t* cls);
/* The scaled FP64 +face conductances of the last assembled system (the
 * operator A^ x = x - sum c^ x_n), 3 x N doubles (i, j, k faces). */
int ls_debug_pressure_faces(ls_ctx* ctx, double* faces);
/* q = S p with S = I - N_br N_rb, the red-black Schur complement of the
 * last assembled FP32 scaled operator, on host vectors compacted to the black
 * voxels (N/2 floats: black voxel v at index v >> 1), with *dot = p.q.
 * form 1: the fused kernel (nk % 8 == 0, nk <= 1024); form 2: the gather and apply passes
 * on full-length vectors (nk % 4 == 0); form 4: the same passes on full-length host
 * vectors (N floats; the red entries of p are ignored and come back zero in q), for any
 * nk (one voxel per thread when nk % 4 != 0, as the solver runs them). */
int ls_debug_schur_apply(ls_ctx* ctx, int form, const float* p, float* q, double* dot);
/* Time `reps` launches of a pressure kernel on the context's stream with HIP
 * events: the stencil y = A x (precision 64 or 32), or one of the FP32
 * red-black Schur passes (LS_TIME_RB_GATHER: w = N_rb p; LS_TIME_RB_APPLY:
 * q = p - N_br w with its dot, vectorised when nk % 4 == 0, else one voxel per
 * thread; LS_TIME_SCHUR: both fused); returns ms per launch and the kernel's
 * algorithmic bytes per launch. */
#define LS_TIME_RB_GATHER 1
#define LS_TIME_RB_APPLY 2
#define LS_TIME_SCHUR 3   /* fused q = S p on compacted vectors (nk % 8 == 0) */
/* The transport's Schur passes on the rows the last transport solve left
 * (run ls_step or the transport phase first; the pressure phase reuses those
 * bytes): LS_TIME_TRED the red pass w = R' p, R' = D_r^-1 A_rb D_b^-1 (32 B
 * per voxel of rows, p and w), LS_TIME_TBLACK the black pass z = p - A_br w
 * with its dot (40 B per voxel), both for the conc system alone; the same
 * two passes of the mixed-precision form on its FP32 rows and vectors
 * (LS_TIME_TRED32 16 B, LS_TIME_TBLACK32 20 B per voxel; adr_mixed with
 * nk % 8 == 0). */
#define LS_TIME_TRED 4
#define LS_TIME_TBLACK 5
#define LS_TIME_TRED32 6
#define LS_TIME_TBLACK32 7
int ls_time_stencil(ls_ctx* ctx, int precision, int reps, double* ms_per_launch,
                    double* bytes_per_launch);
/* Copy bytes within device memory on the context's stream, timed with HIP
 * events (a copy kernel of known size, for bandwidth calibration). */
int ls_time_copy(ls_ctx* ctx, size_t bytes, int reps, double* ms_per_launch);
/* Device memory the context holds now: grid_bytes for arrays that scale with
 * the grid (state, snapshot, properties, the phase pool, the warm-start
 * history, the transport's previous solution, the learned solver's grids),
 * other_bytes for the rest (reduction partials, scalars, network weights).
 * No counterpart in the reference (its PETSc objects live in host memory). */
int ls_device_bytes(const ls_ctx* ctx, unsigned long long* grid_bytes,
                    unsigned long long* other_bytes);

/* ---- Allocation-free entry point.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is the
 * saturation phase of ls_phase (LS_PH_SATURATION, updateSaturation,
 * src/Simulation.cpp:1672-1856) with the same kernels and the same bits, on
 * caller-owned DEVICE fields of an ni x nj x nk grid (k-fastest, no
 * context): labels as uint8 (1..5), the three cell fluxes, and the
 * saturation, updated in place.  It calls no hipMalloc/hipFree: the scratch
 * of its global MAX (block partials) and sub-step plan is a static device
 * array of LS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
 * Limits: 1 <= ni <= 65535, nj, nk >= 1, ni*nj*nk < 2^32 (else -2); the
 * partials never exceed the scratch at any such size.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU race (serialise
 * them, e.g. on one stream).  `stream` is a hipStream_t (NULL: the null
 * stream); the kernels are enqueued on it.  nsub / max_rate, if not NULL,
 * receive the sub-step count and max |div q| + evap; reading them
 * synchronises the stream, so pass NULL for both under graph capture.
 * Errors: ls_noalloc_last_error() (per thread). */
#define LS_NOALLOC_SCRATCH_BYTES 65792
int ls_update_saturation_noalloc(int ni, int nj, int nk, const uint8_t* img, const double* qx,
                                 const double* qy, const double* qz, double* s, double dt,
                                 double dx, double porosity, double evap_flux, void* stream,
                                 int* nsub, double* max_rate);
const char* ls_noalloc_last_error(void);

/* ---- The learned part: a 3-level 3D U-Net (16/32/64 channels, 3x3x3
 * convolutions, BF16 on MFMA with FP32 accumulation), no reference
 * counterpart.  Layers and the flat FP32 weight layout are those of
 * leachingsim_amd/unet_spec.py.  These entry points train it and run it on
 * host buffers; the pressure solver uses it through ls_set_network(). ---- */
typedef struct ls_net ls_net;

size_t ls_net_num_weights(void);
/* A network on an ni x nj x nk grid (multiples of 16) for `batch` samples;
 * trainable != 0 also allocates gradient and optimiser buffers. */
int ls_net_create(ls_net** out, int ni, int nj, int nk, int batch, int trainable);
void ls_net_destroy(ls_net* net);
const char* ls_net_last_error(const ls_net* net);
int ls_net_set_weights(ls_net* net, const float* w, size_t count);
int ls_net_get_weights(ls_net* net, float* w, size_t count);
/* in: [batch][ni][nj][nk][16] FP32; out: [batch][ni][nj][nk] (channel 0). */
int ls_net_forward(ls_net* net, const float* in, float* out);
/* Weight gradient of sum(gout * out) for the input of the last forward. */
int ls_net_backward(ls_net* net, const float* gout, float* grad_w);
/* One Adam step on the masked mean squared error of output channel 0. */
int ls_net_train_step(ls_net* net, const float* in, const float* target, const float* mask,
                      float lr, double* loss);
/* One Adam step for an arbitrary loss: backward of the last forward with
 * gout = d(loss)/d(output channel 0) ([batch][N]). */
int ls_net_step(ls_net* net, const float* gout, float lr);
/* Average time of `reps` forward passes on the device-resident input. */
int ls_net_time_forward(ls_net* net, int reps, double* ms_per_forward);
/* Convolution kernels (forward and backward-data): 1 (default) the
 * LDS-tiled kernels for the layers that have one (at most 48 input
 * channels), 0 the general implicit-GEMM kernel everywhere.  Both give the
 * same bits; the switch is for tests and timing. */
int ls_net_set_conv(ls_net* net, int tiled);
/* Training on the Galerkin energy loss with the samples held on the device
 * (no reference counterpart: the training of the learned preconditioner).
 * With u = P out (the coarse output copied to the solved fine children),
 * loss = 1 - (u.r)^2 / (u.A u  e.A e) for r = A e, evaluated exactly on the
 * network's grid from rc = P^T r and the Galerkin operator A_c = P^T A P.
 * ls_net_store_alloc: `samples` sample slots and `operators` operator slots
 * (trainable network, batch <= 8).  ls_net_store_operator: A_c as [4][n]
 * FP32 (diagonal; couplings of voxel q to q + i, q + j, q + k).
 * ls_net_store_sample: input [n][16] FP32 (kept as BF16), rc [n] FP32 and
 * e.A e, with the slot of its operator.  ls_net_energy_step: one batch of
 * stored samples (net batch size) through the forward pass, the loss and
 * its output gradient on the device; for lr > 0 the backward pass and an
 * Adam step, for lr <= 0 with grad_w given the backward pass alone, the
 * weight gradient of the mean loss copied to grad_w (nw floats); losses[b]
 * is each sample's loss. */
int ls_net_store_alloc(ls_net* net, int samples, int operators);
int ls_net_store_operator(ls_net* net, int slot, const float* ac);
int ls_net_store_sample(ls_net* net, int slot, int op_slot, const float* x, const float* rc,
                        double eae);
int ls_net_energy_step(ls_net* net, const int* slots, float lr, double* losses,
                       float* grad_w);

/* Weights of the learned coarse-level preconditioner used by
 * LS_SOLVER_LEARNED (flat FP32 layout above).  The network runs on the grid
 * given by ls_network_grid(): the pressure grid halved (rounded up) and
 * padded to multiples of 16. */
int ls_set_network(ls_ctx* ctx, const float* weights, size_t count);
int ls_network_grid(ls_ctx* ctx, int* dims3);
/* The network input built from a scaled residual r (N floats, host) by the
 * learned preconditioner's restriction (for tests and training data):
 * out = [coarse grid][16] FP32 (values rounded to BF16). */
int ls_debug_learned_pack(ls_ctx* ctx, const float* r, float* out, size_t out_count);
/* Training data of the learned coarse correction (no reference counterpart;
 * used by leachingsim_amd/learn.py).  ls_debug_learned_galerkin: the
 * Galerkin operator P^T A^ P of the assembled scaled pressure operator on the
 * network grid, [4][cn] (diagonal; couplings to the +i, +j, +k neighbours),
 * count = 4 cn.  ls_debug_learned_sample: with the current pressure field p
 * an iterate and p_star the reference solution (N doubles), the network
 * input of r = A^ e, e = sqrt(d) (p_star - p) on the solved rows
 * ([cn][16], x_count = 16 cn), rc = P^T r ([cn]) and e.A^ e. */
int ls_debug_learned_galerkin(ls_ctx* ctx, float* ac, size_t count);
int ls_debug_learned_sample(ls_ctx* ctx, const double* p_star, float* x, size_t x_count,
                            float* rc, double* eae);

#ifdef __cplusplus
}
#endif

#endif /* LSGPU_H_ */
