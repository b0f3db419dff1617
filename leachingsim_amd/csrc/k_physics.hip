// k_physics.hip — pointwise and stencil kernels of the coupled step other than
// the linear solvers: initial state, properties, pressure operator assembly
// and post-solve, Darcy flux, explicit saturation, surface effects,
// precipitation, derived-field output and step summaries.
//
// Each kernel restates one loop of src/Simulation.cpp (cited per kernel) with
// the same floating-point expressions in the same order.  The library is
// compiled with -ffp-contract=off so no multiply-add is fused: results round
// exactly as the reference's scalar C++ does.
#include <cmath>

#include "ls_context.h"

namespace ls {

// std::max / std::min semantics (NaN handling included): max(a,b) = a<b ? b : a.
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

struct Geo {
    int ni, nj, nk;
    uint32_t si, sj;  // strides of i and j (k is contiguous)
};

__device__ __forceinline__ void decode(uint32_t v, const Geo& g, int& i, int& j, int& k) {
    uint32_t ij = v / (uint32_t)g.nk;
    k = (int)(v - ij * (uint32_t)g.nk);
    i = (int)(ij / (uint32_t)g.nj);
    j = (int)(ij - (uint32_t)i * (uint32_t)g.nj);
}

// Neighbour f in the reference order +i, -i, +j, -j, +k, -k
// (src/Simulation.cpp:771-776); false if outside the grid.
__device__ __forceinline__ bool nbr(const Geo& g, int i, int j, int k, uint32_t v, int f,
                                    uint32_t& n) {
    switch (f) {
        case 0: if (i + 1 >= g.ni) return false; n = v + g.si; return true;
        case 1: if (i == 0) return false; n = v - g.si; return true;
        case 2: if (j + 1 >= g.nj) return false; n = v + g.sj; return true;
        case 3: if (j == 0) return false; n = v - g.sj; return true;
        case 4: if (k + 1 >= g.nk) return false; n = v + 1; return true;
        default: if (k == 0) return false; n = v - 1; return true;
    }
}

// All six neighbour indices of v at once, in the reference order; an index
// outside the grid is replaced by v itself (in bounds, never used: ok[f]
// is false), so that a kernel can issue every neighbour load before it
// looks at any of them.
__device__ __forceinline__ void nbr6(const Geo& g, int i, int j, int k, uint32_t v,
                                     uint32_t n[6], bool ok[6]) {
    ok[0] = i + 1 < g.ni; n[0] = ok[0] ? v + g.si : v;
    ok[1] = i > 0;        n[1] = ok[1] ? v - g.si : v;
    ok[2] = j + 1 < g.nj; n[2] = ok[2] ? v + g.sj : v;
    ok[3] = j > 0;        n[3] = ok[3] ? v - g.sj : v;
    ok[4] = k + 1 < g.nk; n[4] = ok[4] ? v + 1 : v;
    ok[5] = k > 0;        n[5] = ok[5] ? v - 1 : v;
}

static Geo geo_of(const ls_ctx* c) {
    Geo g;
    g.ni = c->ni;
    g.nj = c->nj;
    g.nk = c->nk;
    g.sj = (uint32_t)c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    return g;
}

// ---------------------------------------------------------------------------
// Properties: krel (src/Simulation.cpp:1634-1664), Pc (:1596-1628),
// K and mu (:187-248).
// ---------------------------------------------------------------------------
struct PropConst {
    double m, s_res, vg_alpha, vg_n, log_k_rock, log_k_pore;
    // 10^log_k_pore and 10^log_k_rock, evaluated once on the host (the
    // intrinsic permeability of a pore with no inventory, of a sulphide with
    // fraction 0 or 1, and of a fully clogged pore)
    double k_pore, k_rock;
    // vg_n == 2 (the reference's value, src/main.cpp:89), so m = 1/2: the
    // van Genuchten powers are x^2 = x * x, x^(1/2) = sqrt(x) and x^-2 =
    // 1 / (x * x), each one rounding (correctly rounded, as pow is to within
    // an ulp) instead of a general pow
    int vg2;
};

__device__ __forceinline__ double rel_perm(uint8_t t, double s, const PropConst& pc) {
    if (t == PORE || t == SUL) {
        double Se = (s - pc.s_res) / (1.0 - pc.s_res);
        Se = dmax(0.0, dmin(1.0, Se));
        if (Se <= 0.0) return 0.0;
        double term1 = pc.vg2 ? 1.0 - sqrt(1.0 - Se * Se)
                              : 1.0 - pow(1.0 - pow(Se, 1.0 / pc.m), pc.m);
        return sqrt(Se) * term1 * term1;
    }
    return 0.0;
}

__device__ __forceinline__ double cap_pressure(uint8_t t, double s, const PropConst& pc) {
    if (t == PORE || t == SUL) {
        double Se = (s - pc.s_res) / (1.0 - pc.s_res);
        Se = dmax(1e-6, dmin(1.0, Se));
        if (Se >= 1.0) return 0.0;
        if (pc.vg2) return (1.0 / pc.vg_alpha) * sqrt(1.0 / (Se * Se) - 1.0);
        double Se_inv_m = pow(Se, -1.0 / pc.m);
        return (1.0 / pc.vg_alpha) * pow(Se_inv_m - 1.0, 1.0 / pc.vg_n);
    }
    return 0.0;
}

__device__ __forceinline__ double viscosity(uint8_t t) { return (t == AIR) ? 1.8e-5 : 1.0e-3; }

__device__ __forceinline__ double permeability(uint8_t type, double krel, double inv, float fr,
                                               const PropConst& pc) {
    const double pore_permeability = 1.0e-12;
    const double rock_permeability = 1.0e-18;
    double K_intrinsic = rock_permeability;
    if (type == AIR) {
        K_intrinsic = pore_permeability;
    } else if (type == PORE) {
        double p_frac = dmax(0.0, dmin(1.0, inv));
        // p_frac 0 or 1 makes log_k_eff exactly log_k_pore or log_k_rock
        if (p_frac == 0.0) {
            K_intrinsic = pc.k_pore;
        } else if (p_frac == 1.0) {
            K_intrinsic = pc.k_rock;
        } else {
            double log_k_eff = (p_frac * pc.log_k_rock) + ((1.0 - p_frac) * pc.log_k_pore);
            K_intrinsic = exp10(log_k_eff);
        }
    } else if (type == SUL) {
        double f = fr;
        if (f == 1.0) {
            K_intrinsic = pc.k_rock;
        } else if (f == 0.0) {
            K_intrinsic = pc.k_pore;
        } else {
            double log_k_eff = (f * pc.log_k_rock) + ((1.0 - f) * pc.log_k_pore);
            K_intrinsic = exp10(log_k_eff);
        }
    }
    double K_rel = dmax(krel, 1e-15);
    return K_intrinsic * K_rel;
}

static PropConst prop_const(const ls_ctx* c) {
    PropConst pc;
    pc.m = 1.0 - (1.0 / c->P.vg_n);
    pc.s_res = c->P.s_res;
    pc.vg_alpha = c->P.vg_alpha;
    pc.vg_n = c->P.vg_n;
    pc.log_k_rock = c->log_k_rock;
    pc.log_k_pore = c->log_k_pore;
    pc.k_pore = std::pow(10.0, c->log_k_pore);
    pc.k_rock = std::pow(10.0, c->log_k_rock);
    pc.vg2 = c->P.vg_n == 2.0 ? 1 : 0;
    return pc;
}

__global__ void k_props(const uint8_t* __restrict__ img, const double* __restrict__ s,
                        const double* __restrict__ inv, const float* __restrict__ frac,
                        double* __restrict__ perm, double* __restrict__ pcap, uint32_t N,
                        PropConst pc) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        uint8_t t = img[v];
        double sv = s[v];
        double kr = rel_perm(t, sv, pc);
        __builtin_nontemporal_store(cap_pressure(t, sv, pc), pcap + v);
        __builtin_nontemporal_store(permeability(t, kr, inv[v], frac[v], pc), perm + v);
    }
}

__global__ void k_derived(const uint8_t* __restrict__ img, const double* __restrict__ s,
                          const double* __restrict__ inv, const float* __restrict__ frac,
                          double* __restrict__ out, uint32_t N, int field, PropConst pc) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        uint8_t t = img[v];
        double kr = rel_perm(t, s[v], pc);
        double r;
        if (field == LS_F_KREL) r = kr;
        else if (field == LS_F_PC) r = cap_pressure(t, s[v], pc);
        else if (field == LS_F_PERM) r = permeability(t, kr, inv[v], frac[v], pc);
        else r = viscosity(t);
        out[v] = r;
    }
}

int phase_properties(ls_ctx* c) {
    uint32_t N = (uint32_t)c->N;
    hipLaunchKernelGGL(k_props, dim3(ew_blocks(N)), dim3(EW_BLOCK), 0, c->stream, c->img, c->s,
                       c->inv, c->frac, c->perm, c->pc, N, prop_const(c));
    LS_LAUNCH_CHECK();
    return 0;
}

int derived_field(ls_ctx* c, int field, double* out_dev) {
    uint32_t N = (uint32_t)c->N;
    hipLaunchKernelGGL(k_derived, dim3(ew_blocks(N)), dim3(EW_BLOCK), 0, c->stream, c->img, c->s,
                       c->inv, c->frac, out_dev, N, field, prop_const(c));
    LS_LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------
// Initial state: initFields / initFrac / initSaturation / initAcid
// (src/Simulation.cpp:165-180, 254-310, 1573-1590).
// ---------------------------------------------------------------------------
__global__ void k_init(const uint8_t* __restrict__ img, float* frac, double* inv, double* s,
                       double* p, double* conc, double* acid, double* qx, double* qy, double* qz,
                       uint32_t N, double* partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double cnt = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        uint8_t t = img[v];
        p[v] = 0.0; conc[v] = 0.0; acid[v] = 0.0;
        qx[v] = 0.0; qy[v] = 0.0; qz[v] = 0.0;
        inv[v] = 0.0;
        frac[v] = (t == SUL) ? 1.0f : 0.0f;
        s[v] = (t == PORE || t == SUL) ? 0.1 : 0.0;
        cnt += (t == SUL) ? 1.0 : 0.0;
    }
    double a[1] = {cnt};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

template <int OP>
__global__ void k_finalize_one(const double* __restrict__ partials, int n, double* out) {
    __shared__ double lds[1024 / WAVE];
    double acc = red_identity<OP>();
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc = red_combine<OP>(acc, partials[i]);
    double a[1] = {acc};
    block_reduce<1, OP>(a, lds);
    if (threadIdx.x == 0) *out = a[0];
}

int init_state(ls_ctx* c, long long* nsul) {
    uint32_t N = (uint32_t)c->N;
    int nb = ew_blocks(N);
    hipLaunchKernelGGL(k_init, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->img, c->frac, c->inv,
                       c->s, c->p, c->conc, c->acid, c->qx, c->qy, c->qz, N, c->partials);
    LS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_finalize_one<RED_SUM>, dim3(1), dim3(1024), 0, c->stream, c->partials,
                       nb, &c->ss->nsul);
    LS_LAUNCH_CHECK();
    LS_HIP(hipMemcpyAsync(c->h_ss, c->ss, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    if (nsul) *nsul = (long long)c->h_ss->nsul;
    c->have_prev = false;
    c->n_hist = 0;
    c->press_assembled = false;
    c->acid_zero = true;      // k_init: acid = 0, frac in {0, 1}
    c->tprev_n = 0;
    c->frac_clean = true;
    return phase_properties(c);
}

// ---------------------------------------------------------------------------
// Pressure operator (src/Simulation.cpp:658-798), reduced to the solved Pore
// rows.  Two passes over the grid: k_press_diag stores per voxel the row
// class, the Dirichlet value and sqrt(d) of the coupling sum d on the solved
// rows; k_press_rhs_scale
// then builds, per solved row, the right-hand side and the Jacobi-scaled
// +face couplings (FP64 and FP32).  The face conductances themselves are
// recomputed from the permeability where needed, never stored: a face has
// the same bits from either side (see face_cond), so every pass sees the
// coupling the reference's Pore row computes.
// ---------------------------------------------------------------------------

// Harmonic-mean Pore link (src/Simulation.cpp:737-768).
__device__ __forceinline__ double pore_link(double perm_self, double k_neighbor, double mu_self) {
    double K_self = dmax(perm_self, 1e-25);
    double K_neigh = dmax(k_neighbor, 1e-25);
    double K_avg = (2.0 * K_self * K_neigh) / (K_self + K_neigh);
    return K_avg / mu_self;
}

// Conductance of the face between v and n as seen by the Pore side.  Between
// two Pore voxels it is symmetric bit for bit: 2 K_a is exact, so
// (2 K_a) K_b and (2 K_b) K_a round the same product, and the sum commutes.
__device__ __forceinline__ double face_cond(uint8_t tv, uint8_t tn, double pv, double pn) {
    if (tv == PORE) return pore_link(pv, (tn == AIR) ? 1.0e-9 : pn, viscosity(tv));
    if (tn == PORE) return pore_link(pn, (tv == AIR) ? 1.0e-9 : pv, viscosity(tn));
    return 0.0;
}

// One voxel of k_press_diag: its class, Dirichlet value and sqrt(d) from its
// label, its neighbours' labels and permeabilities (ok[f]: the neighbour
// exists) and its own permeability and pressure.
struct PressDiag {
    int8_t cls;
    double pD, sd;
};
__device__ __forceinline__ PressDiag press_diag_core(uint8_t tv, const uint8_t (&tn)[6],
                                                     const bool (&ok)[6], double pv,
                                                     const double (&kn)[6], double pvv, int i,
                                                     int ni, int zfaces, double ptop) {
    PressDiag o;
    if (tv == PORE) {
        double diagonal_term = 0.0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            if (!ok[f]) continue;
            const double K_neighbor = (tn[f] == AIR) ? 1.0e-9 : kn[f];
            diagonal_term -= pore_link(pv, K_neighbor, viscosity(tv));
        }
        // the virtual inlet / outlet planes of the z-face boundary
        if (zfaces && i == 0) diagonal_term -= pore_link(pv, 1.0e-9, viscosity(tv));
        if (zfaces && i == ni - 1) diagonal_term -= pore_link(pv, 1.0e-9, viscosity(tv));
        const int8_t cl = (fabs(diagonal_term) < 1e-16) ? 2 : 1;
        o.cls = cl;
        o.pD = (cl == 2) ? pvv : 0.0;  // floored rows: lagged value
        o.sd = (cl == 1) ? sqrt(-diagonal_term) : 0.0;
    } else {
        o.cls = 0;
        o.sd = 0.0;
        double val = 0.0;
        if (tv == AIR) {  // surface Air: any in-grid neighbour is not Air (:696-711)
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (ok[f] && tn[f] != AIR) val = ptop;
        }
        o.pD = val;
    }
    return o;
}

__global__ void k_press_diag(const uint8_t* __restrict__ img, const double* __restrict__ perm,
                             const double* __restrict__ p,
                             int8_t* __restrict__ cls, double* __restrict__ pD,
                             double* __restrict__ sd, Geo g, int zfaces, double ptop) {
    const int k = blockIdx.x * TK + threadIdx.x;
    const int j = blockIdx.y * TJ + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nk || j >= g.nj) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    const uint8_t tv = img[v];
    uint32_t n[6];
    bool ok[6];
    nbr6(g, i, j, k, v, n, ok);
    // every neighbour's label (and for a Pore voxel its permeability) is
    // loaded before any is used: one memory round trip per wave
    uint8_t tn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) tn[f] = img[n[f]];
    double kn[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double pv = 0.0;
    if (tv == PORE) {
#pragma unroll
        for (int f = 0; f < 6; ++f) kn[f] = perm[n[f]];
        pv = perm[v];
    }
    const PressDiag o = press_diag_core(tv, tn, ok, pv, kn, tv == PORE ? p[v] : 0.0, i, g.ni,
                                        zfaces, ptop);
    cls[v] = o.cls;
    pD[v] = o.pD;
    sd[v] = o.sd;
}

// One solved voxel of k_press_rhs_scale: the scaled right-hand side and the
// scaled +faces (sc), from the neighbours' labels, classes, permeabilities,
// Dirichlet values and sqrt(d).
__device__ __forceinline__ double press_rhs_core(const uint8_t (&tn)[6], const int8_t (&cn)[6],
                                                 const bool (&ok)[6], double pv, double sv,
                                                 const double (&kn)[6], const double (&dn)[6],
                                                 const double (&sn)[6], int i, int zfaces,
                                                 double ptop, double evap_dx, double (&sc)[3]) {
    double source_term = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f)
        if (ok[f] && tn[f] == AIR) source_term -= evap_dx;
    double rhs = -source_term;
    sc[0] = sc[1] = sc[2] = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (!ok[f]) continue;
        const int8_t cl = cn[f];
        if (cl == 1 && (f & 1)) continue;       // a solved -face: the neighbour's +face
        const double fc = face_cond(PORE, tn[f], pv, kn[f]);
        if (cl != 1) rhs += fc * dn[f];
        else sc[f >> 1] = fc / (sv * sn[f]);
    }
    if (zfaces && i == 0) rhs += pore_link(pv, 1.0e-9, viscosity(PORE)) * ptop;
    return rhs / sv;
}

// Right-hand side (Air neighbours' evaporation source, Dirichlet neighbours,
// the inlet plane), the scaled warm start, and the scaled +faces
// c / (sqrt(d_v) sqrt(d_n)) between two solved rows.  Only the faces these
// need are evaluated: those to unsolved neighbours and the solved +faces.
__global__ void k_press_rhs_scale(const uint8_t* __restrict__ img, const double* __restrict__ perm,
                                  const int8_t* __restrict__ cls, const double* __restrict__ pD,
                                  const double* __restrict__ p, const double* __restrict__ sd,
                                  double* __restrict__ bhat, double* __restrict__ x,
                                  double* __restrict__ hx, double* __restrict__ hy,
                                  double* __restrict__ hz, float* __restrict__ fx,
                                  float* __restrict__ fy, float* __restrict__ fz, Geo g,
                                  int zfaces, double ptop, double evap_dx) {
    const int k = blockIdx.x * TK + threadIdx.x;
    const int j = blockIdx.y * TJ + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nk || j >= g.nj) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    if (cls[v] != 1) {
        bhat[v] = 0.0;
        x[v] = 0.0;
        hx[v] = 0.0; hy[v] = 0.0; hz[v] = 0.0;
        fx[v] = 0.0f; fy[v] = 0.0f; fz[v] = 0.0f;
        return;
    }
    uint32_t n[6];
    bool ok[6];
    nbr6(g, i, j, k, v, n, ok);
    // all neighbour data first (one memory round trip), then the rows
    uint8_t tn[6];
    int8_t cn[6];
    double kn[6], dn[6], sn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        tn[f] = img[n[f]];
        cn[f] = cls[n[f]];
        kn[f] = perm[n[f]];
        dn[f] = pD[n[f]];
        sn[f] = sd[n[f]];
    }
    const double pv = perm[v];
    const double sv = sd[v];
    const double pvv = p[v];
    double sc[3];
    bhat[v] = press_rhs_core(tn, cn, ok, pv, sv, kn, dn, sn, i, zfaces, ptop, evap_dx, sc);
    x[v] = sv * pvv;  // warm start from the previous pressure, scaled variables
    hx[v] = sc[0]; hy[v] = sc[1]; hz[v] = sc[2];
    fx[v] = (float)sc[0]; fy[v] = (float)sc[1]; fz[v] = (float)sc[2];
}

// The same two passes with a voxel pair (k = 2m, 2m + 1) per lane, nk even:
// labels and classes as 2-byte pairs, doubles as 16-byte pairs, the k
// neighbours outside the pair from the neighbouring lanes (the wave's edge
// lanes load theirs), and the per-voxel arithmetic of press_diag_core /
// press_rhs_core (the same bits).  A block is 64 lanes x PP_TR waves, each
// wave walks PP_RPW rows (i, j).
constexpr int PP_TR = 4, PP_RPW = 4;
typedef double pd2 __attribute__((ext_vector_type(2)));
typedef float pf2 __attribute__((ext_vector_type(2)));

struct PairGeo {
    uint32_t nhr, nrows;   // pairs per row (nk / 2), rows (ni nj)
    uint32_t si, sj;
    int ni, nj;
};

static PairGeo pair_geo(const ls_ctx* c) {
    PairGeo pg;
    pg.nhr = (uint32_t)(c->nk / 2);
    pg.nrows = (uint32_t)c->ni * (uint32_t)c->nj;
    pg.sj = (uint32_t)c->nk;
    pg.si = (uint32_t)c->nk * (uint32_t)c->nj;
    pg.ni = c->ni;
    pg.nj = c->nj;
    return pg;
}
static dim3 pair_grid(const PairGeo& pg) {
    return row_grid((pg.nhr + WAVE - 1) / WAVE, (pg.nrows + PP_TR * PP_RPW - 1) / (PP_TR * PP_RPW));
}
// the voxel-pair kernels: nk even and one block partial each within the
// partials array
static bool use_pairs(const ls_ctx* c) {
    const dim3 gr = pair_grid(pair_geo(c));
    return c->nk % 2 == 0 && (size_t)gr.x * gr.y * gr.z <= (size_t)MAX_PARTIALS;
}

__device__ __forceinline__ uint32_t ld_b2(const void* __restrict__ a, uint32_t v) {
    return (uint32_t)(*(const uint16_t*)((const uint8_t*)a + v));
}
__device__ __forceinline__ pd2 ld_d2(const double* __restrict__ a, uint32_t v) {
    return *(const pd2*)(a + v);
}
// the k neighbours outside the pair: the left lane's second value and the
// right lane's first (the edge lanes load them, m the lane's pair index)
template <typename T>
__device__ __forceinline__ void k_out(T lo, T hi, const T* __restrict__ a, uint32_t v0,
                                      uint32_t m, uint32_t nhr, bool ok, T& left, T& right) {
    const int lane = threadIdx.x;
    left = __shfl_up(hi, 1, WAVE);
    right = __shfl_down(lo, 1, WAVE);
    if (lane == 0) left = (ok && m >= 1) ? a[v0 - 1] : T(0);
    if (lane == WAVE - 1) right = (ok && m + 1 < nhr) ? a[v0 + 2] : T(0);
}

__global__ __launch_bounds__(WAVE * PP_TR) void k_press_diag_pair(
    const uint8_t* __restrict__ img, const double* __restrict__ perm,
    const double* __restrict__ p, int8_t* __restrict__ cls, double* __restrict__ pD,
    double* __restrict__ sd, PairGeo g, int zfaces, double ptop) {
    const uint32_t m = blockIdx.x * WAVE + threadIdx.x;
    const bool ok_ = m < g.nhr;
    const int nk = 2 * (int)g.nhr;
    const pd2 z2 = pd2{0.0, 0.0};
#pragma unroll 1
    for (int q8 = 0; q8 < PP_RPW; ++q8) {
        const uint32_t row = (row_block() * PP_RPW + q8) * PP_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const uint32_t v0 = row * (uint32_t)nk + 2 * m;
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        const uint32_t lc = ok_ ? ld_b2(img, v0) : 0u;
        const uint32_t lip = (ok_ && hip_) ? ld_b2(img, v0 + g.si) : 0u;
        const uint32_t lim = (ok_ && him) ? ld_b2(img, v0 - g.si) : 0u;
        const uint32_t ljp = (ok_ && hjp) ? ld_b2(img, v0 + g.sj) : 0u;
        const uint32_t ljm = (ok_ && hjm) ? ld_b2(img, v0 - g.sj) : 0u;
        const pd2 kc = ok_ ? ld_d2(perm, v0) : z2;
        const pd2 kip = (ok_ && hip_) ? ld_d2(perm, v0 + g.si) : z2;
        const pd2 kim = (ok_ && him) ? ld_d2(perm, v0 - g.si) : z2;
        const pd2 kjp = (ok_ && hjp) ? ld_d2(perm, v0 + g.sj) : z2;
        const pd2 kjm = (ok_ && hjm) ? ld_d2(perm, v0 - g.sj) : z2;
        const pd2 pc = ok_ ? ld_d2(p, v0) : z2;
        int l_left, l_right;
        k_out<int>((int)(lc & 0xffu), (int)(lc >> 8), (const int*)nullptr, 0u, m, g.nhr, false,
                   l_left, l_right);
        uint32_t kl = (uint32_t)l_left, kr = (uint32_t)l_right;
        if (threadIdx.x == 0) kl = (ok_ && m >= 1) ? (uint32_t)img[v0 - 1] : 0u;
        if (threadIdx.x == WAVE - 1) kr = (ok_ && m + 1 < g.nhr) ? (uint32_t)img[v0 + 2] : 0u;
        double kkl, kkr;
        k_out<double>(kc.x, kc.y, perm, v0, m, g.nhr, ok_, kkl, kkr);
        if (!ok_) continue;
        PressDiag o[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int k = 2 * (int)m + e;
            const bool okf[6] = {hip_, him, hjp, hjm, k + 1 < nk, k > 0};
            const uint8_t tn[6] = {(uint8_t)(e ? (lip >> 8) : (lip & 0xffu)),
                                   (uint8_t)(e ? (lim >> 8) : (lim & 0xffu)),
                                   (uint8_t)(e ? (ljp >> 8) : (ljp & 0xffu)),
                                   (uint8_t)(e ? (ljm >> 8) : (ljm & 0xffu)),
                                   (uint8_t)(e ? kr : (lc >> 8)),
                                   (uint8_t)(e ? (lc & 0xffu) : kl)};
            const double kn[6] = {e ? kip.y : kip.x, e ? kim.y : kim.x, e ? kjp.y : kjp.x,
                                  e ? kjm.y : kjm.x, e ? kkr : kc.y, e ? kc.x : kkl};
            const uint8_t tv = (uint8_t)(e ? (lc >> 8) : (lc & 0xffu));
            o[e] = press_diag_core(tv, tn, okf, e ? kc.y : kc.x, kn, e ? pc.y : pc.x, i, g.ni,
                                   zfaces, ptop);
        }
        *(uint16_t*)(cls + v0) = (uint16_t)((uint8_t)o[0].cls | ((uint32_t)(uint8_t)o[1].cls << 8));
        __builtin_nontemporal_store(pd2{o[0].pD, o[1].pD}, (pd2*)(pD + v0));
        __builtin_nontemporal_store(pd2{o[0].sd, o[1].sd}, (pd2*)(sd + v0));
    }
}

__global__ __launch_bounds__(WAVE * PP_TR) void k_press_rhs_scale_pair(
    const uint8_t* __restrict__ img, const double* __restrict__ perm,
    const int8_t* __restrict__ cls, const double* __restrict__ pD,
    const double* __restrict__ p, const double* __restrict__ sd, double* __restrict__ bhat,
    double* __restrict__ x, double* __restrict__ hx, double* __restrict__ hy,
    double* __restrict__ hz, float* __restrict__ fx, float* __restrict__ fy,
    float* __restrict__ fz, PairGeo g, int zfaces, double ptop, double evap_dx) {
    const uint32_t m = blockIdx.x * WAVE + threadIdx.x;
    const bool ok_ = m < g.nhr;
    const int nk = 2 * (int)g.nhr;
    const pd2 z2 = pd2{0.0, 0.0};
#pragma unroll 1
    for (int q8 = 0; q8 < PP_RPW; ++q8) {
        const uint32_t row = (row_block() * PP_RPW + q8) * PP_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const uint32_t v0 = row * (uint32_t)nk + 2 * m;
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        const uint32_t cc = ok_ ? ld_b2(cls, v0) : 0u;
        // the pair's solved voxels need their neighbours; a pair with none
        // writes zeros
        const bool any = ok_ && (((cc & 0xffu) == 1u) || ((cc >> 8) == 1u));
        uint32_t lip = 0, lim = 0, ljp = 0, ljm = 0, lc = 0;
        uint32_t cip = 0, cim = 0, cjp = 0, cjm = 0;
        pd2 kc = z2, kip = z2, kim = z2, kjp = z2, kjm = z2;
        pd2 dc = z2, dip = z2, dim = z2, djp = z2, djm = z2;
        pd2 sc2 = z2, sip = z2, sim = z2, sjp = z2, sjm = z2, pc = z2;
        // the pair's own values serve the neighbouring lanes' k faces too
        if (ok_) {
            lc = ld_b2(img, v0);
            kc = ld_d2(perm, v0);
            dc = ld_d2(pD, v0);
            sc2 = ld_d2(sd, v0);
        }
        if (any) {
            pc = ld_d2(p, v0);
            if (hip_) {
                lip = ld_b2(img, v0 + g.si); cip = ld_b2(cls, v0 + g.si);
                kip = ld_d2(perm, v0 + g.si); dip = ld_d2(pD, v0 + g.si); sip = ld_d2(sd, v0 + g.si);
            }
            if (him) {
                lim = ld_b2(img, v0 - g.si); cim = ld_b2(cls, v0 - g.si);
                kim = ld_d2(perm, v0 - g.si); dim = ld_d2(pD, v0 - g.si); sim = ld_d2(sd, v0 - g.si);
            }
            if (hjp) {
                ljp = ld_b2(img, v0 + g.sj); cjp = ld_b2(cls, v0 + g.sj);
                kjp = ld_d2(perm, v0 + g.sj); djp = ld_d2(pD, v0 + g.sj); sjp = ld_d2(sd, v0 + g.sj);
            }
            if (hjm) {
                ljm = ld_b2(img, v0 - g.sj); cjm = ld_b2(cls, v0 - g.sj);
                kjm = ld_d2(perm, v0 - g.sj); djm = ld_d2(pD, v0 - g.sj); sjm = ld_d2(sd, v0 - g.sj);
            }
        }
        // k neighbours outside the pair (every lane takes part in the
        // shuffles; the edge lanes load theirs when their pair needs them)
        int t_l, t_r, c_l, c_r;
        k_out<int>((int)(lc & 0xffu), (int)(lc >> 8), (const int*)nullptr, 0u, m, g.nhr, false,
                   t_l, t_r);
        k_out<int>((int)(cc & 0xffu), (int)(cc >> 8), (const int*)nullptr, 0u, m, g.nhr, false,
                   c_l, c_r);
        double k_l, k_r, d_l, d_r, s_l, s_r;
        k_out<double>(kc.x, kc.y, perm, v0, m, g.nhr, any, k_l, k_r);
        k_out<double>(dc.x, dc.y, pD, v0, m, g.nhr, any, d_l, d_r);
        k_out<double>(sc2.x, sc2.y, sd, v0, m, g.nhr, any, s_l, s_r);
        if (threadIdx.x == 0) {
            t_l = (any && m >= 1) ? (int)img[v0 - 1] : 0;
            c_l = (any && m >= 1) ? (int)(uint8_t)cls[v0 - 1] : 0;
        }
        if (threadIdx.x == WAVE - 1) {
            t_r = (any && m + 1 < g.nhr) ? (int)img[v0 + 2] : 0;
            c_r = (any && m + 1 < g.nhr) ? (int)(uint8_t)cls[v0 + 2] : 0;
        }
        if (!ok_) continue;
        double bh[2] = {0.0, 0.0}, xs[2] = {0.0, 0.0}, f0[2] = {0.0, 0.0}, f1[2] = {0.0, 0.0},
               f2[2] = {0.0, 0.0};
        if (any) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const uint32_t ce = e ? (cc >> 8) : (cc & 0xffu);
                if (ce != 1u) continue;
                const int k = 2 * (int)m + e;
                const bool okf[6] = {hip_, him, hjp, hjm, k + 1 < nk, k > 0};
                const uint8_t tn[6] = {(uint8_t)(e ? (lip >> 8) : (lip & 0xffu)),
                                       (uint8_t)(e ? (lim >> 8) : (lim & 0xffu)),
                                       (uint8_t)(e ? (ljp >> 8) : (ljp & 0xffu)),
                                       (uint8_t)(e ? (ljm >> 8) : (ljm & 0xffu)),
                                       (uint8_t)(e ? (uint32_t)t_r : (lc >> 8)),
                                       (uint8_t)(e ? (lc & 0xffu) : (uint32_t)t_l)};
                const int8_t cn[6] = {(int8_t)(e ? (cip >> 8) : (cip & 0xffu)),
                                      (int8_t)(e ? (cim >> 8) : (cim & 0xffu)),
                                      (int8_t)(e ? (cjp >> 8) : (cjp & 0xffu)),
                                      (int8_t)(e ? (cjm >> 8) : (cjm & 0xffu)),
                                      (int8_t)(e ? (uint32_t)c_r : (cc >> 8)),
                                      (int8_t)(e ? (cc & 0xffu) : (uint32_t)c_l)};
                const double kn[6] = {e ? kip.y : kip.x, e ? kim.y : kim.x, e ? kjp.y : kjp.x,
                                      e ? kjm.y : kjm.x, e ? k_r : kc.y, e ? kc.x : k_l};
                const double dn[6] = {e ? dip.y : dip.x, e ? dim.y : dim.x, e ? djp.y : djp.x,
                                      e ? djm.y : djm.x, e ? d_r : dc.y, e ? dc.x : d_l};
                const double sn[6] = {e ? sip.y : sip.x, e ? sim.y : sim.x, e ? sjp.y : sjp.x,
                                      e ? sjm.y : sjm.x, e ? s_r : sc2.y, e ? sc2.x : s_l};
                const double sv = e ? sc2.y : sc2.x;
                double scf[3];
                bh[e] = press_rhs_core(tn, cn, okf, e ? kc.y : kc.x, sv, kn, dn, sn, i, zfaces,
                                       ptop, evap_dx, scf);
                xs[e] = sv * (e ? pc.y : pc.x);
                f0[e] = scf[0];
                f1[e] = scf[1];
                f2[e] = scf[2];
            }
        }
        // written for later kernels only: nontemporal stores
        __builtin_nontemporal_store(pd2{bh[0], bh[1]}, (pd2*)(bhat + v0));
        __builtin_nontemporal_store(pd2{xs[0], xs[1]}, (pd2*)(x + v0));
        __builtin_nontemporal_store(pd2{f0[0], f0[1]}, (pd2*)(hx + v0));
        __builtin_nontemporal_store(pd2{f1[0], f1[1]}, (pd2*)(hy + v0));
        __builtin_nontemporal_store(pd2{f2[0], f2[1]}, (pd2*)(hz + v0));
        __builtin_nontemporal_store(pf2{(float)f0[0], (float)f0[1]}, (pf2*)(fx + v0));
        __builtin_nontemporal_store(pf2{(float)f1[0], (float)f1[1]}, (pf2*)(fy + v0));
        __builtin_nontemporal_store(pf2{(float)f2[0], (float)f2[1]}, (pf2*)(fz + v0));
    }
}

__global__ void k_count_floored(const int8_t* __restrict__ cls, uint32_t N, double* partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double cnt = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        cnt += (cls[v] == 2) ? 1.0 : 0.0;
    double a[1] = {cnt};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

static double effective_top_pressure(double top, double t) {
    const double ramp_duration = 3600.0;  // src/Simulation.cpp:662-670
    double eff = top;
    if (t < ramp_duration) eff = top * (t / ramp_duration) + 1e-6;
    return eff;
}

int assemble_pressure(ls_ctx* c, double t) {
    c->trans_valid = false;     // the pressure phase lays its buffers over the pool
    Geo g = geo_of(c);
    const double ptop = effective_top_pressure(c->P.top_pressure, t);
    const int zf = c->P.bc_mode == LS_BC_ZFACES;
    dim3 blk(TK, TJ), grd((c->nk + TK - 1) / TK, (c->nj + TJ - 1) / TJ, c->ni);
    const PairGeo pg = pair_geo(c);
    const dim3 pgrid = pair_grid(pg);
    const bool pair = use_pairs(c);
    if (pair)
        hipLaunchKernelGGL(k_press_diag_pair, pgrid, dim3(WAVE, PP_TR), 0, c->stream, c->img,
                           c->perm, c->p, c->cls, c->pD, c->sd, pg, zf, ptop);
    else
        hipLaunchKernelGGL(k_press_diag, grd, blk, 0, c->stream, c->img, c->perm, c->p, c->cls,
                           c->pD, c->sd, g, zf, ptop);
    LS_LAUNCH_CHECK();
    if (pair)
        hipLaunchKernelGGL(k_press_rhs_scale_pair, pgrid, dim3(WAVE, PP_TR), 0, c->stream, c->img,
                           c->perm, c->cls, c->pD, c->p, c->sd, c->bhat, c->x64, c->c64[0],
                           c->c64[1], c->c64[2], c->c32[0], c->c32[1], c->c32[2], pg, zf, ptop,
                           c->P.evap_flux * c->P.dx);
    else
        hipLaunchKernelGGL(k_press_rhs_scale, grd, blk, 0, c->stream, c->img, c->perm, c->cls,
                           c->pD, c->p, c->sd, c->bhat, c->x64, c->c64[0], c->c64[1], c->c64[2],
                           c->c32[0], c->c32[1], c->c32[2], g, zf, ptop, c->P.evap_flux * c->P.dx);
    LS_LAUNCH_CHECK();
    uint32_t N = (uint32_t)c->N;
    int nb = ew_blocks(N);
    hipLaunchKernelGGL(k_count_floored, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->cls, N,
                       c->partials);
    LS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_finalize_one<RED_SUM>, dim3(1), dim3(1024), 0, c->stream, c->partials,
                       nb, &c->ss->n_floored);
    LS_LAUNCH_CHECK();
    c->press_assembled = true;
    return 0;
}

// After the solve: p = x/sqrt(d) on solved rows, the floored rows from their
// own reference row 1e-16 p_f + sum_n c p_n = source (src/Simulation.cpp:
// 783-797), the Dirichlet value everywhere else.
__global__ void k_press_final(const uint8_t* __restrict__ img, const double* __restrict__ perm,
                              const int8_t* __restrict__ cls, const double* __restrict__ pD,
                              const double* __restrict__ sd, const double* __restrict__ x,
                              double* __restrict__ p, Geo g, int zfaces, double ptop,
                              double evap_dx) {
    const int k = blockIdx.x * TK + threadIdx.x;
    const int j = blockIdx.y * TJ + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nk || j >= g.nj) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    const int8_t cl = cls[v];
    if (cl == 1) {
        __builtin_nontemporal_store(x[v] / sd[v], p + v);
    } else if (cl == 2) {
        const double pv = perm[v];
        double source_term = 0.0, sum = 0.0;
        for (int f = 0; f < 6; ++f) {
            uint32_t n;
            if (!nbr(g, i, j, k, v, f, n)) continue;
            if (img[n] == AIR) source_term -= evap_dx;
            double pn = (cls[n] == 1) ? x[n] / sd[n] : pD[n];
            sum += face_cond(PORE, img[n], pv, perm[n]) * pn;
        }
        if (zfaces && i == 0) sum += pore_link(pv, 1.0e-9, viscosity(PORE)) * ptop;
        __builtin_nontemporal_store((source_term - sum) / 1e-16, p + v);
    } else {
        __builtin_nontemporal_store(pD[v], p + v);
    }
}

int finalize_pressure(ls_ctx* c, double ptop) {
    // after the step's snapshot p still holds the snapshot's values: the new
    // field goes into the backup's buffer and the two swap (k_press_final
    // writes every voxel and reads no p)
    if (c->bp_stale) {
        std::swap(c->p, c->b_p);
        c->bp_stale = false;
    }
    Geo g = geo_of(c);
    dim3 blk(TK, TJ), grd((c->nk + TK - 1) / TK, (c->nj + TJ - 1) / TJ, c->ni);
    hipLaunchKernelGGL(k_press_final, grd, blk, 0, c->stream, c->img, c->perm, c->cls, c->pD,
                       c->sd, c->x64, c->p, g, c->P.bc_mode == LS_BC_ZFACES, ptop,
                       c->P.evap_flux * c->P.dx);
    LS_LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------
// Darcy flux (src/Simulation.cpp:866-953); one-sided differences only at the
// global faces of the single domain.
// ---------------------------------------------------------------------------
// Central difference inside, one-sided at the two ends of the axis; fm, fc,
// fp are F at pos - 1, pos, pos + 1 (the missing one unused).
__device__ __forceinline__ double grad1(double fm, double fc, double fp, int pos, int ext,
                                        double dx) {
    if (pos == 0) return (fp - fc) / dx;
    if (pos == ext - 1) return (fc - fm) / dx;
    return (fp - fm) / (2.0 * dx);
}

// One voxel's Darcy flux from p and Pc at the voxel and its neighbours (the
// voxel's own values where a neighbour is missing) and its permeability.
__device__ __forceinline__ void flux_core(uint8_t tv, double pv, double cv, double kv,
                                          const double (&pn)[6], const double (&cn)[6], int i,
                                          int j, int k, const Geo& g, double dx, double mcg,
                                          double& qxo, double& qyo, double& qzo) {
    double gradP_x = grad1(pn[1], pv, pn[0], i, g.ni, dx);
    double gradP_y = grad1(pn[3], pv, pn[2], j, g.nj, dx);
    double gradP_z = grad1(pn[5], pv, pn[4], k, g.nk, dx);
    double gradPc_x = grad1(cn[1], cv, cn[0], i, g.ni, dx);
    double gradPc_y = grad1(cn[3], cv, cn[2], j, g.nj, dx);
    double gradPc_z = grad1(cn[5], cv, cn[4], k, g.nk, dx);
    gradPc_x = dmax(-mcg, dmin(mcg, gradPc_x));
    gradPc_y = dmax(-mcg, dmin(mcg, gradPc_y));
    gradPc_z = dmax(-mcg, dmin(mcg, gradPc_z));
    double transmissivity = kv / viscosity(tv);
    const double rho_g = 1000.0 * 9.81;
    double liq_grad_x = gradP_x - gradPc_x;
    double liq_grad_y = gradP_y - gradPc_y;
    double liq_grad_z = gradP_z - gradPc_z;
    qxo = -transmissivity * (liq_grad_x - rho_g);
    qyo = -transmissivity * liq_grad_y;
    qzo = -transmissivity * liq_grad_z;
}

__global__ void k_flux(const uint8_t* __restrict__ img, const double* __restrict__ p,
                       const double* __restrict__ pcap, const double* __restrict__ perm,
                       double* __restrict__ qx, double* __restrict__ qy, double* __restrict__ qz,
                       Geo g, double dx, double mcg) {
    const int k = blockIdx.x * TK + threadIdx.x;
    const int j = blockIdx.y * TJ + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nk || j >= g.nj) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    uint32_t n[6];
    bool ok[6];
    nbr6(g, i, j, k, v, n, ok);
    // all loads first (a missing neighbour reads v itself, unused)
    double pn[6], cn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        pn[f] = p[n[f]];
        cn[f] = pcap[n[f]];
    }
    double a, b, c;
    flux_core(img[v], p[v], pcap[v], perm[v], pn, cn, i, j, k, g, dx, mcg, a, b, c);
    qx[v] = a;
    qy[v] = b;
    qz[v] = c;
}

// The same with a voxel pair per lane (nk even; rows (i, j) per wave as
// k_press_diag_pair): p, Pc and K as 16-byte pairs, the k neighbours outside
// the pair from the neighbouring lanes, the fluxes written as pairs.
__global__ __launch_bounds__(WAVE * PP_TR) void k_flux_pair(
    const uint8_t* __restrict__ img, const double* __restrict__ p,
    const double* __restrict__ pcap, const double* __restrict__ perm, double* __restrict__ qx,
    double* __restrict__ qy, double* __restrict__ qz, PairGeo pg, Geo g, double dx, double mcg) {
    const uint32_t m = blockIdx.x * WAVE + threadIdx.x;
    const bool ok_ = m < pg.nhr;
    const pd2 z2 = pd2{0.0, 0.0};
#pragma unroll 1
    for (int q8 = 0; q8 < PP_RPW; ++q8) {
        const uint32_t row = (row_block() * PP_RPW + q8) * PP_TR + threadIdx.y;
        if (row >= pg.nrows) continue;                  // uniform per wave
        const int i = (int)(row / (uint32_t)pg.nj), j = (int)(row - (uint32_t)i * pg.nj);
        const uint32_t v0 = row * (uint32_t)g.nk + 2 * m;
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        const pd2 pc = ok_ ? ld_d2(p, v0) : z2, cc = ok_ ? ld_d2(pcap, v0) : z2;
        const pd2 kc = ok_ ? ld_d2(perm, v0) : z2;
        const uint32_t lc = ok_ ? ld_b2(img, v0) : 0u;
        // a missing neighbour takes the voxel's own values (as k_flux)
        const pd2 pip = (ok_ && hip_) ? ld_d2(p, v0 + g.si) : pc;
        const pd2 pim = (ok_ && him) ? ld_d2(p, v0 - g.si) : pc;
        const pd2 pjp = (ok_ && hjp) ? ld_d2(p, v0 + g.sj) : pc;
        const pd2 pjm = (ok_ && hjm) ? ld_d2(p, v0 - g.sj) : pc;
        const pd2 cip = (ok_ && hip_) ? ld_d2(pcap, v0 + g.si) : cc;
        const pd2 cim = (ok_ && him) ? ld_d2(pcap, v0 - g.si) : cc;
        const pd2 cjp = (ok_ && hjp) ? ld_d2(pcap, v0 + g.sj) : cc;
        const pd2 cjm = (ok_ && hjm) ? ld_d2(pcap, v0 - g.sj) : cc;
        double p_l, p_r, c_l, c_r;
        k_out<double>(pc.x, pc.y, p, v0, m, pg.nhr, ok_, p_l, p_r);
        k_out<double>(cc.x, cc.y, pcap, v0, m, pg.nhr, ok_, c_l, c_r);
        if (!ok_) continue;
        const int k0 = 2 * (int)m;
        double qo[3][2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int k = k0 + e;
            const double pv = e ? pc.y : pc.x, cv = e ? cc.y : cc.x;
            const double pkp = e ? (k + 1 < g.nk ? p_r : pv) : pc.y;
            const double pkm = e ? pc.x : (k > 0 ? p_l : pv);
            const double ckp = e ? (k + 1 < g.nk ? c_r : cv) : cc.y;
            const double ckm = e ? cc.x : (k > 0 ? c_l : cv);
            const double pn[6] = {e ? pip.y : pip.x, e ? pim.y : pim.x, e ? pjp.y : pjp.x,
                                  e ? pjm.y : pjm.x, pkp, pkm};
            const double cn[6] = {e ? cip.y : cip.x, e ? cim.y : cim.x, e ? cjp.y : cjp.x,
                                  e ? cjm.y : cjm.x, ckp, ckm};
            flux_core((uint8_t)(e ? (lc >> 8) : (lc & 0xffu)), pv, cv, e ? kc.y : kc.x, pn, cn, i,
                      j, k, g, dx, mcg, qo[0][e], qo[1][e], qo[2][e]);
        }
        __builtin_nontemporal_store(pd2{qo[0][0], qo[0][1]}, (pd2*)(qx + v0));
        __builtin_nontemporal_store(pd2{qo[1][0], qo[1][1]}, (pd2*)(qy + v0));
        __builtin_nontemporal_store(pd2{qo[2][0], qo[2][1]}, (pd2*)(qz + v0));
    }
}

int phase_flux(ls_ctx* c) {
    Geo g = geo_of(c);
    dim3 blk(TK, TJ), grd((c->nk + TK - 1) / TK, (c->nj + TJ - 1) / TJ, c->ni);
    if (use_pairs(c))
        hipLaunchKernelGGL(k_flux_pair, pair_grid(pair_geo(c)), dim3(WAVE, PP_TR), 0, c->stream,
                           c->img, c->p, c->pc, c->perm, c->qx, c->qy, c->qz, pair_geo(c), g,
                           c->P.dx, c->P.max_cap_grad);
    else
        hipLaunchKernelGGL(k_flux, grd, blk, 0, c->stream, c->img, c->p, c->pc, c->perm, c->qx,
                           c->qy, c->qz, g, c->P.dx, c->P.max_cap_grad);
    LS_LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------
// Explicit saturation (src/Simulation.cpp:1672-1856).  The sub-step increment
// of a voxel is the same in every sub-step (q is fixed), so the N clamped
// updates run in registers: bit-identical to N full sweeps.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sat_inactive(uint8_t t) {
    return t == AIR || t == ROCK || t == SUL || t == PREC;
}

// One voxel's div q (cell fluxes averaged to its faces) and evaporation
// sink, from its own fluxes, the neighbours' fluxes along each axis (qn: +x,
// -x, +y, -y, +z, -z) and labels; ok[f]: the neighbour exists.
__device__ __forceinline__ void sat_rates_core(uint8_t tv, const uint8_t (&tn)[6],
                                               const bool (&ok)[6], double qxv, double qyv,
                                               double qzv, const double (&qn)[6], double dx,
                                               double evap, double& div_q, double& evap_sink) {
    double q_face_xp = ok[0] ? (qxv + qn[0]) / 2.0 : 0.0;
    double q_face_xm = ok[1] ? (qxv + qn[1]) / 2.0 : 0.0;
    double q_face_yp = ok[2] ? (qyv + qn[2]) / 2.0 : 0.0;
    double q_face_ym = ok[3] ? (qyv + qn[3]) / 2.0 : 0.0;
    double q_face_zp = ok[4] ? (qzv + qn[4]) / 2.0 : 0.0;
    double q_face_zm = ok[5] ? (qzv + qn[5]) / 2.0 : 0.0;
    div_q = ((q_face_xp - q_face_xm) + (q_face_yp - q_face_ym) + (q_face_zp - q_face_zm)) / dx;
    evap_sink = 0.0;
    if (tv == PORE) {
        int air_faces = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            if (ok[f] && tn[f] == AIR) air_faces++;
        evap_sink = (air_faces * evap) / dx;
    }
}

__device__ __forceinline__ void sat_rates(const uint8_t* img, const double* qx, const double* qy,
                                          const double* qz, const Geo& g, int i, int j, int k,
                                          uint32_t v, double dx, double evap, double& div_q,
                                          double& evap_sink) {
    uint32_t n[6];
    bool ok[6];
    nbr6(g, i, j, k, v, n, ok);
    // every load first (neighbour indices outside the grid are v itself)
    const double qxv = qx[v], qyv = qy[v], qzv = qz[v];
    const double qn[6] = {qx[n[0]], qx[n[1]], qy[n[2]], qy[n[3]], qz[n[4]], qz[n[5]]};
    const uint8_t tv = img[v];
    uint8_t tn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) tn[f] = img[n[f]];
    sat_rates_core(tv, tn, ok, qxv, qyv, qzv, qn, dx, evap, div_q, evap_sink);
}

// A voxel pair's rates (voxel-pair layout of k_press_diag_pair): fluxes as
// 16-byte pairs, labels as 2-byte pairs, the k neighbours outside the pair
// from the neighbouring lanes.  Every lane of the wave must call it.
struct SatPair {
    double div_q[2], evap_sink[2];
    uint32_t lc;           // the pair's labels
};
__device__ __forceinline__ SatPair sat_rates_pair(const uint8_t* __restrict__ img,
                                                  const double* __restrict__ qx,
                                                  const double* __restrict__ qy,
                                                  const double* __restrict__ qz,
                                                  const PairGeo& pg, int nk, uint32_t m,
                                                  bool ok_, int i, int j, uint32_t v0, double dx,
                                                  double evap) {
    const pd2 z2 = pd2{0.0, 0.0};
    const bool hip_ = i + 1 < pg.ni, him = i > 0, hjp = j + 1 < pg.nj, hjm = j > 0;
    const pd2 qxc = ok_ ? ld_d2(qx, v0) : z2, qyc = ok_ ? ld_d2(qy, v0) : z2;
    const pd2 qzc = ok_ ? ld_d2(qz, v0) : z2;
    const pd2 qxp = (ok_ && hip_) ? ld_d2(qx, v0 + pg.si) : z2;
    const pd2 qxm = (ok_ && him) ? ld_d2(qx, v0 - pg.si) : z2;
    const pd2 qyp = (ok_ && hjp) ? ld_d2(qy, v0 + pg.sj) : z2;
    const pd2 qym = (ok_ && hjm) ? ld_d2(qy, v0 - pg.sj) : z2;
    SatPair o;
    o.lc = ok_ ? ld_b2(img, v0) : 0u;
    const uint32_t lip = (ok_ && hip_) ? ld_b2(img, v0 + pg.si) : 0u;
    const uint32_t lim = (ok_ && him) ? ld_b2(img, v0 - pg.si) : 0u;
    const uint32_t ljp = (ok_ && hjp) ? ld_b2(img, v0 + pg.sj) : 0u;
    const uint32_t ljm = (ok_ && hjm) ? ld_b2(img, v0 - pg.sj) : 0u;
    double qz_l, qz_r;
    k_out<double>(qzc.x, qzc.y, qz, v0, m, pg.nhr, ok_, qz_l, qz_r);
    int t_l, t_r;
    k_out<int>((int)(o.lc & 0xffu), (int)(o.lc >> 8), (const int*)nullptr, 0u, m, pg.nhr, false,
               t_l, t_r);
    if (threadIdx.x == 0) t_l = (ok_ && m >= 1) ? (int)img[v0 - 1] : 0;
    if (threadIdx.x == WAVE - 1) t_r = (ok_ && m + 1 < pg.nhr) ? (int)img[v0 + 2] : 0;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int k = 2 * (int)m + e;
        const bool okf[6] = {hip_, him, hjp, hjm, k + 1 < nk, k > 0};
        const uint8_t tn[6] = {(uint8_t)(e ? (lip >> 8) : (lip & 0xffu)),
                               (uint8_t)(e ? (lim >> 8) : (lim & 0xffu)),
                               (uint8_t)(e ? (ljp >> 8) : (ljp & 0xffu)),
                               (uint8_t)(e ? (ljm >> 8) : (ljm & 0xffu)),
                               (uint8_t)(e ? (uint32_t)t_r : (o.lc >> 8)),
                               (uint8_t)(e ? (o.lc & 0xffu) : (uint32_t)t_l)};
        const double qn[6] = {e ? qxp.y : qxp.x, e ? qxm.y : qxm.x, e ? qyp.y : qyp.x,
                              e ? qym.y : qym.x, e ? qz_r : qzc.y, e ? qzc.x : qz_l};
        sat_rates_core((uint8_t)(e ? (o.lc >> 8) : (o.lc & 0xffu)), tn, okf, e ? qxc.y : qxc.x,
                       e ? qyc.y : qyc.x, e ? qzc.y : qzc.x, qn, dx, evap, o.div_q[e],
                       o.evap_sink[e]);
    }
    return o;
}

__global__ void k_sat_max(const uint8_t* __restrict__ img, const double* __restrict__ qx,
                          const double* __restrict__ qy, const double* __restrict__ qz, Geo g,
                          uint32_t N, double dx, double evap, double* partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double mx = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        if (sat_inactive(img[v])) continue;
        int i, j, k;
        decode(v, g, i, j, k);
        double div_q, evap_sink;
        sat_rates(img, qx, qy, qz, g, i, j, k, v, dx, evap, div_q, evap_sink);
        double total_rate = fabs(div_q) + evap_sink;
        if (total_rate > mx) mx = total_rate;
    }
    double a[1] = {mx};
    block_reduce<1, RED_MAX>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

__global__ void k_sat_finalize(const double* __restrict__ partials, int n, double dt,
                               double porosity, StepScalars* ss) {
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc = fmax(acc, partials[i]);
    double a[1] = {acc};
    block_reduce<1, RED_MAX>(a, lds);
    if (threadIdx.x == 0) {
        double global_max_div_q = a[0];
        double max_dsat = 0.1;
        double dt_safe = dt;
        if (global_max_div_q > 1e-20) dt_safe = (max_dsat * porosity) / global_max_div_q;
        double calculated_steps = ceil(dt / dt_safe);
        int num_steps = 1;
        if (isinf(calculated_steps) || isnan(calculated_steps)) num_steps = 1000;
        else if (calculated_steps > 1000) num_steps = 1000;
        else if (calculated_steps > 1.0) num_steps = (int)calculated_steps;
        double dt_sub = dt / num_steps;
        ss->sat_max = global_max_div_q;
        ss->sat_nsub = num_steps;
        ss->sat_dt_sub_over_phi = dt_sub / porosity;
    }
}

__global__ void k_sat_update(const uint8_t* __restrict__ img, const double* __restrict__ qx,
                             const double* __restrict__ qy, const double* __restrict__ qz,
                             double* __restrict__ s, Geo g, double dx, double evap,
                             const StepScalars* __restrict__ ss) {
    const int k = blockIdx.x * TK + threadIdx.x;
    const int j = blockIdx.y * TJ + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nk || j >= g.nj) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    if (sat_inactive(img[v])) return;  // inactive voxels keep their value
    double div_q, evap_sink;
    sat_rates(img, qx, qy, qz, g, i, j, k, v, dx, evap, div_q, evap_sink);
    const int nsub = ss->sat_nsub;
    const double dt_sub_over_phi = ss->sat_dt_sub_over_phi;
    double sv = s[v];
    for (int step = 0; step < nsub; ++step) {
        double new_sat = sv - dt_sub_over_phi * (div_q + evap_sink);
        sv = dmax(0.0, dmin(1.0, new_sat));
    }
    s[v] = sv;
}

__global__ __launch_bounds__(WAVE * PP_TR) void k_sat_max_pair(
    const uint8_t* __restrict__ img, const double* __restrict__ qx,
    const double* __restrict__ qy, const double* __restrict__ qz, PairGeo pg, int nk,
    double dx, double evap, double* partials) {
    __shared__ double lds[PP_TR];
    const uint32_t m = blockIdx.x * WAVE + threadIdx.x;
    const bool ok_ = m < pg.nhr;
    double mx = 0.0;
#pragma unroll 1
    for (int q8 = 0; q8 < PP_RPW; ++q8) {
        const uint32_t row = (row_block() * PP_RPW + q8) * PP_TR + threadIdx.y;
        if (row >= pg.nrows) continue;                  // uniform per wave
        const int i = (int)(row / (uint32_t)pg.nj), j = (int)(row - (uint32_t)i * pg.nj);
        const uint32_t v0 = row * (uint32_t)nk + 2 * m;
        const SatPair r = sat_rates_pair(img, qx, qy, qz, pg, nk, m, ok_, i, j, v0, dx, evap);
        if (!ok_) continue;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (sat_inactive((uint8_t)(e ? (r.lc >> 8) : (r.lc & 0xffu)))) continue;
            const double total_rate = fabs(r.div_q[e]) + r.evap_sink[e];
            if (total_rate > mx) mx = total_rate;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, WAVE));
    if (threadIdx.x == 0) lds[threadIdx.y] = mx;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double a = lds[0];
        for (int w = 1; w < PP_TR; ++w) a = fmax(a, lds[w]);
        partials[block_linear()] = a;
    }
}

__global__ __launch_bounds__(WAVE * PP_TR) void k_sat_update_pair(
    const uint8_t* __restrict__ img, const double* __restrict__ qx,
    const double* __restrict__ qy, const double* __restrict__ qz, double* __restrict__ s,
    PairGeo pg, int nk, double dx, double evap, const StepScalars* __restrict__ ss) {
    const uint32_t m = blockIdx.x * WAVE + threadIdx.x;
    const bool ok_ = m < pg.nhr;
    const int nsub = ss->sat_nsub;
    const double dt_sub_over_phi = ss->sat_dt_sub_over_phi;
#pragma unroll 1
    for (int q8 = 0; q8 < PP_RPW; ++q8) {
        const uint32_t row = (row_block() * PP_RPW + q8) * PP_TR + threadIdx.y;
        if (row >= pg.nrows) continue;                  // uniform per wave
        const int i = (int)(row / (uint32_t)pg.nj), j = (int)(row - (uint32_t)i * pg.nj);
        const uint32_t v0 = row * (uint32_t)nk + 2 * m;
        const SatPair r = sat_rates_pair(img, qx, qy, qz, pg, nk, m, ok_, i, j, v0, dx, evap);
        if (!ok_) continue;
        const bool act0 = !sat_inactive((uint8_t)(r.lc & 0xffu));
        const bool act1 = !sat_inactive((uint8_t)(r.lc >> 8));
        if (!act0 && !act1) continue;                   // inactive voxels keep their value
        pd2 sv = ld_d2(s, v0);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (!(e ? act1 : act0)) continue;
            double x = e ? sv.y : sv.x;
            for (int step = 0; step < nsub; ++step) {
                double new_sat = x - dt_sub_over_phi * (r.div_q[e] + r.evap_sink[e]);
                x = dmax(0.0, dmin(1.0, new_sat));
            }
            if (e) sv.y = x; else sv.x = x;
        }
        *(pd2*)(s + v0) = sv;
    }
}

int phase_saturation(ls_ctx* c, double dt, ls_step_report* rep) {
    if (use_pairs(c)) {
        const PairGeo pg = pair_geo(c);
        const dim3 pgrid = pair_grid(pg);
        const int nbp = (int)(pgrid.x * pgrid.y * pgrid.z);
        hipLaunchKernelGGL(k_sat_max_pair, pgrid, dim3(WAVE, PP_TR), 0, c->stream, c->img, c->qx,
                           c->qy, c->qz, pg, c->nk, c->P.dx, c->P.evap_flux, c->partials);
        hipLaunchKernelGGL(k_sat_finalize, dim3(1), dim3(1024), 0, c->stream, c->partials, nbp,
                           dt, c->P.porosity, c->ss);
        hipLaunchKernelGGL(k_sat_update_pair, pgrid, dim3(WAVE, PP_TR), 0, c->stream, c->img,
                           c->qx, c->qy, c->qz, c->s, pg, c->nk, c->P.dx, c->P.evap_flux, c->ss);
        LS_LAUNCH_CHECK();
        (void)rep;
        return 0;
    }
    Geo g = geo_of(c);
    uint32_t N = (uint32_t)c->N;
    int nb = ew_blocks(N);
    hipLaunchKernelGGL(k_sat_max, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->img, c->qx, c->qy,
                       c->qz, g, N, c->P.dx, c->P.evap_flux, c->partials);
    LS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sat_finalize, dim3(1), dim3(1024), 0, c->stream, c->partials, nb, dt,
                       c->P.porosity, c->ss);
    LS_LAUNCH_CHECK();
    dim3 blk(TK, TJ), grd((c->nk + TK - 1) / TK, (c->nj + TJ - 1) / TJ, c->ni);
    hipLaunchKernelGGL(k_sat_update, grd, blk, 0, c->stream, c->img, c->qx, c->qy, c->qz, c->s, g,
                       c->P.dx, c->P.evap_flux, c->ss);
    LS_LAUNCH_CHECK();
    (void)rep;
    return 0;
}

// ---------------------------------------------------------------------------
// Surface effects (src/Simulation.cpp:2063-2103): Pore with s > 0 touching
// Air becomes Precipitate with conc = 1.  Written to a second label buffer so
// no thread reads a label another thread is changing.
// ---------------------------------------------------------------------------
__global__ void k_surface(const uint8_t* __restrict__ img, uint8_t* __restrict__ img_out,
                          const double* __restrict__ s, double* __restrict__ conc, Geo g) {
    const int k = blockIdx.x * TK + threadIdx.x;
    const int j = blockIdx.y * TJ + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nk || j >= g.nj) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    uint8_t t = img[v];
    if (t == PORE && s[v] > 0.0) {
        bool is_surface_voxel = false;
        for (int f = 0; f < 6; ++f) {
            uint32_t n;
            if (nbr(g, i, j, k, v, f, n) && img[n] == AIR) { is_surface_voxel = true; break; }
        }
        if (is_surface_voxel) {
            conc[v] = 1.0;
            t = PREC;
        }
    }
    img_out[v] = t;
}

// The same with a voxel pair per lane (nk even; the layout of
// k_press_diag_pair): a Pore voxel with s > 0 and an Air neighbour becomes
// Precipitate with conc = 1.
__global__ __launch_bounds__(WAVE * PP_TR) void k_surface_pair(
    const uint8_t* __restrict__ img, uint8_t* __restrict__ img_out,
    const double* __restrict__ s, double* __restrict__ conc, PairGeo pg, int nk) {
    const uint32_t m = blockIdx.x * WAVE + threadIdx.x;
    const bool ok_ = m < pg.nhr;
#pragma unroll 1
    for (int q8 = 0; q8 < PP_RPW; ++q8) {
        const uint32_t row = (row_block() * PP_RPW + q8) * PP_TR + threadIdx.y;
        if (row >= pg.nrows) continue;                  // uniform per wave
        const int i = (int)(row / (uint32_t)pg.nj), j = (int)(row - (uint32_t)i * pg.nj);
        const uint32_t v0 = row * (uint32_t)nk + 2 * m;
        const uint32_t lc = ok_ ? ld_b2(img, v0) : 0u;
        const pd2 sc = ok_ ? ld_d2(s, v0) : pd2{0.0, 0.0};
        const bool c0 = (lc & 0xffu) == PORE && sc.x > 0.0;
        const bool c1 = (lc >> 8) == PORE && sc.y > 0.0;
        const bool need = c0 || c1;
        const bool hip_ = i + 1 < pg.ni, him = i > 0, hjp = j + 1 < pg.nj, hjm = j > 0;
        uint32_t lip = 0, lim = 0, ljp = 0, ljm = 0;
        if (need) {
            if (hip_) lip = ld_b2(img, v0 + pg.si);
            if (him) lim = ld_b2(img, v0 - pg.si);
            if (hjp) ljp = ld_b2(img, v0 + pg.sj);
            if (hjm) ljm = ld_b2(img, v0 - pg.sj);
        }
        int t_l, t_r;
        k_out<int>((int)(lc & 0xffu), (int)(lc >> 8), (const int*)nullptr, 0u, m, pg.nhr, false,
                   t_l, t_r);
        if (threadIdx.x == 0) t_l = (need && m >= 1) ? (int)img[v0 - 1] : 0;
        if (threadIdx.x == WAVE - 1) t_r = (need && m + 1 < pg.nhr) ? (int)img[v0 + 2] : 0;
        if (!ok_) continue;
        uint32_t out = lc;
        if (need) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if (!(e ? c1 : c0)) continue;
                const int k = 2 * (int)m + e;
                const bool okf[6] = {hip_, him, hjp, hjm, k + 1 < nk, k > 0};
                const uint32_t tn[6] = {e ? (lip >> 8) : (lip & 0xffu), e ? (lim >> 8) : (lim & 0xffu),
                                        e ? (ljp >> 8) : (ljp & 0xffu), e ? (ljm >> 8) : (ljm & 0xffu),
                                        e ? (uint32_t)t_r : (lc >> 8),
                                        e ? (lc & 0xffu) : (uint32_t)t_l};
                bool surface = false;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                    if (okf[f] && tn[f] == AIR) surface = true;
                if (surface) {
                    conc[v0 + e] = 1.0;
                    out = e ? ((out & 0xffu) | ((uint32_t)PREC << 8)) : ((out & 0xff00u) | PREC);
                }
            }
        }
        *(uint16_t*)(img_out + v0) = (uint16_t)out;
    }
}

int phase_surface(ls_ctx* c) {
    Geo g = geo_of(c);
    dim3 blk(TK, TJ), grd((c->nk + TK - 1) / TK, (c->nj + TJ - 1) / TJ, c->ni);
    if (use_pairs(c))
        hipLaunchKernelGGL(k_surface_pair, pair_grid(pair_geo(c)), dim3(WAVE, PP_TR), 0,
                           c->stream, c->img, c->img2, c->s, c->conc, pair_geo(c), c->nk);
    else
        hipLaunchKernelGGL(k_surface, grd, blk, 0, c->stream, c->img, c->img2, c->s, c->conc, g);
    LS_LAUNCH_CHECK();
    uint8_t* tmp = c->img;
    c->img = c->img2;
    c->img2 = tmp;
    return 0;
}

// The transport assembly zeroes q at Rock and Precipitate voxels
// (src/Simulation.cpp:1132-1150); the fluxes written back keep that.
__global__ void k_zero_flux(const uint8_t* __restrict__ img, double* qx, double* qy, double* qz,
                            uint32_t N) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        uint8_t t = img[v];
        if (t == PREC || t == ROCK) {
            qx[v] = 0.0;
            qy[v] = 0.0;
            qz[v] = 0.0;
        }
    }
}

int zero_flux_rock_precip(ls_ctx* c) {
    uint32_t N = (uint32_t)c->N;
    hipLaunchKernelGGL(k_zero_flux, dim3(ew_blocks(N)), dim3(EW_BLOCK), 0, c->stream, c->img,
                       c->qx, c->qy, c->qz, N);
    LS_LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------
// Precipitation (src/Simulation.cpp:2114-2165).  Pointwise.
// ---------------------------------------------------------------------------
__global__ void k_precip(uint8_t* __restrict__ img, double* __restrict__ conc,
                         double* __restrict__ inv, const double* __restrict__ s, uint32_t N,
                         double c_sat, int instant, double porosity, double solid_density,
                         double time_scalar) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        double cv = conc[v];
        if (img[v] == PORE && cv > c_sat) {
            double excess_conc = cv - c_sat;
            if (instant) {
                img[v] = PREC;
            } else {
                double fluid_vol_fraction = porosity * s[v];
                double precipitated_vol_fraction = (excess_conc * fluid_vol_fraction) / solid_density;
                double iv = inv[v] + (precipitated_vol_fraction * time_scalar);
                inv[v] = iv;
                double critical_limit = 0.90;
                if (iv > critical_limit) img[v] = PREC;
            }
            conc[v] = c_sat;
        }
    }
}

int phase_precipitation(ls_ctx* c) {
    uint32_t N = (uint32_t)c->N;
    hipLaunchKernelGGL(k_precip, dim3(ew_blocks(N)), dim3(EW_BLOCK), 0, c->stream, c->img,
                       c->conc, c->inv, c->s, N, c->P.c_sat, c->P.instant_precip, c->P.porosity,
                       c->P.solid_density, c->P.time_scalar);
    LS_LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------
// Step summary: min/max of p, conc, acid and FP64 sums of conc, acid, s (the
// per-step mass balance).
// ---------------------------------------------------------------------------
__global__ void k_summary(const double* __restrict__ p, const double* __restrict__ conc,
                          const double* __restrict__ acid, const double* __restrict__ s,
                          uint32_t N, double* partials) {
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    double pmn = 1.0 / 0.0, pmx = -1.0 / 0.0, cmn = 1.0 / 0.0, cmx = -1.0 / 0.0;
    double amn = 1.0 / 0.0, amx = -1.0 / 0.0, sc = 0.0, sa = 0.0, sw = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        double a = p[v], b = conc[v], d = acid != nullptr ? acid[v] : 0.0;
        pmn = fmin(pmn, a); pmx = fmax(pmx, a);
        cmn = fmin(cmn, b); cmx = fmax(cmx, b);
        amn = fmin(amn, d); amx = fmax(amx, d);
        sc += b; sa += d; sw += s[v];
    }
    double r1[4] = {pmn, cmn, amn, 0.0};
    block_reduce<4, RED_MIN, RED_MIN, RED_MIN, RED_SUM>(r1, lds);
    __syncthreads();
    double r2[3] = {pmx, cmx, amx};
    block_reduce<3, RED_MAX, RED_MAX, RED_MAX>(r2, lds);
    __syncthreads();
    double r3[3] = {sc, sa, sw};
    block_reduce<3, RED_SUM, RED_SUM, RED_SUM>(r3, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[0 * nb + blockIdx.x] = r1[0];
        partials[1 * nb + blockIdx.x] = r2[0];
        partials[2 * nb + blockIdx.x] = r1[1];
        partials[3 * nb + blockIdx.x] = r2[1];
        partials[4 * nb + blockIdx.x] = r1[2];
        partials[5 * nb + blockIdx.x] = r2[2];
        partials[6 * nb + blockIdx.x] = r3[0];
        partials[7 * nb + blockIdx.x] = r3[1];
        partials[8 * nb + blockIdx.x] = r3[2];
    }
}

__global__ void k_summary_finalize(const double* __restrict__ partials, int nb, StepScalars* ss) {
    __shared__ double lds[1024 / WAVE];
    for (int q = 0; q < 9; ++q) {
        const int op = (q == 0 || q == 2 || q == 4) ? RED_MIN : (q < 6) ? RED_MAX : RED_SUM;
        double acc = op == RED_MIN ? 1.0 / 0.0 : op == RED_MAX ? -1.0 / 0.0 : 0.0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            double b = partials[q * nb + i];
            acc = op == RED_MIN ? fmin(acc, b) : op == RED_MAX ? fmax(acc, b) : acc + b;
        }
        double a[1] = {acc};
        if (op == RED_MIN) block_reduce<1, RED_MIN>(a, lds);
        else if (op == RED_MAX) block_reduce<1, RED_MAX>(a, lds);
        else block_reduce<1, RED_SUM>(a, lds);
        if (threadIdx.x == 0) ss->sums[q] = a[0];
        __syncthreads();
    }
}

int summarize(ls_ctx* c, ls_step_report* rep) {
    uint32_t N = (uint32_t)c->N;
    int nb = ew_blocks(N);
    if (nb * 9 > MAX_PARTIALS * 8) nb = MAX_PARTIALS * 8 / 9;
    // acid known zero (+0.0 everywhere) reads as zeros without its loads
    hipLaunchKernelGGL(k_summary, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->p, c->conc,
                       c->acid_zero ? nullptr : c->acid, c->s, N, c->partials);
    LS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_summary_finalize, dim3(1), dim3(1024), 0, c->stream, c->partials, nb,
                       c->ss);
    LS_LAUNCH_CHECK();
    (void)rep;
    return 0;
}

}  // namespace ls

// ---------------------------------------------------------------------------
// updateSaturation on caller-owned device fields, allocating nothing.  The
// only device memory beyond the caller's inputs and outputs is the scratch
// of the global MAX (block partials) and the sub-step plan (a StepScalars),
// held in a file-scope device array.  Its device address is looked up once
// per device with hipGetSymbolAddress (the array's name is a host-side
// stand-in, not a device address) and only that address is used from the
// host.  Every call on one device shares the array: calls that overlap on
// one device race.  Nothing in it needs clearing: k_sat_max writes every
// partial that k_sat_finalize reads, and k_sat_finalize writes every field
// of the plan that k_sat_update reads.
// ---------------------------------------------------------------------------
#include <cstdio>
#include <mutex>

namespace {
constexpr int NOALLOC_PLAN_BYTES = 256;
constexpr int NOALLOC_MAX_PARTIALS = 8192;     // ew_blocks() never launches more
constexpr size_t NOALLOC_SCRATCH_BYTES =
    NOALLOC_PLAN_BYTES + sizeof(double) * (size_t)NOALLOC_MAX_PARTIALS;
constexpr int NOALLOC_MAX_DEVICES = 64;
static_assert(sizeof(StepScalars) <= NOALLOC_PLAN_BYTES, "plan slot too small");
static_assert(NOALLOC_SCRATCH_BYTES == LS_NOALLOC_SCRATCH_BYTES, "header and kernel disagree");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device (looked up once per
// device, with that device current); nullptr and a message on failure.
void* scratch_on_current_device(const char** why) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        *why = "no current HIP device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) {
            *why = "hipGetSymbolAddress(scratch) failed";
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// error text of the last failed call on this thread (a fixed buffer: the
// entry point allocates no memory on the host either)
thread_local char g_noalloc_err[256];

void set_err(const char* fmt, const char* a = "", long long b = 0, long long c = 0) {
    snprintf(g_noalloc_err, sizeof g_noalloc_err, fmt, a, b, c);
}
}  // namespace

extern "C" const char* ls_noalloc_last_error(void) { return g_noalloc_err; }

extern "C" int ls_update_saturation_noalloc(int ni, int nj, int nk, const uint8_t* img,
                                            const double* qx, const double* qy,
                                            const double* qz, double* s, double dt, double dx,
                                            double porosity, double evap_flux, void* stream,
                                            int* nsub, double* max_rate) {
    using namespace ls;
    g_noalloc_err[0] = 0;
    if (!img || !qx || !qy || !qz || !s) {
        set_err("ls_update_saturation_noalloc: null field pointer%s");
        return -1;
    }
    if (ni < 1 || nj < 1 || nk < 1 || ni > 65535 || (double)ni * nj * nk >= 4294967295.0) {
        set_err("ls_update_saturation_noalloc: grid must have 1 <= ni <= 65535, "
                "nj, nk >= 1 and ni*nj*nk < 2^32%s");
        return -2;
    }
    if (!(dt > 0.0) || !(dx > 0.0) || !(porosity > 0.0)) {
        set_err("ls_update_saturation_noalloc: dt, dx and porosity must be > 0%s");
        return -2;
    }
    const uint32_t N = (uint32_t)((size_t)ni * nj * nk);
    const int nb = ew_blocks(N);
    if (nb > NOALLOC_MAX_PARTIALS) {
        set_err("ls_update_saturation_noalloc:%s %lld block partials exceed the scratch's %lld",
                "", nb, NOALLOC_MAX_PARTIALS);
        return -2;
    }
    const char* why = "";
    unsigned char* base = (unsigned char*)scratch_on_current_device(&why);
    if (!base) {
        set_err("ls_update_saturation_noalloc: %s", why);
        return -10;
    }
    StepScalars* plan = (StepScalars*)base;
    double* partials = (double*)(base + NOALLOC_PLAN_BYTES);
    hipStream_t st = (hipStream_t)stream;
    Geo g;
    g.ni = ni;
    g.nj = nj;
    g.nk = nk;
    g.sj = (uint32_t)nk;
    g.si = (uint32_t)nk * (uint32_t)nj;
    hipLaunchKernelGGL(k_sat_max, dim3(nb), dim3(EW_BLOCK), 0, st, img, qx, qy, qz, g, N, dx,
                       evap_flux, partials);
    hipLaunchKernelGGL(k_sat_finalize, dim3(1), dim3(1024), 0, st, partials, nb, dt, porosity,
                       plan);
    dim3 blk(TK, TJ), grd((nk + TK - 1) / TK, (nj + TJ - 1) / TJ, ni);
    hipLaunchKernelGGL(k_sat_update, grd, blk, 0, st, img, qx, qy, qz, s, g, dx, evap_flux, plan);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_err("ls_update_saturation_noalloc: kernel launch: %s", hipGetErrorString(e));
        return -11;
    }
    if (nsub || max_rate) {
        // the plan's first fields; a synchronous read, so leave both null
        // where the call must stay asynchronous (under graph capture)
        StepScalars h;
        e = hipMemcpyAsync(&h, plan, sizeof(StepScalars), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_err("ls_update_saturation_noalloc: %s", hipGetErrorString(e));
            return -10;
        }
        if (nsub) *nsub = h.sat_nsub;
        if (max_rate) *max_rate = h.sat_max;
    }
    return 0;
}
