// k_transport.hip — advection-diffusion-reaction transport of conc and acid
// (replaces setupConcentrationEqns + solveConc + solveAcid,
// src/Simulation.cpp:966-1237, 1243-1295, 2167-2224).
//
// The transport matrix depends only on labels, fluxes and dt, so conc and
// acid share it (the reference re-assembles the identical matrix twice).  It
// is applied matrix-free: every matvec rebuilds the row coefficients from the
// labels and the three flux fields, in the reference's accumulation order,
// and applies them to both right-hand sides at once.  The solver is Jacobi
// (right) preconditioned BiCGStab run for the two systems in lock step, with
// per-system scalars on the device.
//
// The same solver also runs the implicit saturation system (KIND 1,
// setupSaturationEqns + solveSaturation, src/Simulation.cpp:1858-2053) with
// its own matrix-free row builder, as system 0 alone.
#include <cmath>
#include <type_traits>

#include "ls_context.h"

namespace ls {

struct TGeo {
    int ni, nj, nk;
    uint32_t si, sj;
};

struct AdrConst {
    double inv_dt, alpha, theta, D, Dpf, dx;
    double phi, evap;      // implicit saturation: porosity, evaporative flux
    // -D_eff / dx^2 next to a Pore and to other phases, as the row computes
    // them (the host evaluates the same expression, so the bits agree)
    double diff_pore, diff_other;
    double inv_dx;         // 1 / dx, exact when dx_pow2
    int dx_pow2;           // dx is a power of two: q / dx == q * (1 / dx)
};

// The extrapolated transport start: y = w[0] x + sum_k w[k + 1] h_k over the
// first n history fields h_0 (newest) .. h_2, Lagrange weights at the new time.
struct Extrap {
    int n;
    double w[4];
};

// One black voxel's start from its current field and the first ex.n
// history fields; the current field then goes into hw (the buffer the host
// makes the newest history field, the oldest one's, read before it is
// overwritten).
__device__ __forceinline__ double extrap_start(double xa, uint32_t m, const double* h0,
                                               const double* h1, const double* h2, double* hw,
                                               const Extrap& ex) {
    double xs = xa;
    if (ex.n >= 1) {
        xs = ex.w[0] * xa + ex.w[1] * h0[m];
        if (ex.n >= 2) xs += ex.w[2] * h1[m];
        if (ex.n >= 3) xs += ex.w[3] * h2[m];
    }
    if (hw != nullptr) hw[m] = xa;
    return xs;
}

__device__ __forceinline__ void tdecode(uint32_t v, const TGeo& g, int& i, int& j, int& k) {
    uint32_t ij = v / (uint32_t)g.nk;
    k = (int)(v - ij * (uint32_t)g.nk);
    i = (int)(ij / (uint32_t)g.nj);
    j = (int)(ij - (uint32_t)i * (uint32_t)g.nj);
}

__device__ __forceinline__ bool tnbr(const TGeo& g, int i, int j, int k, uint32_t v, int f,
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

// One row of the transport matrix (src/Simulation.cpp:1132-1231).
struct AdrRow {
    double diag;
    double off[6];
    double cs[6], cn[6];   // explicit (1 - theta) parts; zero for Rock links
    uint32_t nb[6];
    int mask;              // bit f: neighbour f linked
    int explicit_mask;     // bit f: neighbour f contributes to the RHS
    uint8_t type;
};

// The row of one voxel from its label, its neighbours' labels (has[f]: the
// neighbour exists) and, for a fluid voxel, its face fluxes, which getq
// loads (qp the +faces, qm the -faces) only for such a voxel.  R.nb is the
// caller's.
template <typename QF>
__device__ __forceinline__ void adr_row_core(uint8_t tv, const uint8_t (&nt)[6],
                                             const bool (&has)[6], QF getq, const AdrConst& a,
                                             AdrRow& R) {
    R.type = tv;
    R.mask = 0;
    R.explicit_mask = 0;
    if (tv == PREC) {          // 1. Precipitate: identity, retain the old value
        R.diag = 1.0;
        return;
    }
    R.diag = a.inv_dt;
    if (tv == ROCK) {          // 2. Rock: 0-D Warren-Root bin
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            if (has[f] && nt[f] == PORE) {
                R.diag += a.alpha;
                R.off[f] = -a.alpha;
                R.mask |= 1 << f;
            }
        }
        return;
    }
    // 3. Pore / Air / Sulphide: active fluid network.  Outward face flux:
    // +face q[v], -face -q[v - s] (faces +i, -i, +j, -j, +k, -k)
    double qp[3], qm[3];
    getq(qp, qm);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (!has[f]) continue;
        const double q_face = (f & 1) ? -qm[f >> 1] : qp[f >> 1];
        const uint8_t neigh_type = nt[f];
        R.mask |= 1 << f;
        if (neigh_type == ROCK) {
            R.diag += a.alpha;
            R.off[f] = -a.alpha;
            continue;
        }
        double actual_q_face = q_face;
        if (neigh_type == PREC || neigh_type == SUL) actual_q_face = 0.0;
        // D_eff = D Dpf next to a Pore, else D; diffusion_coeff = -D_eff / dx^2
        double diffusion_coeff = (neigh_type == PORE) ? a.diff_pore : a.diff_other;
        double coeff_neighbor = diffusion_coeff;
        double coeff_self = -diffusion_coeff;
        double adv_coeff = a.dx_pow2 ? actual_q_face * a.inv_dx : actual_q_face / a.dx;
        if (adv_coeff > 0) coeff_self += adv_coeff;
        else coeff_neighbor -= adv_coeff;
        R.off[f] = a.theta * coeff_neighbor;
        R.diag += a.theta * coeff_self;
        R.cs[f] = coeff_self;
        R.cn[f] = coeff_neighbor;
        R.explicit_mask |= 1 << f;
    }
}

__device__ __forceinline__ void adr_row(const uint8_t* __restrict__ img, const double* __restrict__ qx,
                                        const double* __restrict__ qy, const double* __restrict__ qz,
                                        const TGeo& g, int i, int j, int k, uint32_t v,
                                        const AdrConst& a, AdrRow& R) {
    const uint8_t tv = img[v];
    if (tv == PREC) {
        R.type = tv;
        R.mask = 0;
        R.explicit_mask = 0;
        R.diag = 1.0;
        return;
    }
    // The neighbours' labels (and, for a fluid voxel, the face fluxes) are
    // all loaded before any decision on them, so the loads overlap; a missing
    // neighbour points at v itself and is not used.
    bool has[6];
    uint32_t nb[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        has[f] = tnbr(g, i, j, k, v, f, nb[f]);
        if (!has[f]) nb[f] = v;
    }
    uint8_t nt[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) nt[f] = img[nb[f]];
    adr_row_core(tv, nt, has,
                 [&](double (&qp)[3], double (&qm)[3]) {
                     qp[0] = qx[v]; qp[1] = qy[v]; qp[2] = qz[v];
                     qm[0] = qx[nb[1]]; qm[1] = qy[nb[3]]; qm[2] = qz[nb[5]];
                 },
                 a, R);
#pragma unroll
    for (int f = 0; f < 6; ++f)
        if (R.mask & (1 << f)) R.nb[f] = nb[f];
}

// The explicit (1 - theta) part of the right-hand sides of both systems
// (src/Simulation.cpp:1221-1222): the neighbours' old values are loaded
// together before the sum, which runs over the faces in order.  At theta = 1
// (fully implicit, the default) every term is 0 x a finite value and leaves
// the right-hand side as it is, so the callers skip it and its six
// neighbour loads per system.
__device__ __forceinline__ void explicit_rhs(const AdrRow& R, uint32_t v,
                                             const double* __restrict__ x0,
                                             const double* __restrict__ x1, double xa, double xb,
                                             const AdrConst& a, double& ba, double& bbv) {
    double n0[6], n1[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const uint32_t n = (R.explicit_mask & (1 << f)) ? R.nb[f] : v;
        n0[f] = x0[n];
        n1[f] = x1[n];
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (!(R.explicit_mask & (1 << f))) continue;
        double e0 = -(R.cs[f] * xa + R.cn[f] * n0[f]);
        double e1 = -(R.cs[f] * xb + R.cn[f] * n1[f]);
        ba += (1.0 - a.theta) * e0;
        bbv += (1.0 - a.theta) * e1;
    }
}

// One row of the implicit saturation matrix and its right-hand side
// (src/Simulation.cpp:1870-1955): inactive voxels keep their value; a Pore
// row is phi/dt plus the upwinded mobility advection, the outward face flux
// being the mean of the two cells' flux and the upwind divisor the
// assembly-time saturation sf (floored at 1e-6); rhs phi/dt sf - evap sink.
__device__ __forceinline__ void sat_row(const uint8_t* __restrict__ img, const double* __restrict__ qx,
                                        const double* __restrict__ qy, const double* __restrict__ qz,
                                        const double* __restrict__ sf, const TGeo& g, int i, int j,
                                        int k, uint32_t v, const AdrConst& a, AdrRow& R,
                                        double& rhs) {
    const uint8_t tv = img[v];
    R.type = tv;
    R.mask = 0;
    R.explicit_mask = 0;
    if (tv == AIR || tv == ROCK || tv == SUL || tv == PREC) {
        R.diag = 1.0;
        rhs = sf[v];
        return;
    }
    int air_faces = 0;
    for (int f = 0; f < 6; ++f) {
        uint32_t n;
        if (tnbr(g, i, j, k, v, f, n) && img[n] == AIR) ++air_faces;
    }
    const double evap_sink = (air_faces * a.evap) / a.dx;
    double coeff_self = a.phi * a.inv_dt;
    rhs = (a.phi * a.inv_dt * sf[v]) - evap_sink;
    for (int f = 0; f < 6; ++f) {      // +i, -i, +j, -j, +k, -k
        uint32_t n;
        if (!tnbr(g, i, j, k, v, f, n)) continue;
        const double* q = f < 2 ? qx : (f < 4 ? qy : qz);
        const double q_out = (f % 2 == 0) ? (q[v] + q[n]) / 2.0 : -(q[v] + q[n]) / 2.0;
        if (q_out > 0) {
            const double v_eff = q_out / fmax(sf[v], 1e-6);
            coeff_self += v_eff / a.dx;
        } else {
            const double v_eff = q_out / fmax(sf[n], 1e-6);
            R.off[f] = v_eff / a.dx;
            R.nb[f] = n;
            R.mask |= 1 << f;
        }
    }
    R.diag = coeff_self;
}

// Right-hand sides and the Jacobi diagonal; initial residual r = b - A x,
// r0 = r, p = v = 0.  Partials: |b|^2 and |r|^2 of both systems.
template <int KIND>
__global__ void k_adr_start(const uint8_t* __restrict__ img, const double* __restrict__ qx,
                            const double* __restrict__ qy, const double* __restrict__ qz,
                            const double* __restrict__ sf,
                            const double* __restrict__ x0, const double* __restrict__ x1,
                            double* __restrict__ b0, double* __restrict__ b1,
                            double* __restrict__ invd, double* __restrict__ r_0,
                            double* __restrict__ r_1, double* __restrict__ rh0,
                            double* __restrict__ rh1, double* __restrict__ p0,
                            double* __restrict__ p1, double* __restrict__ vv0,
                            double* __restrict__ vv1, TGeo g, uint32_t N, AdrConst a,
                            int build_rhs, double* __restrict__ partials) {
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    double bb0 = 0.0, bb1 = 0.0, rr0 = 0.0, rr1 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int i, j, k;
        tdecode(v, g, i, j, k);
        AdrRow R;
        double srhs = 0.0;
        if (KIND == 1) sat_row(img, qx, qy, qz, sf, g, i, j, k, v, a, R, srhs);
        else adr_row(img, qx, qy, qz, g, i, j, k, v, a, R);
        const double xa = x0[v], xb = x1[v];
        double ba, bbv;
        if (build_rhs && KIND == 1) {
            ba = srhs;      // system 1 idles with b = 0
            bbv = 0.0;
            b0[v] = ba;
            b1[v] = bbv;
            invd[v] = 1.0 / R.diag;
        } else if (build_rhs) {
            if (R.type == PREC) {
                ba = xa;
                bbv = xb;
            } else {
                ba = xa * a.inv_dt;
                bbv = xb * a.inv_dt;
                if (a.theta != 1.0) explicit_rhs(R, v, x0, x1, xa, xb, a, ba, bbv);
            }
            b0[v] = ba;
            b1[v] = bbv;
            invd[v] = 1.0 / R.diag;
        } else {
            ba = b0[v];
            bbv = b1[v];
        }
        double ya = R.diag * xa, yb = R.diag * xb;
        for (int f = 0; f < 6; ++f) {
            if (!(R.mask & (1 << f))) continue;
            ya += R.off[f] * x0[R.nb[f]];
            yb += R.off[f] * x1[R.nb[f]];
        }
        const double ra = ba - ya, rb = bbv - yb;
        r_0[v] = ra;
        r_1[v] = rb;
        rh0[v] = ra;
        rh1[v] = rb;
        p0[v] = 0.0;
        p1[v] = 0.0;
        vv0[v] = 0.0;
        vv1[v] = 0.0;
        bb0 += ba * ba;
        bb1 += bbv * bbv;
        rr0 += ra * ra;
        rr1 += rb * rb;
    }
    double q[4] = {bb0, bb1, rr0, rr1};
    block_reduce<4, RED_SUM, RED_SUM, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[blockIdx.x] = q[0];
        partials[nb + blockIdx.x] = q[1];
        partials[2 * nb + blockIdx.x] = q[2];
        partials[3 * nb + blockIdx.x] = q[3];
    }
}

__device__ __forceinline__ void sum_partials(const double* partials, int n, int nv, double* out,
                                             double* lds) {
    // deterministic: each thread sums a strided subset, then a fixed tree
    for (int q = 0; q < nv; ++q) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[q * n + i];
        double a[1] = {acc};
        block_reduce<1, RED_SUM>(a, lds);
        if (threadIdx.x == 0) out[q] = a[0];
        __syncthreads();
    }
}

__global__ void k_bicg_init(const double* __restrict__ partials, int n, BicgScalars* bs,
                            double rtol, int maxit, int reset_iter) {
    __shared__ double lds[1024 / WAVE];
    __shared__ double tot[4];
    sum_partials(partials, n, 4, tot, lds);
    if (threadIdx.x == 0) {
        bs->maxit = maxit;
        int all = 1;
        for (int f = 0; f < 2; ++f) {
            bs->bnorm2[f] = tot[f];
            bs->tol2[f] = (rtol * rtol) * tot[f];
            bs->rr[f] = tot[2 + f];
            bs->rho[f] = 1.0;
            bs->alpha[f] = 1.0;
            bs->omega[f] = 1.0;
            bs->beta[f] = 0.0;
            bs->r0r[f] = tot[2 + f];  // r0 = r
            bs->s_conv[f] = 0;
            if (reset_iter) bs->iter[f] = 0;
            const int ok = tot[2 + f] <= bs->tol2[f];
            bs->done[f] = ok;
            bs->converged[f] = ok;
            all &= ok;
        }
        bs->all_done = all;
    }
}

// p = r + beta (p - omega v); with PRE (Jacobi outside the operator) also
// ph = p / diag.  beta from the last rho update.
template <bool PRE>
__global__ void k_bicg_p(double* __restrict__ p0, double* __restrict__ p1,
                         const double* __restrict__ r0_, const double* __restrict__ r1_,
                         const double* __restrict__ v0, const double* __restrict__ v1,
                         double* __restrict__ ph0, double* __restrict__ ph1,
                         const double* __restrict__ invd, const BicgScalars* __restrict__ bs,
                         uint32_t N) {
    if (bs->all_done) return;
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const double be0 = bs->beta[0], om0 = bs->omega[0], be1 = bs->beta[1], om1 = bs->omega[1];
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const double id = PRE ? invd[v] : 1.0;
        if (a0) {
            double pv = r0_[v] + be0 * (p0[v] - om0 * v0[v]);
            p0[v] = pv;
            if (PRE) ph0[v] = pv * id;
        }
        if (a1) {
            double pv = r1_[v] + be1 * (p1[v] - om1 * v1[v]);
            p1[v] = pv;
            if (PRE) ph1[v] = pv * id;
        }
    }
}

// y = A x for both systems; partials of rh.y (MODE 0) or (y.y, y.s) (MODE 1).
template <int MODE, int KIND>
__global__ void k_adr_matvec(const uint8_t* __restrict__ img, const double* __restrict__ qx,
                             const double* __restrict__ qy, const double* __restrict__ qz,
                             const double* __restrict__ sf,
                             const double* __restrict__ x0, const double* __restrict__ x1,
                             double* __restrict__ y0, double* __restrict__ y1,
                             const double* __restrict__ w0, const double* __restrict__ w1,
                             TGeo g, uint32_t N, AdrConst a, const BicgScalars* __restrict__ bs,
                             double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int i, j, k;
        tdecode(v, g, i, j, k);
        AdrRow R;
        if (KIND == 1) {
            double unused;
            sat_row(img, qx, qy, qz, sf, g, i, j, k, v, a, R, unused);
        } else {
            adr_row(img, qx, qy, qz, g, i, j, k, v, a, R);
        }
        double ya = R.diag * x0[v], yb = R.diag * x1[v];
        for (int f = 0; f < 6; ++f) {
            if (!(R.mask & (1 << f))) continue;
            ya += R.off[f] * x0[R.nb[f]];
            yb += R.off[f] * x1[R.nb[f]];
        }
        if (a0) y0[v] = ya;
        if (a1) y1[v] = yb;
        if (MODE == 0) {
            s0 += w0[v] * ya;
            s1 += w1[v] * yb;
        } else {
            s0 += ya * ya;
            s1 += yb * yb;
            s2 += ya * w0[v];
            s3 += yb * w1[v];
        }
    }
    double q[4] = {s0, s1, s2, s3};
    block_reduce<4, RED_SUM, RED_SUM, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[blockIdx.x] = q[0];
        partials[nb + blockIdx.x] = q[1];
        partials[2 * nb + blockIdx.x] = q[2];
        partials[3 * nb + blockIdx.x] = q[3];
    }
}

// ---------------------------------------------------------------------------
// Stored rows.  The matrix is fixed for a whole solve, so its rows are built
// once (the row builder above, scalar) into 7 arrays of N doubles: the
// diagonal and the six face coefficients in face order (+i, -i, +j, -j, +k,
// -k), 0 where a face is not linked.  The matvec then streams them: a
// vectorised i-march like the pressure stencil (2 doubles of one k-row per
// lane, x[i-1], x[i], x[i+1] in registers, k neighbours by shuffles),
// 7 x 8 B of rows + 2 x (x, y, w) = 104 B per voxel, against ~73 B and a
// scalar, branchy row rebuild per matvec.  Terms are added in the row
// builder's order, so y is the same as the matrix-free matvec's.
// ---------------------------------------------------------------------------
// Also builds the right-hand sides and the Jacobi inverse diagonal from the
// current x0, x1, as k_adr_start does with build_rhs.
template <int KIND>
__global__ __launch_bounds__(EW_BLOCK) void k_adr_rows(
    const uint8_t* __restrict__ img, const double* __restrict__ qx,
    const double* __restrict__ qy, const double* __restrict__ qz,
    const double* __restrict__ sf, TGeo g, uint32_t N, AdrConst a,
    double* __restrict__ rows, const double* __restrict__ x0, const double* __restrict__ x1,
    double* __restrict__ b0, double* __restrict__ b1, double* __restrict__ invd) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int i, j, k;
        tdecode(v, g, i, j, k);
        AdrRow R;
        double srhs = 0.0;
        if (KIND == 1) {
            sat_row(img, qx, qy, qz, sf, g, i, j, k, v, a, R, srhs);
        } else {
            adr_row(img, qx, qy, qz, g, i, j, k, v, a, R);
        }
        const double xa = x0[v], xb = x1[v];
        double ba, bbv;
        if (KIND == 1) {
            ba = srhs;      // system 1 idles with b = 0
            bbv = 0.0;
        } else if (R.type == PREC) {
            ba = xa;
            bbv = xb;
        } else {
            ba = xa * a.inv_dt;
            bbv = xb * a.inv_dt;
            if (a.theta != 1.0) explicit_rhs(R, v, x0, x1, xa, xb, a, ba, bbv);
        }
        b0[v] = ba;
        b1[v] = bbv;
        invd[v] = 1.0 / R.diag;
        rows[v] = R.diag;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            rows[(size_t)(f + 1) * N + v] = (R.mask & (1 << f)) ? R.off[f] : 0.0;
    }
}

constexpr int ADR_TJ = 4, ADR_ICH = 8;

static dim3 adr_vec_grid(const ls_ctx* c);

// The vectorised Schur passes' blocks: TSV_TR waves, each over TSV_RPW rows.
constexpr int TSV_TR = 4;     // waves (rows in flight) per block
constexpr int TSV_RPW = 4;    // rows per wave

// Stored rows for nk even when the row kernels' block partials fit; the
// red-black Schur form on them unless switched off.
// With mixed: the Schur form's vectorised passes (nk % 4 == 0, the row count
// within the grid's y limit) and, where nk % 8 == 0 and adr_mixed is set,
// FP32 BiCGStab cycles inside FP64 refinement.
void transport_form(const ls_ctx* c, bool* stored, bool* schur, bool* mixed) {
    const dim3 vg = adr_vec_grid(c);
    const size_t nbm = (size_t)vg.x * vg.y * vg.z;
    *stored = c->nk % 2 == 0 && 4 * nbm <= 8 * (size_t)MAX_PARTIALS;
    *schur = *stored && c->P.adr_schur != 0;
    if (mixed) {
        const size_t gx = ((size_t)c->nk / 4 + WAVE - 1) / WAVE;
        const size_t gy = ((size_t)c->ni * c->nj + TSV_TR * TSV_RPW - 1) / (TSV_TR * TSV_RPW);
        const bool vec = *schur && c->nk % 4 == 0 && 4 * gx * gy <= 8 * (size_t)MAX_PARTIALS;
        *mixed = vec && c->nk % 8 == 0 && c->P.adr_mixed != 0;
    }
}

static dim3 adr_vec_grid(const ls_ctx* c) {
    return dim3((c->nk / 2 + WAVE - 1) / WAVE, (c->nj + ADR_TJ - 1) / ADR_TJ,
                (c->ni + ADR_ICH - 1) / ADR_ICH);
}

// y = A x from the stored rows for the systems in SYS (bit 0: system 0,
// bit 1: system 1; the host drops a system once it has converged); partials
// of rh.y (MODE 0) or (y.y, y.s) (MODE 1), as k_adr_matvec.  MODE 2 is the
// solver start: y = b - A x (w = b) into r, and also rh = r, p = v = 0 and
// the partials (b.b, r.r) of k_adr_start.
template <int MODE, int SYS>
__global__ __launch_bounds__(WAVE * ADR_TJ) void k_adr_matvec_rows(
    const double* __restrict__ rows, const double* __restrict__ x0,
    const double* __restrict__ x1, double* __restrict__ y0, double* __restrict__ y1,
    const double* __restrict__ w0, const double* __restrict__ w1, TGeo g, uint32_t N,
    const BicgScalars* __restrict__ bs, double* __restrict__ partials,
    double* __restrict__ rh0, double* __restrict__ rh1, double* __restrict__ p0,
    double* __restrict__ p1, double* __restrict__ v0, double* __restrict__ v1) {
    constexpr bool S0 = (SYS & 1) != 0, S1 = (SYS & 2) != 0;
    if (MODE != 2 && bs->all_done) return;
    typedef double D2 __attribute__((ext_vector_type(2)));
    __shared__ double lds[4 * ADR_TJ];
    const bool a0 = MODE == 2 || !bs->done[0], a1 = MODE == 2 || !bs->done[1];
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nb = nbx * nby * gridDim.z;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int t = xcd_remap(hw, nb);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int lane = threadIdx.x;
    const int k0 = (bx * WAVE + lane) * 2;
    const int j = by * ADR_TJ + threadIdx.y;
    const int i0 = bz * ADR_ICH, i1 = min(i0 + ADR_ICH, g.ni);
    const bool ok = j < g.nj && k0 < g.nk;
    const bool hj = j > 0, pj = j + 1 < g.nj;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const D2 zero = D2(0.0);
    uint32_t v = ok ? (uint32_t)i0 * g.si + (uint32_t)j * g.sj + (uint32_t)k0 : 0u;
    D2 am = zero, bm = zero, ac = zero, bc = zero;
    if (ok && i0 > 0) {
        if (S0) am = *(const D2*)(x0 + v - g.si);
        if (S1) bm = *(const D2*)(x1 + v - g.si);
    }
    if (ok) {
        if (S0) ac = *(const D2*)(x0 + v);
        if (S1) bc = *(const D2*)(x1 + v);
    }
    for (int i = i0; i < i1; ++i, v += g.si) {
        D2 ap = zero, bp = zero, ajm = zero, bjm = zero, ajp = zero, bjp = zero;
        D2 r[7];
#pragma unroll
        for (int f = 0; f < 7; ++f) r[f] = zero;
        double la = 0.0, lb = 0.0, ra = 0.0, rb = 0.0;
        if (ok) {
            if (i + 1 < g.ni) {
                if (S0) ap = *(const D2*)(x0 + v + g.si);
                if (S1) bp = *(const D2*)(x1 + v + g.si);
            }
            if (hj) {
                if (S0) ajm = *(const D2*)(x0 + v - g.sj);
                if (S1) bjm = *(const D2*)(x1 + v - g.sj);
            }
            if (pj) {
                if (S0) ajp = *(const D2*)(x0 + v + g.sj);
                if (S1) bjp = *(const D2*)(x1 + v + g.sj);
            }
#pragma unroll
            for (int f = 0; f < 7; ++f) r[f] = *(const D2*)(rows + (size_t)f * N + v);
            if (lane == 0 && k0 > 0) {
                if (S0) la = x0[v - 1];
                if (S1) lb = x1[v - 1];
            }
            if (lane == WAVE - 1 && k0 + 2 < g.nk) {
                if (S0) ra = x0[v + 2];
                if (S1) rb = x1[v + 2];
            }
        }
        double sla = 0.0, slb = 0.0, sra = 0.0, srb = 0.0;
        if (S0) {
            sla = __shfl_up(ac[1], 1, WAVE);
            sra = __shfl_down(ac[0], 1, WAVE);
        }
        if (S1) {
            slb = __shfl_up(bc[1], 1, WAVE);
            srb = __shfl_down(bc[0], 1, WAVE);
        }
        if (lane != 0) { la = sla; lb = slb; }
        if (lane != WAVE - 1) { ra = sra; rb = srb; }
        D2 ya = zero, yb = zero;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            if (S0) {
                const double akp = w == 0 ? ac[1] : ra, akm = w == 0 ? la : ac[0];
                double u = r[0][w] * ac[w];
                u += r[1][w] * ap[w];
                u += r[2][w] * am[w];
                u += r[3][w] * ajp[w];
                u += r[4][w] * ajm[w];
                u += r[5][w] * akp;
                u += r[6][w] * akm;
                ya[w] = u;
            }
            if (S1) {
                const double bkp = w == 0 ? bc[1] : rb, bkm = w == 0 ? lb : bc[0];
                double z = r[0][w] * bc[w];
                z += r[1][w] * bp[w];
                z += r[2][w] * bm[w];
                z += r[3][w] * bjp[w];
                z += r[4][w] * bjm[w];
                z += r[5][w] * bkp;
                z += r[6][w] * bkm;
                yb[w] = z;
            }
        }
        if (ok) {
            D2 wa = zero, wb = zero;
            if (S0) wa = *(const D2*)(w0 + v);
            if (S1) wb = *(const D2*)(w1 + v);
            if (MODE == 2) {
                // r = b - A x, rh = r, p = v = 0
                const D2 ra2 = wa - ya, rb2 = wb - yb;
                if (S0) {
                    *(D2*)(y0 + v) = ra2;
                    *(D2*)(rh0 + v) = ra2;
                    *(D2*)(p0 + v) = zero;
                    *(D2*)(v0 + v) = zero;
                }
                if (S1) {
                    *(D2*)(y1 + v) = rb2;
                    *(D2*)(rh1 + v) = rb2;
                    *(D2*)(p1 + v) = zero;
                    *(D2*)(v1 + v) = zero;
                }
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    s0 += wa[w] * wa[w];
                    s1 += wb[w] * wb[w];
                    s2 += ra2[w] * ra2[w];
                    s3 += rb2[w] * rb2[w];
                }
            } else {
                if (S0 && a0) *(D2*)(y0 + v) = ya;
                if (S1 && a1) *(D2*)(y1 + v) = yb;
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    if (MODE == 0) {
                        s0 += wa[w] * ya[w];
                        s1 += wb[w] * yb[w];
                    } else {
                        s0 += ya[w] * ya[w];
                        s1 += yb[w] * yb[w];
                        s2 += ya[w] * wa[w];
                        s3 += yb[w] * wb[w];
                    }
                }
            }
        }
        am = ac;
        bm = bc;
        ac = ap;
        bc = bp;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    s3 = wave_sum(s3);
    if (lane == 0) {
        lds[4 * threadIdx.y + 0] = s0;
        lds[4 * threadIdx.y + 1] = s1;
        lds[4 * threadIdx.y + 2] = s2;
        lds[4 * threadIdx.y + 3] = s3;
    }
    __syncthreads();
    if (flat_tid() == 0) {
        double q[4] = {0.0, 0.0, 0.0, 0.0};
        for (int y = 0; y < ADR_TJ; ++y)
            for (int u = 0; u < 4; ++u) q[u] += lds[4 * y + u];
        for (int u = 0; u < 4; ++u) partials[(size_t)u * nb + t] = q[u];
    }
}

// ---------------------------------------------------------------------------
// Red-black Schur complement of the stored rows.  The 7-point graph is
// bipartite (colour (i + j + k) & 1, black 0), so with the rows split by
// colour, A = [D_b A_br; A_rb D_r] with diagonal D_b, D_r, and
//   S x_b = b_b - A_br D_r^-1 b_r,   S = D_b - A_br D_r^-1 A_rb,
//   x_r   = D_r^-1 (b_r - A_rb x_b).
// BiCGStab with Jacobi on D_b runs on the black unknowns: the spectral
// radius of the Jacobi iteration matrix squares, so about half the
// iterations, on half-length vectors.  (diag(S) = D_b - sum a_bf a_fb / d_f
// lies within 0.83..1 of D_b on the transport rows, and BiCGStab took the
// same iteration counts with either on the oracle's systems of the
// fractured workload at 48^3.)  With the red values implied by the black
// ones the full residual is (b_b - A_bb x_b - A_br x_r, 0) = (r_S, 0), so the
// solver stops on ||r_S|| <= rtol ||b|| with the full system's b.  b_S is
// never formed: a start pass computes u = D_r^-1 (b_r - A_rb x) at the red
// voxels and r_S = b_b - D_b x - A_br u at the black ones; the final one
// also writes u and x into the fields.
//
// The Jacobi preconditioner is folded into the operator: right-preconditioned
// BiCGStab on S is BiCGStab on S' = S D_b^-1 = I - A_br R' with
// R' = D_r^-1 A_rb D_b^-1, in the unknown y = D_b x_b (the same iterates in
// exact arithmetic).  The first start pass rescales the red rows in place to
// R' (and their diagonal slot to 1 / d_r), so an apply reads six
// coefficients per colour and no diagonal, and the solver keeps no
// preconditioned copies of its direction vectors.
//
// Layout (nk even): voxel v of colour c is stored at v >> 1 of its colour's
// half-length arrays; the pair (2m, 2m + 1) holds one voxel of each colour.
// Rows: 7 coefficients (diagonal, then faces +i, -i, +j, -j, +k, -k) x 2
// colours x N/2 doubles.
// ---------------------------------------------------------------------------
struct SPair {
    uint32_t vb, vr;   // full indices of the pair's black and red voxel
    int i, j, kb, kr;
};

__device__ __forceinline__ SPair spair(uint32_t m, const TGeo& g) {
    SPair s;
    const uint32_t v0 = 2u * m;
    const uint32_t row = v0 / (uint32_t)g.nk;
    s.i = (int)(row / (uint32_t)g.nj);
    s.j = (int)(row - (uint32_t)s.i * (uint32_t)g.nj);
    const int k0 = (int)(v0 - row * (uint32_t)g.nk);
    const int par = (s.i + s.j) & 1;          // k0 is even
    s.vb = v0 + (uint32_t)par;
    s.vr = v0 + 1u - (uint32_t)par;
    s.kb = k0 + par;
    s.kr = k0 + 1 - par;
    return s;
}

// Order of the extrapolated transport start (1 linear, 2 quadratic, 3 cubic;
// one history field of N/2 doubles each).  On the 512^3 fractured workload
// the quadratic start took FP64 BiCGStab from 5 to 4 iterations per step
// from step 13 on (steps 5-24: 69.3 -> 65.7 ms per step), and the cubic one
// the mixed-precision form's FP32 cycle from 5 to 4 iterations from step 11
// on (steps 5-24: 50.0 -> 47.2 ms per step).
#ifndef LS_ADR_EXTRAP_ORDER
#define LS_ADR_EXTRAP_ORDER 3
#endif

// The 14 colour-split row arrays are TROW_PAD doubles apart beyond their
// N/2 length, so that arrays streamed together start at different offsets
// modulo the HBM channel interleave (at 512^3, N/2 doubles are exactly
// 2^29 B, and unpadded every row array of one voxel would map to the same
// channel).
constexpr uint32_t TROW_PAD = 544;     // 4352 B
__device__ __forceinline__ size_t rstride(uint32_t Nh) { return (size_t)Nh + TROW_PAD; }
// The FP32 copies of the twelve off-diagonal arrays (mixed-precision
// transport), the same 4352 B apart.
constexpr uint32_t TROW_PAD32 = 1088;
__host__ __device__ __forceinline__ size_t rstride32(uint32_t Nh) { return (size_t)Nh + TROW_PAD32; }

// Doubles of the stored rows: 7 N for the full-system layout, 14 (N/2 +
// TROW_PAD) for the colour-split one.
// Floats of the FP32 off-diagonal copies of the mixed-precision form.
size_t transport_rows32_floats(const ls_ctx* c) {
    return 12 * rstride32((uint32_t)(c->N / 2));
}

size_t transport_rows_doubles(const ls_ctx* c) {
    return std::max(7 * c->N, 14 * (c->N / 2 + TROW_PAD));
}

__device__ __forceinline__ double srow(const double* __restrict__ rows, uint32_t Nh, int f,
                                       int colour, uint32_t m) {
    return rows[((size_t)f * 2 + colour) * rstride(Nh) + m];
}

// The stored rows, colour-split, with b (full layout) as k_adr_rows; at the
// black voxels also the solver's start y = D_b x.  one: system 1 is known to
// be 0 with b = 0 (its b and x are not touched).  With a history (the black
// part of system 0's field at the last transport solves, hist0 the newest)
// system 0 starts from its extrapolation in time through the current field
// and ex.n history fields (extrap_start), and the history shifts by one.
// Transport (KIND 0) also zeroes the flux at Rock and Precipitate voxels
// here, as the reference's assembly does (src/Simulation.cpp:1134-1136,
// 1147-1149): no row reads the flux of such a voxel (a Rock neighbour is an
// exchange link, a Precipitate neighbour a flux barrier), so the writes race
// with no read whose value is used.
template <int KIND>
__global__ __launch_bounds__(EW_BLOCK) void k_adr_rows_split(
    const uint8_t* __restrict__ img, double* qx, double* qy, double* qz,
    const double* __restrict__ sf, TGeo g, uint32_t N, AdrConst a,
    double* __restrict__ rows, const double* __restrict__ x0, const double* __restrict__ x1,
    double* __restrict__ b0, double* __restrict__ b1, int one, double* __restrict__ xs0,
    double* __restrict__ xs1, const double* hist0,
    const double* hist1, const double* hist2, double* hw, Extrap ex) {
    const uint32_t Nh = N / 2;
    const double* x1r = one ? x0 : x1;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int i, j, k;
        tdecode(v, g, i, j, k);
        AdrRow R;
        double srhs = 0.0;
        if (KIND == 1) {
            sat_row(img, qx, qy, qz, sf, g, i, j, k, v, a, R, srhs);
        } else {
            adr_row(img, qx, qy, qz, g, i, j, k, v, a, R);
            if (R.type == PREC || R.type == ROCK) {
                qx[v] = 0.0;
                qy[v] = 0.0;
                qz[v] = 0.0;
            }
        }
        const double xa = x0[v], xb = x1r[v];
        double ba, bbv;
        if (KIND == 1) {
            ba = srhs;
            bbv = 0.0;
        } else if (R.type == PREC) {
            ba = xa;
            bbv = xb;
        } else {
            ba = xa * a.inv_dt;
            bbv = xb * a.inv_dt;
            if (a.theta != 1.0) explicit_rhs(R, v, x0, x1r, xa, xb, a, ba, bbv);
        }
        b0[v] = ba;
        if (!one) b1[v] = bbv;
        const int colour = (i + j + k) & 1;
        const uint32_t m = v >> 1;
        rows[((size_t)0 * 2 + colour) * rstride(Nh) + m] = R.diag;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            rows[((size_t)(f + 1) * 2 + colour) * rstride(Nh) + m] =
                (R.mask & (1 << f)) ? R.off[f] : 0.0;
        if (colour == 0) {
            const double xstart = extrap_start(xa, m, hist0, hist1, hist2, hw, ex);
            xs0[m] = R.diag * xstart;
            if (!one) xs1[m] = R.diag * xb;
        }
    }
}

// The red pass for the systems in SYS, on the scaled unknowns y = D_b x
// (the red rows hold R' = D_r^-1 A_rb D_b^-1 and 1 / d_r once the first start
// pass has rescaled them).  MODE 0 (apply): w = R' p.  MODE 1 (first start):
// rescales the red rows of the unscaled assembly (a_f -> a_f / (d_r d_f)
// with d_f the black neighbour's diagonal, d_r -> 1 / d_r), then as MODE 2.
// MODE 2 (start): w = D_r^-1 b_r - R' y = D_r^-1 (b_r - A_rb x), the red
// unknowns implied by the black ones, also written into the full field xf
// at the red voxels (every start after the first).
template <int SYS, int MODE>
__global__ __launch_bounds__(EW_BLOCK) void k_tschur_red(
    double* __restrict__ rows, TGeo g, uint32_t Nh, const double* __restrict__ p0,
    const double* __restrict__ p1, double* __restrict__ w0, double* __restrict__ w1,
    const BicgScalars* __restrict__ bs, const double* __restrict__ bf0,
    const double* __restrict__ bf1, double* __restrict__ xf0, double* __restrict__ xf1) {
    if (MODE == 0 && bs->all_done) return;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < Nh; m += gridDim.x * blockDim.x) {
        const SPair P = spair(m, g);
        double inv;
        if (MODE == 1) {
            inv = 1.0 / srow(rows, Nh, 0, 1, m);
            rows[((size_t)0 * 2 + 1) * rstride(Nh) + m] = inv;
        } else {
            inv = srow(rows, Nh, 0, 1, m);
        }
        double a0 = 0.0, a1 = 0.0;
        for (int f = 0; f < 6; ++f) {
            uint32_t n;
            if (!tnbr(g, P.i, P.j, P.kr, P.vr, f, n)) continue;
            double o = srow(rows, Nh, f + 1, 1, m);
            if (MODE == 1) {
                o = (o * inv) / srow(rows, Nh, 0, 0, n >> 1);
                rows[((size_t)(f + 1) * 2 + 1) * rstride(Nh) + m] = o;
            }
            if (SYS & 1) a0 += o * p0[n >> 1];
            if (SYS & 2) a1 += o * p1[n >> 1];
        }
        if (MODE == 0) {
            if (SYS & 1) w0[m] = a0;
            if (SYS & 2) w1[m] = a1;
        } else {
            if (SYS & 1) {
                const double u = bf0[P.vr] * inv - a0;
                w0[m] = u;
                if (MODE == 2) xf0[P.vr] = u;
            }
            if (SYS & 2) {
                const double u = bf1[P.vr] * inv - a1;
                w1[m] = u;
                if (MODE == 2) xf1[P.vr] = u;
            }
        }
    }
}

// The black pass, on the scaled unknowns y = D_b x.  MODE 0/1 (apply):
// z = S' y = y - A_br w at the black voxels, with partials rh.z (MODE 0) or
// (z.z, z.s) (MODE 1), w from the red pass in mode 0.  MODE 2 (start): the
// Schur residual r = b_b - y - A_br u = b_b - D_b x - A_br u with u from a
// red start pass (r_S = b_S - S x, b_S = b_b - A_br D_r^-1 b_r), rh = r,
// p = v = 0, partials (b.b over the whole grid, r.r); MODE 3 (final): the
// same, also writing x = y / d_b into the full field xf at the black
// voxels.
template <int MODE, int SYS>
__global__ __launch_bounds__(EW_BLOCK) void k_tschur_black(
    const double* __restrict__ rows, TGeo g, uint32_t Nh, const double* __restrict__ x0,
    const double* __restrict__ x1, const double* __restrict__ wr0, const double* __restrict__ wr1,
    double* __restrict__ y0, double* __restrict__ y1, const double* __restrict__ w0,
    const double* __restrict__ w1, const BicgScalars* __restrict__ bs,
    double* __restrict__ partials, double* __restrict__ rh0, double* __restrict__ rh1,
    double* __restrict__ p0, double* __restrict__ p1, double* __restrict__ v0,
    double* __restrict__ v1, const double* __restrict__ bf0, const double* __restrict__ bf1,
    double* __restrict__ xf0, double* __restrict__ xf1) {
    constexpr bool S0 = (SYS & 1) != 0, S1 = (SYS & 2) != 0;
    constexpr bool START = MODE >= 2;
    if (!START && bs->all_done) return;
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    const bool a0 = START || !bs->done[0], a1 = START || !bs->done[1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < Nh; m += gridDim.x * blockDim.x) {
        const SPair P = spair(m, g);
        const double xa = S0 ? x0[m] : 0.0, xb = S1 ? x1[m] : 0.0;
        double ya = xa, yb = xb;
        for (int f = 0; f < 6; ++f) {
            uint32_t n;
            if (!tnbr(g, P.i, P.j, P.kb, P.vb, f, n)) continue;
            const double o = srow(rows, Nh, f + 1, 0, m);
            if (START) {
                if (S0) ya += o * wr0[n >> 1];
                if (S1) yb += o * wr1[n >> 1];
            } else {
                if (S0) ya -= o * wr0[n >> 1];
                if (S1) yb -= o * wr1[n >> 1];
            }
        }
        if (START) {
            const double ra = S0 ? bf0[P.vb] - ya : 0.0, rb = S1 ? bf1[P.vb] - yb : 0.0;
            const double d = MODE == 3 ? srow(rows, Nh, 0, 0, m) : 1.0;
            if (S0) {
                y0[m] = ra;
                rh0[m] = ra;
                p0[m] = 0.0;
                v0[m] = 0.0;
                if (MODE == 3) xf0[P.vb] = xa / d;
            }
            if (S1) {
                y1[m] = rb;
                rh1[m] = rb;
                p1[m] = 0.0;
                v1[m] = 0.0;
                if (MODE == 3) xf1[P.vb] = xb / d;
            }
            // ||b|| of the full system: the pair (2m, 2m + 1)
            if (S0) {
                const double ba = bf0[2 * m], bb = bf0[2 * m + 1];
                s0 += ba * ba + bb * bb;
            }
            if (S1) {
                const double bc = bf1[2 * m], bd = bf1[2 * m + 1];
                s1 += bc * bc + bd * bd;
            }
            s2 += ra * ra;
            s3 += rb * rb;
        } else {
            if (S0 && a0) y0[m] = ya;
            if (S1 && a1) y1[m] = yb;
            if (MODE == 0) {
                if (S0) s0 += w0[m] * ya;
                if (S1) s1 += w1[m] * yb;
            } else {
                if (S0) {
                    s0 += ya * ya;
                    s2 += ya * w0[m];
                }
                if (S1) {
                    s1 += yb * yb;
                    s3 += yb * w1[m];
                }
            }
        }
    }
    double q[4] = {s0, s1, s2, s3};
    block_reduce<4, RED_SUM, RED_SUM, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[blockIdx.x] = q[0];
        partials[nb + blockIdx.x] = q[1];
        partials[2 * nb + blockIdx.x] = q[2];
        partials[3 * nb + blockIdx.x] = q[3];
    }
}

// ---------------------------------------------------------------------------
// The same two passes, vectorised, for nk % 4 == 0: a lane owns two
// consecutive compacted values (16 B) of one row (i, j), so every row array
// and every i / j neighbour row is read as aligned double2 streams, with no
// index division per value.  In a row of parity par = (i + j) & 1 a voxel
// with compacted index m sits at k = 2m + q (q = par for black, 1 - par for
// red); its k neighbours (the other colour) are at m - 1 + q and m + q,
// which for the lane's pair (m0, m0 + 1) are the pair itself and one value
// of the neighbouring lane (shuffles; the wave's edge lanes load it).  A
// block is 64 lanes x TSV_TR waves; each wave walks TSV_RPW rows.  The sums
// run over the faces in the order of the scalar passes (+i, -i, +j, -j, +k,
// -k), absent neighbours contributing 0 x 0.
// ---------------------------------------------------------------------------

typedef double d2 __attribute__((ext_vector_type(2)));

// Row coefficients are read once per pass: nontemporal loads keep them from
// displacing the reused vectors in L2 and the last-level cache.
__device__ __forceinline__ d2 ld_once(const double* p) {
    return __builtin_nontemporal_load((const d2*)p);
}

struct TVGeo {
    uint32_t nhr;       // compacted values per row (nk / 2, even)
    uint32_t nrows;     // ni * nj
    uint32_t plane;     // compacted values per i-plane (nj * nhr)
    int ni, nj;
};

// the other colour's values at the k neighbours of the lane's pair: (km, kp)
// for each of the two values; q is the own k parity
__device__ __forceinline__ void tsv_kn(const double* __restrict__ src, uint32_t idx, uint32_t m0,
                                       uint32_t nhr, int q, bool ok, d2 c, d2& km, d2& kp) {
    const int lane = threadIdx.x;
    double prev = __shfl_up(c.y, 1, WAVE);
    double next = __shfl_down(c.x, 1, WAVE);
    if (lane == 0) prev = (ok && q == 0 && m0 >= 1) ? src[idx - 1] : 0.0;
    if (lane == WAVE - 1) next = (ok && q == 1 && m0 + 2 < nhr) ? src[idx + 2] : 0.0;
    if (q == 0) {
        km = d2{m0 >= 1 ? prev : 0.0, c.x};
        kp = d2{c.x, c.y};
    } else {
        km = d2{c.x, c.y};
        kp = d2{c.y, m0 + 2 < nhr ? next : 0.0};
    }
}

// The conc system's rows (KIND 0 of k_adr_rows_split) with a voxel pair per
// lane: the pair (2m, 2m + 1) of a row (i, j) holds one black and one red
// voxel (the black one at 2m + par, par = (i + j) & 1), so labels are read
// as 2-byte pairs and fluxes and fields as 16-byte pairs, the k neighbours
// come from the neighbouring lanes, and each colour-split row array is
// written by all lanes of a wave at consecutive m.  Rows are computed by
// adr_row_core, as in k_adr_rows_split: the same bits.  A block is 64 lanes
// x TSV_TR waves, each wave walks TSV_RPW rows (i, j); nk even.
__device__ __forceinline__ uint32_t ld_lab2(const uint8_t* __restrict__ img, uint32_t v) {
    return (uint32_t)(*(const uint16_t*)(img + v));
}

template <bool EXPL>
__global__ __launch_bounds__(WAVE * TSV_TR) void k_adr_rows_pair(
    const uint8_t* __restrict__ img, double* qx, double* qy, double* qz, TVGeo g, uint32_t N,
    AdrConst a, double* __restrict__ rows, const double* __restrict__ x0,
    const double* __restrict__ x1, double* __restrict__ b0, double* __restrict__ b1, int one,
    double* __restrict__ xs0, double* __restrict__ xs1, const double* hist0,
    const double* hist1, const double* hist2, double* hw, Extrap ex) {
    const uint32_t Nh = N / 2;
    const double* x1r = one ? x0 : x1;
    const int lane = threadIdx.x;
    const uint32_t m = blockIdx.x * WAVE + lane;
    const bool ok = m < g.nhr;
    const uint32_t nk = 2 * g.nhr;
    const uint32_t sj = nk, si = nk * (uint32_t)g.nj;
    const d2 zero = d2{0.0, 0.0};
    constexpr bool expl = EXPL;     // theta != 1: the explicit part of the old field
#pragma unroll 1
    for (int q8 = 0; q8 < TSV_RPW; ++q8) {
        const uint32_t row = (row_block() * TSV_RPW + q8) * TSV_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const int par = (i + j) & 1;
        const uint32_t v0 = row * nk + 2 * m;
        const uint32_t mg = row * g.nhr + m;            // the pair's compacted index
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        // every load of the pair first
        const uint32_t lc = ok ? ld_lab2(img, v0) : 0u;
        const uint32_t lip = (ok && hip_) ? ld_lab2(img, v0 + si) : 0u;
        const uint32_t lim = (ok && him) ? ld_lab2(img, v0 - si) : 0u;
        const uint32_t ljp = (ok && hjp) ? ld_lab2(img, v0 + sj) : 0u;
        const uint32_t ljm = (ok && hjm) ? ld_lab2(img, v0 - sj) : 0u;
        const d2 qxp = ok ? *(const d2*)(qx + v0) : zero;
        const d2 qyp = ok ? *(const d2*)(qy + v0) : zero;
        const d2 qzp = ok ? *(const d2*)(qz + v0) : zero;
        const d2 qxm = (ok && him) ? *(const d2*)(qx + v0 - si) : zero;
        const d2 qym = (ok && hjm) ? *(const d2*)(qy + v0 - sj) : zero;
        const d2 xa2 = ok ? *(const d2*)(x0 + v0) : zero;
        const d2 xb2 = ok ? *(const d2*)(x1r + v0) : zero;
        // the k neighbours outside the pair: voxel 2m - 1 (the left lane's
        // second) and 2m + 2 (the right lane's first); the wave's edge lanes
        // load theirs
        uint32_t kl = (uint32_t)__shfl_up((int)lc, 1, WAVE) >> 8;
        uint32_t kr = (uint32_t)__shfl_down((int)lc, 1, WAVE) & 0xffu;
        double qzl = __shfl_up(qzp.y, 1, WAVE);
        if (lane == 0) {
            kl = (ok && m >= 1) ? (uint32_t)img[v0 - 1] : 0u;
            qzl = (ok && m >= 1) ? qz[v0 - 1] : 0.0;
        }
        if (lane == WAVE - 1) kr = (ok && m + 1 < g.nhr) ? (uint32_t)img[v0 + 2] : 0u;
        // the old field at the neighbours (the explicit part, theta != 1)
        d2 nip = zero, nim = zero, njp = zero, njm = zero, mip = zero, mim = zero, mjp = zero,
           mjm = zero;
        double nkl = 0.0, nkr = 0.0, mkl = 0.0, mkr = 0.0;
        if (expl) {
            if (ok && hip_) { nip = *(const d2*)(x0 + v0 + si); mip = *(const d2*)(x1r + v0 + si); }
            if (ok && him) { nim = *(const d2*)(x0 + v0 - si); mim = *(const d2*)(x1r + v0 - si); }
            if (ok && hjp) { njp = *(const d2*)(x0 + v0 + sj); mjp = *(const d2*)(x1r + v0 + sj); }
            if (ok && hjm) { njm = *(const d2*)(x0 + v0 - sj); mjm = *(const d2*)(x1r + v0 - sj); }
            nkl = __shfl_up(xa2.y, 1, WAVE);
            mkl = __shfl_up(xb2.y, 1, WAVE);
            nkr = __shfl_down(xa2.x, 1, WAVE);
            mkr = __shfl_down(xb2.x, 1, WAVE);
            if (lane == 0) {
                nkl = (ok && m >= 1) ? x0[v0 - 1] : 0.0;
                mkl = (ok && m >= 1) ? x1r[v0 - 1] : 0.0;
            }
            if (lane == WAVE - 1) {
                nkr = (ok && m + 1 < g.nhr) ? x0[v0 + 2] : 0.0;
                mkr = (ok && m + 1 < g.nhr) ? x1r[v0 + 2] : 0.0;
            }
        }
        if (!ok) continue;
        double bav[2], bbv[2];
        bool zq[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int k = 2 * (int)m + e;
            const uint8_t tv = (uint8_t)(e ? (lc >> 8) : (lc & 0xffu));
            const bool has[6] = {hip_, him, hjp, hjm, k + 1 < (int)nk, k > 0};
            const uint8_t nt[6] = {(uint8_t)(e ? (lip >> 8) : (lip & 0xffu)),
                                   (uint8_t)(e ? (lim >> 8) : (lim & 0xffu)),
                                   (uint8_t)(e ? (ljp >> 8) : (ljp & 0xffu)),
                                   (uint8_t)(e ? (ljm >> 8) : (ljm & 0xffu)),
                                   (uint8_t)(e ? kr : (lc >> 8)),
                                   (uint8_t)(e ? (lc & 0xffu) : kl)};
            AdrRow R;
            adr_row_core(tv, nt, has,
                         [&](double (&qp)[3], double (&qm)[3]) {
                             qp[0] = e ? qxp.y : qxp.x;
                             qp[1] = e ? qyp.y : qyp.x;
                             qp[2] = e ? qzp.y : qzp.x;
                             qm[0] = e ? qxm.y : qxm.x;
                             qm[1] = e ? qym.y : qym.x;
                             qm[2] = e ? qzp.x : qzl;
                         },
                         a, R);
            zq[e] = R.type == PREC || R.type == ROCK;
            const double xa = e ? xa2.y : xa2.x, xb = e ? xb2.y : xb2.x;
            double ba, bb;
            if (R.type == PREC) {
                ba = xa;
                bb = xb;
            } else {
                ba = xa * a.inv_dt;
                bb = xb * a.inv_dt;
                if (expl) {
                    const double n0[6] = {e ? nip.y : nip.x, e ? nim.y : nim.x,
                                          e ? njp.y : njp.x, e ? njm.y : njm.x,
                                          e ? nkr : xa2.y, e ? xa2.x : nkl};
                    const double n1[6] = {e ? mip.y : mip.x, e ? mim.y : mim.x,
                                          e ? mjp.y : mjp.x, e ? mjm.y : mjm.x,
                                          e ? mkr : xb2.y, e ? xb2.x : mkl};
#pragma unroll
                    for (int f = 0; f < 6; ++f) {
                        if (!(R.explicit_mask & (1 << f))) continue;
                        double ea = -(R.cs[f] * xa + R.cn[f] * n0[f]);
                        double eb = -(R.cs[f] * xb + R.cn[f] * n1[f]);
                        ba += (1.0 - a.theta) * ea;
                        bb += (1.0 - a.theta) * eb;
                    }
                }
            }
            bav[e] = ba;
            bbv[e] = bb;
            const int colour = (par + e) & 1;
            // the rows, b and the start are read by later kernels only:
            // nontemporal stores keep them from evicting the flux and label
            // rows that the neighbouring waves still read (4.24 -> 3.66 ms
            // at 512^3)
            __builtin_nontemporal_store(R.diag, rows + (size_t)colour * rstride(Nh) + mg);
#pragma unroll
            for (int f = 0; f < 6; ++f)
                __builtin_nontemporal_store((R.mask & (1 << f)) ? R.off[f] : 0.0,
                                            rows + ((size_t)(f + 1) * 2 + colour) * rstride(Nh) + mg);
            if (colour == 0) {
                const double xstart = extrap_start(xa, mg, hist0, hist1, hist2, hw, ex);
                xs0[mg] = R.diag * xstart;
                if (!one) xs1[mg] = R.diag * xb;
            }
        }
        __builtin_nontemporal_store(d2{bav[0], bav[1]}, (d2*)(b0 + v0));
        if (!one) __builtin_nontemporal_store(d2{bbv[0], bbv[1]}, (d2*)(b1 + v0));
        // the flux at Rock and Precipitate voxels is zeroed, as in
        // k_adr_rows_split (no row uses it)
        if (zq[0] || zq[1]) {
            *(d2*)(qx + v0) = d2{zq[0] ? 0.0 : qxp.x, zq[1] ? 0.0 : qxp.y};
            *(d2*)(qy + v0) = d2{zq[0] ? 0.0 : qyp.x, zq[1] ? 0.0 : qyp.y};
            *(d2*)(qz + v0) = d2{zq[0] ? 0.0 : qzp.x, zq[1] ? 0.0 : qzp.y};
        }
    }
}

// Red pass on the scaled unknowns, the modes of k_tschur_red: MODE 0
// w = R' p; MODE 1 rescales the red rows to R' and 1 / d_r, then as MODE 2;
// MODE 2 w = D_r^-1 b_r - R' y (the final black pass writes it into the
// field).  The mixed-precision form keeps the FP64 red rows unscaled: MODE 3
// is MODE 1 that writes only the FP32 copies of R', MODE 4 is MODE 2 on the
// unscaled rows, rescaled in registers as MODE 1 does (the same bits).
template <int SYS, int MODE>
__global__ __launch_bounds__(WAVE * TSV_TR) void k_tschur_red_v(
    double* __restrict__ rows, TVGeo g, uint32_t Nh, const double* __restrict__ p0,
    const double* __restrict__ p1, double* __restrict__ w0, double* __restrict__ w1,
    const BicgScalars* __restrict__ bs, const double* __restrict__ bf0,
    const double* __restrict__ bf1, double* __restrict__ xf0, double* __restrict__ xf1,
    float* __restrict__ rows32) {
    if (MODE == 0 && bs->all_done) return;
    const uint32_t m0 = 2u * (blockIdx.x * WAVE + threadIdx.x);
    const bool ok = m0 < g.nhr;
    const d2 zero = d2{0.0, 0.0};
#pragma unroll
    for (int q8 = 0; q8 < TSV_RPW; ++q8) {
        const uint32_t row = (row_block() * TSV_RPW + q8) * TSV_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const int q = 1 - ((i + j) & 1);
        const uint32_t idx = row * g.nhr + m0;
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        d2 o[6], inv = zero;
        for (int f = 0; f < 6; ++f) o[f] = zero;
        d2 a = zero, b = zero;
        d2 c0 = zero, c1 = zero;
        if (ok) {
            if (MODE != 0) inv = *(const d2*)(rows + (size_t)1 * rstride(Nh) + idx);
#pragma unroll
            for (int f = 0; f < 6; ++f)
                o[f] = ld_once(rows + ((size_t)(f + 1) * 2 + 1) * rstride(Nh) + idx);
            if (SYS & 1) c0 = *(const d2*)(p0 + idx);
            if (SYS & 2) c1 = *(const d2*)(p1 + idx);
        }
        if (MODE == 1 || MODE == 3 || MODE == 4) {
            // the black neighbours' diagonals (black slot 0), in the order
            // of the faces; an absent neighbour has coefficient 0
            const double* db = rows;
            const d2 dc = ok ? *(const d2*)(db + idx) : zero;
            d2 dn[6];
            dn[0] = (ok && hip_) ? *(const d2*)(db + idx + g.plane) : zero;
            dn[1] = (ok && him) ? *(const d2*)(db + idx - g.plane) : zero;
            dn[2] = (ok && hjp) ? *(const d2*)(db + idx + g.nhr) : zero;
            dn[3] = (ok && hjm) ? *(const d2*)(db + idx - g.nhr) : zero;
            tsv_kn(db, idx, m0, g.nhr, q, ok, dc, dn[5], dn[4]);
            inv = d2{1.0 / inv.x, 1.0 / inv.y};
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                o[f].x = dn[f].x != 0.0 ? (o[f].x * inv.x) / dn[f].x : 0.0;
                o[f].y = dn[f].y != 0.0 ? (o[f].y * inv.y) / dn[f].y : 0.0;
            }
            if (ok && MODE == 1) {
                *(d2*)(rows + (size_t)1 * rstride(Nh) + idx) = inv;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                    *(d2*)(rows + ((size_t)(f + 1) * 2 + 1) * rstride(Nh) + idx) = o[f];
            }
            if (ok && (MODE == 1 || MODE == 3)) {
                if (rows32 != nullptr) {
                    typedef float f2v __attribute__((ext_vector_type(2)));
#pragma unroll
                    for (int f = 0; f < 6; ++f)
                        __builtin_nontemporal_store(
                            f2v{(float)o[f].x, (float)o[f].y},
                            (f2v*)(rows32 + ((size_t)f * 2 + 1) * rstride32(Nh) + idx));
                }
            }
        }
        if (SYS & 1) {
            d2 ip = (ok && hip_) ? *(const d2*)(p0 + idx + g.plane) : zero;
            d2 im = (ok && him) ? *(const d2*)(p0 + idx - g.plane) : zero;
            d2 jp = (ok && hjp) ? *(const d2*)(p0 + idx + g.nhr) : zero;
            d2 jm = (ok && hjm) ? *(const d2*)(p0 + idx - g.nhr) : zero;
            d2 km, kp;
            tsv_kn(p0, idx, m0, g.nhr, q, ok, c0, km, kp);
            a = o[0] * ip;
            a = a + o[1] * im;
            a = a + o[2] * jp;
            a = a + o[3] * jm;
            a = a + o[4] * kp;
            a = a + o[5] * km;
        }
        if (SYS & 2) {
            d2 ip = (ok && hip_) ? *(const d2*)(p1 + idx + g.plane) : zero;
            d2 im = (ok && him) ? *(const d2*)(p1 + idx - g.plane) : zero;
            d2 jp = (ok && hjp) ? *(const d2*)(p1 + idx + g.nhr) : zero;
            d2 jm = (ok && hjm) ? *(const d2*)(p1 + idx - g.nhr) : zero;
            d2 km, kp;
            tsv_kn(p1, idx, m0, g.nhr, q, ok, c1, km, kp);
            b = o[0] * ip;
            b = b + o[1] * im;
            b = b + o[2] * jp;
            b = b + o[3] * jm;
            b = b + o[4] * kp;
            b = b + o[5] * km;
        }
        if (!ok) continue;
        if (MODE == 0) {
            if (SYS & 1) *(d2*)(w0 + idx) = a;
            if (SYS & 2) *(d2*)(w1 + idx) = b;
        } else {
            // the pair's voxels of both colours are the full-layout quad
            // 2 idx .. 2 idx + 3; the red ones are at 2 idx + q, 2 idx + 2 + q
            if (SYS & 1) {
                const d2 lo = *(const d2*)(bf0 + 2u * idx), hi = *(const d2*)(bf0 + 2u * idx + 2);
                const d2 br = q ? d2{lo.y, hi.y} : d2{lo.x, hi.x};
                *(d2*)(w0 + idx) = br * inv - a;
            }
            if (SYS & 2) {
                const d2 lo = *(const d2*)(bf1 + 2u * idx), hi = *(const d2*)(bf1 + 2u * idx + 2);
                const d2 br = q ? d2{lo.y, hi.y} : d2{lo.x, hi.x};
                *(d2*)(w1 + idx) = br * inv - b;
            }
        }
    }
}

// Black pass on the scaled unknowns, the modes of k_tschur_black; the final
// one (MODE 3) writes the full-layout quad of both colours.
template <int MODE, int SYS>
__global__ __launch_bounds__(WAVE * TSV_TR) void k_tschur_black_v(
    const double* __restrict__ rows, TVGeo g, uint32_t Nh, const double* __restrict__ x0,
    const double* __restrict__ x1, const double* __restrict__ wr0, const double* __restrict__ wr1,
    double* __restrict__ y0, double* __restrict__ y1, const double* __restrict__ w0,
    const double* __restrict__ w1, const BicgScalars* __restrict__ bs,
    double* __restrict__ partials, double* __restrict__ rh0, double* __restrict__ rh1,
    double* __restrict__ p0, double* __restrict__ p1, double* __restrict__ v0,
    double* __restrict__ v1, const double* __restrict__ bf0, const double* __restrict__ bf1,
    double* __restrict__ xf0, double* __restrict__ xf1, float* __restrict__ rows32) {
    constexpr bool S0 = (SYS & 1) != 0, S1 = (SYS & 2) != 0;
    constexpr bool START = MODE >= 2;
    if (!START && bs->all_done) return;
    __shared__ double lds[4 * TSV_TR];
    const bool a0 = START || !bs->done[0], a1 = START || !bs->done[1];
    const uint32_t m0 = 2u * (blockIdx.x * WAVE + threadIdx.x);
    const bool ok = m0 < g.nhr;
    const d2 zero = d2{0.0, 0.0};
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int q8 = 0; q8 < TSV_RPW; ++q8) {
        const uint32_t row = (row_block() * TSV_RPW + q8) * TSV_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const int q = (i + j) & 1;
        const uint32_t idx = row * g.nhr + m0;
        d2 o[6], d = d2{1.0, 1.0};
        for (int f = 0; f < 6; ++f) o[f] = zero;
        d2 xa = zero, xb = zero, c0 = zero, c1 = zero;
        if (ok) {
            if (MODE == 3) d = *(const d2*)(rows + idx);
#pragma unroll
            for (int f = 0; f < 6; ++f)
                o[f] = ld_once(rows + ((size_t)(f + 1) * 2) * rstride(Nh) + idx);
            if (S0) {
                xa = *(const d2*)(x0 + idx);
                c0 = *(const d2*)(wr0 + idx);
            }
            if (S1) {
                xb = *(const d2*)(x1 + idx);
                c1 = *(const d2*)(wr1 + idx);
            }
        }
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        d2 ya = zero, yb = zero;
        if (S0) {
            d2 ip = (ok && hip_) ? *(const d2*)(wr0 + idx + g.plane) : zero;
            d2 im = (ok && him) ? *(const d2*)(wr0 + idx - g.plane) : zero;
            d2 jp = (ok && hjp) ? *(const d2*)(wr0 + idx + g.nhr) : zero;
            d2 jm = (ok && hjm) ? *(const d2*)(wr0 + idx - g.nhr) : zero;
            d2 km, kp;
            tsv_kn(wr0, idx, m0, g.nhr, q, ok, c0, km, kp);
            ya = xa;
            if (START) {
                ya = ya + o[0] * ip;
                ya = ya + o[1] * im;
                ya = ya + o[2] * jp;
                ya = ya + o[3] * jm;
                ya = ya + o[4] * kp;
                ya = ya + o[5] * km;
            } else {
                ya = ya - o[0] * ip;
                ya = ya - o[1] * im;
                ya = ya - o[2] * jp;
                ya = ya - o[3] * jm;
                ya = ya - o[4] * kp;
                ya = ya - o[5] * km;
            }
        }
        if (S1) {
            d2 ip = (ok && hip_) ? *(const d2*)(wr1 + idx + g.plane) : zero;
            d2 im = (ok && him) ? *(const d2*)(wr1 + idx - g.plane) : zero;
            d2 jp = (ok && hjp) ? *(const d2*)(wr1 + idx + g.nhr) : zero;
            d2 jm = (ok && hjm) ? *(const d2*)(wr1 + idx - g.nhr) : zero;
            d2 km, kp;
            tsv_kn(wr1, idx, m0, g.nhr, q, ok, c1, km, kp);
            yb = xb;
            if (START) {
                yb = yb + o[0] * ip;
                yb = yb + o[1] * im;
                yb = yb + o[2] * jp;
                yb = yb + o[3] * jm;
                yb = yb + o[4] * kp;
                yb = yb + o[5] * km;
            } else {
                yb = yb - o[0] * ip;
                yb = yb - o[1] * im;
                yb = yb - o[2] * jp;
                yb = yb - o[3] * jm;
                yb = yb - o[4] * kp;
                yb = yb - o[5] * km;
            }
        }
        if (!ok) continue;
        if (START) {
            if (MODE == 2 && rows32 != nullptr) {
                typedef float f2v __attribute__((ext_vector_type(2)));
#pragma unroll
                for (int f = 0; f < 6; ++f)
                    __builtin_nontemporal_store(f2v{(float)o[f].x, (float)o[f].y},
                                                (f2v*)(rows32 + ((size_t)f * 2) * rstride32(Nh) + idx));
            }
            // the pair's voxels of both colours are the full-layout quad
            // 2 idx .. 2 idx + 3, the black ones at 2 idx + q, 2 idx + 2 + q;
            // the final pass writes the whole quad: x (black) and the red
            // values u the red pass left in wr at the same compacted index.
            // rh, p and v (FP64 BiCGStab) are not written when null (the
            // mixed-precision form lays its FP32 vectors over them)
            if (S0) {
                const d2 lo = *(const d2*)(bf0 + 2u * idx), hi = *(const d2*)(bf0 + 2u * idx + 2);
                const d2 bbk = q ? d2{lo.y, hi.y} : d2{lo.x, hi.x};
                const d2 ra = bbk - ya;
                *(d2*)(y0 + idx) = ra;
                if (rh0 != nullptr) {
                    *(d2*)(rh0 + idx) = ra;
                    *(d2*)(p0 + idx) = zero;
                    *(d2*)(v0 + idx) = zero;
                }
                if (MODE == 3) {
                    const d2 xu = xa / d;
                    *(d2*)(xf0 + 2u * idx) = q ? d2{c0.x, xu.x} : d2{xu.x, c0.x};
                    *(d2*)(xf0 + 2u * idx + 2) = q ? d2{c0.y, xu.y} : d2{xu.y, c0.y};
                }
                // ||b|| of the full system over the quad
                s0 += lo.x * lo.x + lo.y * lo.y;
                s0 += hi.x * hi.x + hi.y * hi.y;
                s2 += ra.x * ra.x;
                s2 += ra.y * ra.y;
            }
            if (S1) {
                const d2 lo = *(const d2*)(bf1 + 2u * idx), hi = *(const d2*)(bf1 + 2u * idx + 2);
                const d2 bbk = q ? d2{lo.y, hi.y} : d2{lo.x, hi.x};
                const d2 rb = bbk - yb;
                *(d2*)(y1 + idx) = rb;
                if (rh1 != nullptr) {
                    *(d2*)(rh1 + idx) = rb;
                    *(d2*)(p1 + idx) = zero;
                    *(d2*)(v1 + idx) = zero;
                }
                if (MODE == 3) {
                    const d2 xu = xb / d;
                    *(d2*)(xf1 + 2u * idx) = q ? d2{c1.x, xu.x} : d2{xu.x, c1.x};
                    *(d2*)(xf1 + 2u * idx + 2) = q ? d2{c1.y, xu.y} : d2{xu.y, c1.y};
                }
                s1 += lo.x * lo.x + lo.y * lo.y;
                s1 += hi.x * hi.x + hi.y * hi.y;
                s3 += rb.x * rb.x;
                s3 += rb.y * rb.y;
            }
        } else {
            if (S0 && a0) *(d2*)(y0 + idx) = ya;
            if (S1 && a1) *(d2*)(y1 + idx) = yb;
            if (MODE == 0) {
                if (S0) {
                    const d2 r = *(const d2*)(w0 + idx);
                    s0 += r.x * ya.x;
                    s0 += r.y * ya.y;
                }
                if (S1) {
                    const d2 r = *(const d2*)(w1 + idx);
                    s1 += r.x * yb.x;
                    s1 += r.y * yb.y;
                }
            } else {
                if (S0) {
                    const d2 sv = *(const d2*)(w0 + idx);
                    s0 += ya.x * ya.x;
                    s0 += ya.y * ya.y;
                    s2 += ya.x * sv.x;
                    s2 += ya.y * sv.y;
                }
                if (S1) {
                    const d2 sv = *(const d2*)(w1 + idx);
                    s1 += yb.x * yb.x;
                    s1 += yb.y * yb.y;
                    s3 += yb.x * sv.x;
                    s3 += yb.y * sv.y;
                }
            }
        }
    }
    // block partials: wave sums, then the waves in order
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    s3 = wave_sum(s3);
    if (threadIdx.x == 0) {
        lds[threadIdx.y] = s0;
        lds[TSV_TR + threadIdx.y] = s1;
        lds[2 * TSV_TR + threadIdx.y] = s2;
        lds[3 * TSV_TR + threadIdx.y] = s3;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int nb = block_count();
        const int b = block_linear();
        for (int a = 0; a < 4; ++a) {
            double t = 0.0;
            for (int w = 0; w < TSV_TR; ++w) t += lds[a * TSV_TR + w];
            partials[a * nb + b] = t;
        }
    }
}

__global__ void k_bicg_alpha(const double* __restrict__ partials, int n, BicgScalars* bs) {
    if (bs->all_done) return;
    __shared__ double lds[1024 / WAVE];
    __shared__ double tot[2];
    sum_partials(partials, n, 2, tot, lds);
    if (threadIdx.x == 0) {
        for (int f = 0; f < 2; ++f) {
            if (bs->done[f]) continue;
            bs->r0v[f] = tot[f];
            if (tot[f] == 0.0 || !isfinite(tot[f])) {  // breakdown: restart from x
                bs->done[f] = 1;
                bs->converged[f] = 0;
                bs->alpha[f] = 0.0;
            } else {
                bs->alpha[f] = bs->rho[f] / tot[f];
            }
        }
    }
}

// s = r - alpha v; with PRE sh = s / diag; partial s.s
template <bool PRE>
__global__ void k_bicg_s(const double* __restrict__ r0_, const double* __restrict__ r1_,
                         const double* __restrict__ v0, const double* __restrict__ v1,
                         double* __restrict__ s0, double* __restrict__ s1, double* __restrict__ sh0,
                         double* __restrict__ sh1, const double* __restrict__ invd,
                         const BicgScalars* __restrict__ bs, uint32_t N,
                         double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const double al0 = bs->alpha[0], al1 = bs->alpha[1];
    double q0 = 0.0, q1 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const double id = PRE ? invd[v] : 1.0;
        if (a0) {
            double sv = r0_[v] - al0 * v0[v];
            s0[v] = sv;
            if (PRE) sh0[v] = sv * id;
            q0 += sv * sv;
        }
        if (a1) {
            double sv = r1_[v] - al1 * v1[v];
            s1[v] = sv;
            if (PRE) sh1[v] = sv * id;
            q1 += sv * sv;
        }
    }
    double q[2] = {q0, q1};
    block_reduce<2, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q[0];
        partials[gridDim.x + blockIdx.x] = q[1];
    }
}

__global__ void k_bicg_scheck(const double* __restrict__ partials, int n, BicgScalars* bs) {
    if (bs->all_done) return;
    __shared__ double lds[1024 / WAVE];
    __shared__ double tot[2];
    sum_partials(partials, n, 2, tot, lds);
    if (threadIdx.x == 0)
        for (int f = 0; f < 2; ++f) {
            if (bs->done[f]) continue;
            bs->ss[f] = tot[f];
            bs->s_conv[f] = tot[f] <= bs->tol2[f];
        }
}

__global__ void k_bicg_omega(const double* __restrict__ partials, int n, BicgScalars* bs) {
    if (bs->all_done) return;
    __shared__ double lds[1024 / WAVE];
    __shared__ double tot[4];
    sum_partials(partials, n, 4, tot, lds);
    if (threadIdx.x == 0)
        for (int f = 0; f < 2; ++f) {
            if (bs->done[f]) continue;
            const double tt = tot[f], ts = tot[2 + f];
            bs->tt[f] = tt;
            bs->ts[f] = ts;
            bs->omega[f] = (bs->s_conv[f] || !(tt > 0.0)) ? 0.0 : ts / tt;
        }
}

// x += alpha ph + omega sh (without PRE: x += alpha p + omega s); r = s -
// omega t; partials r.r, r0.r
template <bool PRE>
__global__ void k_bicg_x(double* __restrict__ x0, double* __restrict__ x1,
                         double* __restrict__ r0_, double* __restrict__ r1_,
                         const double* __restrict__ ph0, const double* __restrict__ ph1,
                         const double* __restrict__ sh0, const double* __restrict__ sh1,
                         const double* __restrict__ s0, const double* __restrict__ s1,
                         const double* __restrict__ t0, const double* __restrict__ t1,
                         const double* __restrict__ rh0, const double* __restrict__ rh1,
                         const BicgScalars* __restrict__ bs, uint32_t N,
                         double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const double al0 = bs->alpha[0], al1 = bs->alpha[1], om0 = bs->omega[0], om1 = bs->omega[1];
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        if (a0) {
            const double sv = s0[v];
            x0[v] += al0 * ph0[v] + om0 * (PRE ? sh0[v] : sv);
            double rv = sv - om0 * t0[v];
            r0_[v] = rv;
            q0 += rv * rv;
            q2 += rh0[v] * rv;
        }
        if (a1) {
            const double sv = s1[v];
            x1[v] += al1 * ph1[v] + om1 * (PRE ? sh1[v] : sv);
            double rv = sv - om1 * t1[v];
            r1_[v] = rv;
            q1 += rv * rv;
            q3 += rh1[v] * rv;
        }
    }
    double q[4] = {q0, q1, q2, q3};
    block_reduce<4, RED_SUM, RED_SUM, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[blockIdx.x] = q[0];
        partials[nb + blockIdx.x] = q[1];
        partials[2 * nb + blockIdx.x] = q[2];
        partials[3 * nb + blockIdx.x] = q[3];
    }
}

// The Schur form's three vector updates on pairs of values, 16 B per lane
// and array (k_bicg_p, k_bicg_s, k_bicg_x without PRE otherwise): an odd
// length leaves its last value to thread 0 of block 0.  Buffers start on
// 16-byte boundaries.
__device__ __forceinline__ d2 ld2(const double* __restrict__ a, uint32_t u) {
    return *(const d2*)(a + 2 * (size_t)u);
}
__device__ __forceinline__ void st2(double* __restrict__ a, uint32_t u, d2 x) {
    *(d2*)(a + 2 * (size_t)u) = x;
}

__global__ __launch_bounds__(EW_BLOCK) void k_bicg_p_v(
    double* __restrict__ p0, double* __restrict__ p1, const double* __restrict__ r0_,
    const double* __restrict__ r1_, const double* __restrict__ v0,
    const double* __restrict__ v1, const BicgScalars* __restrict__ bs, uint32_t N) {
    if (bs->all_done) return;
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const double be0 = bs->beta[0], om0 = bs->omega[0], be1 = bs->beta[1], om1 = bs->omega[1];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t u = gid; u < N / 2; u += gridDim.x * blockDim.x) {
        if (a0) st2(p0, u, ld2(r0_, u) + be0 * (ld2(p0, u) - om0 * ld2(v0, u)));
        if (a1) st2(p1, u, ld2(r1_, u) + be1 * (ld2(p1, u) - om1 * ld2(v1, u)));
    }
    if ((N & 1) && gid == 0) {
        const uint32_t v = N - 1;
        if (a0) p0[v] = r0_[v] + be0 * (p0[v] - om0 * v0[v]);
        if (a1) p1[v] = r1_[v] + be1 * (p1[v] - om1 * v1[v]);
    }
}

__global__ __launch_bounds__(EW_BLOCK) void k_bicg_s_v(
    const double* __restrict__ r0_, const double* __restrict__ r1_,
    const double* __restrict__ v0, const double* __restrict__ v1, double* __restrict__ s0,
    double* __restrict__ s1, const BicgScalars* __restrict__ bs, uint32_t N,
    double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const double al0 = bs->alpha[0], al1 = bs->alpha[1];
    double q0 = 0.0, q1 = 0.0;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t u = gid; u < N / 2; u += gridDim.x * blockDim.x) {
        if (a0) {
            const d2 sv = ld2(r0_, u) - al0 * ld2(v0, u);
            st2(s0, u, sv);
            q0 += sv.x * sv.x;
            q0 += sv.y * sv.y;
        }
        if (a1) {
            const d2 sv = ld2(r1_, u) - al1 * ld2(v1, u);
            st2(s1, u, sv);
            q1 += sv.x * sv.x;
            q1 += sv.y * sv.y;
        }
    }
    if ((N & 1) && gid == 0) {
        const uint32_t v = N - 1;
        if (a0) {
            const double sv = r0_[v] - al0 * v0[v];
            s0[v] = sv;
            q0 += sv * sv;
        }
        if (a1) {
            const double sv = r1_[v] - al1 * v1[v];
            s1[v] = sv;
            q1 += sv * sv;
        }
    }
    double q[2] = {q0, q1};
    block_reduce<2, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q[0];
        partials[gridDim.x + blockIdx.x] = q[1];
    }
}

// x += alpha p + omega s; r = s - omega t; partials r.r, r0.r
__global__ __launch_bounds__(EW_BLOCK) void k_bicg_x_v(
    double* __restrict__ x0, double* __restrict__ x1, double* __restrict__ r0_,
    double* __restrict__ r1_, const double* __restrict__ p0, const double* __restrict__ p1,
    const double* __restrict__ s0, const double* __restrict__ s1,
    const double* __restrict__ t0, const double* __restrict__ t1,
    const double* __restrict__ rh0, const double* __restrict__ rh1,
    const BicgScalars* __restrict__ bs, uint32_t N, double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const double al0 = bs->alpha[0], al1 = bs->alpha[1], om0 = bs->omega[0], om1 = bs->omega[1];
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t u = gid; u < N / 2; u += gridDim.x * blockDim.x) {
        if (a0) {
            const d2 sv = ld2(s0, u);
            st2(x0, u, ld2(x0, u) + (al0 * ld2(p0, u) + om0 * sv));
            const d2 rv = sv - om0 * ld2(t0, u);
            st2(r0_, u, rv);
            const d2 hv = ld2(rh0, u);
            q0 += rv.x * rv.x;
            q0 += rv.y * rv.y;
            q2 += hv.x * rv.x;
            q2 += hv.y * rv.y;
        }
        if (a1) {
            const d2 sv = ld2(s1, u);
            st2(x1, u, ld2(x1, u) + (al1 * ld2(p1, u) + om1 * sv));
            const d2 rv = sv - om1 * ld2(t1, u);
            st2(r1_, u, rv);
            const d2 hv = ld2(rh1, u);
            q1 += rv.x * rv.x;
            q1 += rv.y * rv.y;
            q3 += hv.x * rv.x;
            q3 += hv.y * rv.y;
        }
    }
    if ((N & 1) && gid == 0) {
        const uint32_t v = N - 1;
        if (a0) {
            const double sv = s0[v];
            x0[v] += al0 * p0[v] + om0 * sv;
            const double rv = sv - om0 * t0[v];
            r0_[v] = rv;
            q0 += rv * rv;
            q2 += rh0[v] * rv;
        }
        if (a1) {
            const double sv = s1[v];
            x1[v] += al1 * p1[v] + om1 * sv;
            const double rv = sv - om1 * t1[v];
            r1_[v] = rv;
            q1 += rv * rv;
            q3 += rh1[v] * rv;
        }
    }
    double q[4] = {q0, q1, q2, q3};
    block_reduce<4, RED_SUM, RED_SUM, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[blockIdx.x] = q[0];
        partials[nb + blockIdx.x] = q[1];
        partials[2 * nb + blockIdx.x] = q[2];
        partials[3 * nb + blockIdx.x] = q[3];
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision transport on the Schur complement: BiCGStab in FP32 on an
// FP32 copy of the scaled operator S' = I - A_br R' solves for a correction
// e of the FP64 iterate, inside FP64 refinement cycles.  A cycle takes the
// FP64 Schur residual of the start pass, scaled by a power of two to unit
// size, solves S' e = r to the FP64 tolerance (with a margin) or to a fixed
// reduction of its own start, adds scale * e to y in FP64, and recomputes
// the FP64 residual with the start pass; that pass's stop test is the
// solve's.  FP32 copies of the off-diagonal coefficients (12 arrays of
// N/2 + TROW_PAD32 floats: the red rows of R' and the black rows of A_br)
// are written by the first start pass.  Vectors are FP32, dot products are
// accumulated in FP64.
// ---------------------------------------------------------------------------
template <int VW> struct FVec;
template <> struct FVec<2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct FVec<4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct FVec<8> { typedef float type __attribute__((ext_vector_type(8))); };

// Blocks of the FP32 passes: TMX_TR waves, each over TMX_RPW rows; TMX_VW
// values per lane (nk % (2 TMX_VW) == 0; 4 otherwise).  At 512^3 the red
// pass took 404-405 us with 8 x 2 or 8 x 4 or 4 x 2, 410 us with 4 x 4,
// 436 us with 2 x 8, and 476 us with 8 values per lane.
#ifndef LS_TMIX_TR
#define LS_TMIX_TR 8
#endif
#ifndef LS_TMIX_RPW
#define LS_TMIX_RPW 2
#endif
#ifndef LS_TMIX_VW
#define LS_TMIX_VW 4
#endif
constexpr int TMX_TR = LS_TMIX_TR, TMX_RPW = LS_TMIX_RPW, TMX_VW = LS_TMIX_VW;

// The other colour's values at the k neighbours of the lane's VW values
// (km at -k, kp at +k); q is the own k parity (as tsv_kn).
template <int VW, typename V = typename FVec<VW>::type>
__device__ __forceinline__ void tmx_kn(const float* __restrict__ src, uint32_t idx, uint32_t m0,
                                       uint32_t nhr, int q, bool ok, V c, V& km, V& kp) {
    const int lane = threadIdx.x;
    float prev = __shfl_up(c[VW - 1], 1, WAVE);
    float next = __shfl_down(c[0], 1, WAVE);
    if (lane == 0) prev = (ok && q == 0 && m0 >= 1) ? src[idx - 1] : 0.0f;
    if (lane == WAVE - 1) next = (ok && q == 1 && m0 + VW < nhr) ? src[idx + VW] : 0.0f;
    if (q == 0) {
        km[0] = m0 >= 1 ? prev : 0.0f;
#pragma unroll
        for (int e = 1; e < VW; ++e) km[e] = c[e - 1];
        kp = c;
    } else {
        km = c;
#pragma unroll
        for (int e = 0; e + 1 < VW; ++e) kp[e] = c[e + 1];
        kp[VW - 1] = m0 + VW < nhr ? next : 0.0f;
    }
}

template <int VW, typename V = typename FVec<VW>::type>
__device__ __forceinline__ V tmx_ld(const float* __restrict__ a, size_t i) {
    return *(const V*)(a + i);
}

// The six neighbours of the lane's values in src, o-weighted: sum_f o_f n_f.
template <int VW, typename V = typename FVec<VW>::type>
__device__ __forceinline__ V tmx_gather(const float* __restrict__ src, const V (&o)[6], V c,
                                        uint32_t idx, uint32_t m0, const TVGeo& g, int q, bool ok,
                                        bool hip_, bool him, bool hjp, bool hjm) {
    const V zero = (V)(0.0f);
    const V ip = (ok && hip_) ? tmx_ld<VW>(src, idx + g.plane) : zero;
    const V im = (ok && him) ? tmx_ld<VW>(src, idx - g.plane) : zero;
    const V jp = (ok && hjp) ? tmx_ld<VW>(src, idx + g.nhr) : zero;
    const V jm = (ok && hjm) ? tmx_ld<VW>(src, idx - g.nhr) : zero;
    V km, kp;
    tmx_kn<VW>(src, idx, m0, g.nhr, q, ok, c, km, kp);
    V a = o[0] * ip;
    a = a + o[1] * im;
    a = a + o[2] * jp;
    a = a + o[3] * jm;
    a = a + o[4] * kp;
    a = a + o[5] * km;
    return a;
}

// Red pass: w = R' p on the FP32 rows.
template <int VW, int SYS>
__global__ __launch_bounds__(WAVE * TMX_TR) void k_tmix_red(
    const float* __restrict__ rows, TVGeo g, uint32_t Nh, const float* __restrict__ p0,
    const float* __restrict__ p1, float* __restrict__ w0, float* __restrict__ w1,
    const BicgScalars* __restrict__ bs) {
    typedef typename FVec<VW>::type V;
    if (bs->all_done) return;
    const uint32_t m0 = VW * (blockIdx.x * WAVE + threadIdx.x);
    const bool ok = m0 < g.nhr;
    const V zero = (V)(0.0f);
    const size_t rs = rstride32(Nh);
#pragma unroll
    for (int q8 = 0; q8 < TMX_RPW; ++q8) {
        const uint32_t row = (row_block() * TMX_RPW + q8) * TMX_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const int q = 1 - ((i + j) & 1);
        const uint32_t idx = row * g.nhr + m0;
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        V o[6];
        V c0 = zero, c1 = zero;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            o[f] = ok ? __builtin_nontemporal_load((const V*)(rows + ((size_t)f * 2 + 1) * rs + idx))
                      : zero;
        if (ok) {
            if (SYS & 1) c0 = tmx_ld<VW>(p0, idx);
            if (SYS & 2) c1 = tmx_ld<VW>(p1, idx);
        }
        V a = zero, b = zero;
        if (SYS & 1) a = tmx_gather<VW>(p0, o, c0, idx, m0, g, q, ok, hip_, him, hjp, hjm);
        if (SYS & 2) b = tmx_gather<VW>(p1, o, c1, idx, m0, g, q, ok, hip_, him, hjp, hjm);
        if (!ok) continue;
        if (SYS & 1) *(V*)(w0 + idx) = a;
        if (SYS & 2) *(V*)(w1 + idx) = b;
    }
}

// Black pass: z = p - A_br w on the FP32 rows, with partials (FP64) of
// rh.z (MODE 0) or (z.z, z.s) (MODE 1), the vector of the dot in d0 / d1.
template <int MODE, int VW, int SYS>
__global__ __launch_bounds__(WAVE * TMX_TR) void k_tmix_black(
    const float* __restrict__ rows, TVGeo g, uint32_t Nh, const float* __restrict__ x0,
    const float* __restrict__ x1, const float* __restrict__ wr0, const float* __restrict__ wr1,
    float* __restrict__ y0, float* __restrict__ y1, const float* __restrict__ d0,
    const float* __restrict__ d1, const BicgScalars* __restrict__ bs,
    double* __restrict__ partials) {
    typedef typename FVec<VW>::type V;
    constexpr bool S0 = (SYS & 1) != 0, S1 = (SYS & 2) != 0;
    if (bs->all_done) return;
    __shared__ double lds[4 * TMX_TR];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const uint32_t m0 = VW * (blockIdx.x * WAVE + threadIdx.x);
    const bool ok = m0 < g.nhr;
    const V zero = (V)(0.0f);
    const size_t rs = rstride32(Nh);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int q8 = 0; q8 < TMX_RPW; ++q8) {
        const uint32_t row = (row_block() * TMX_RPW + q8) * TMX_TR + threadIdx.y;
        if (row >= g.nrows) continue;                   // uniform per wave
        const int i = (int)(row / (uint32_t)g.nj), j = (int)(row - (uint32_t)i * g.nj);
        const int q = (i + j) & 1;
        const uint32_t idx = row * g.nhr + m0;
        const bool hip_ = i + 1 < g.ni, him = i > 0, hjp = j + 1 < g.nj, hjm = j > 0;
        V o[6];
        V xa = zero, xb = zero, c0 = zero, c1 = zero;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            o[f] = ok ? __builtin_nontemporal_load((const V*)(rows + ((size_t)f * 2) * rs + idx))
                      : zero;
        if (ok) {
            if (S0) {
                xa = tmx_ld<VW>(x0, idx);
                c0 = tmx_ld<VW>(wr0, idx);
            }
            if (S1) {
                xb = tmx_ld<VW>(x1, idx);
                c1 = tmx_ld<VW>(wr1, idx);
            }
        }
        V ya = zero, yb = zero;
        if (S0) ya = xa - tmx_gather<VW>(wr0, o, c0, idx, m0, g, q, ok, hip_, him, hjp, hjm);
        if (S1) yb = xb - tmx_gather<VW>(wr1, o, c1, idx, m0, g, q, ok, hip_, him, hjp, hjm);
        if (!ok) continue;
        if (S0 && a0) *(V*)(y0 + idx) = ya;
        if (S1 && a1) *(V*)(y1 + idx) = yb;
        if (S0) {
            const V dv = tmx_ld<VW>(d0, idx);
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                if (MODE == 0) {
                    s0 += (double)dv[e] * (double)ya[e];
                } else {
                    s0 += (double)ya[e] * (double)ya[e];
                    s2 += (double)ya[e] * (double)dv[e];
                }
            }
        }
        if (S1) {
            const V dv = tmx_ld<VW>(d1, idx);
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                if (MODE == 0) {
                    s1 += (double)dv[e] * (double)yb[e];
                } else {
                    s1 += (double)yb[e] * (double)yb[e];
                    s3 += (double)yb[e] * (double)dv[e];
                }
            }
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    s3 = wave_sum(s3);
    if (threadIdx.x == 0) {
        lds[threadIdx.y] = s0;
        lds[TMX_TR + threadIdx.y] = s1;
        lds[2 * TMX_TR + threadIdx.y] = s2;
        lds[3 * TMX_TR + threadIdx.y] = s3;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int nb = block_count();
        const int b = block_linear();
        for (int a = 0; a < 4; ++a) {
            double t = 0.0;
            for (int w = 0; w < TMX_TR; ++w) t += lds[a * TMX_TR + w];
            partials[a * nb + b] = t;
        }
    }
}

// The scale of a cycle's residual: the power of two at or below its norm.
__device__ __forceinline__ double tmix_scale(double rr) {
    return ldexp(1.0, ilogb(sqrt(rr)));
}

typedef float f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4 ldf4(const float* __restrict__ a, uint32_t u) {
    return *(const f4*)(a + 4 * (size_t)u);
}
__device__ __forceinline__ void stf4(float* __restrict__ a, uint32_t u, f4 x) {
    *(f4*)(a + 4 * (size_t)u) = x;
}
__device__ __forceinline__ double dot4(f4 a, f4 b) {
    double s = (double)a.x * (double)b.x;
    s += (double)a.y * (double)b.y;
    s += (double)a.z * (double)b.z;
    s += (double)a.w * (double)b.w;
    return s;
}

// The FP32 vectors of one system, passed by value: r, rh, p, v, s, t, e, w.
struct TMixVecs {
    float* v[8];
};

// Start of a cycle for the systems still iterating in FP64: r = r64 / scale
// (FP32), rh = r, p = v = e = 0; partials r.r.  N % 4 == 0.
__global__ __launch_bounds__(EW_BLOCK) void k_tmix_start(
    const double* __restrict__ g0, const double* __restrict__ g1, TMixVecs V0, TMixVecs V1,
    const BicgScalars* __restrict__ bs64, uint32_t N,
    double* __restrict__ partials) {
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    const bool a0 = !bs64->done[0], a1 = !bs64->done[1];
    const double i0 = a0 ? 1.0 / tmix_scale(bs64->rr[0]) : 0.0;
    const double i1 = a1 ? 1.0 / tmix_scale(bs64->rr[1]) : 0.0;
    double q0 = 0.0, q1 = 0.0;
    const f4 z = (f4)(0.0f);
    float *__restrict__ r0 = V0.v[0], *__restrict__ h0 = V0.v[1], *__restrict__ p0 = V0.v[2],
          *__restrict__ w0 = V0.v[3], *__restrict__ e0 = V0.v[6];
    float *__restrict__ r1 = V1.v[0], *__restrict__ h1 = V1.v[1], *__restrict__ p1 = V1.v[2],
          *__restrict__ w1 = V1.v[3], *__restrict__ e1 = V1.v[6];
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N / 4;
         u += gridDim.x * blockDim.x) {
        if (a0) {
            const d2 lo = ld2(g0, 2 * u), hi = ld2(g0, 2 * u + 1);
            const f4 r = f4{(float)(lo.x * i0), (float)(lo.y * i0), (float)(hi.x * i0),
                            (float)(hi.y * i0)};
            stf4(r0, u, r);
            stf4(h0, u, r);
            stf4(p0, u, z);
            stf4(w0, u, z);
            stf4(e0, u, z);
            q0 += dot4(r, r);
        }
        if (a1) {
            const d2 lo = ld2(g1, 2 * u), hi = ld2(g1, 2 * u + 1);
            const f4 r = f4{(float)(lo.x * i1), (float)(lo.y * i1), (float)(hi.x * i1),
                            (float)(hi.y * i1)};
            stf4(r1, u, r);
            stf4(h1, u, r);
            stf4(p1, u, z);
            stf4(w1, u, z);
            stf4(e1, u, z);
            q1 += dot4(r, r);
        }
    }
    double q[2] = {q0, q1};
    block_reduce<2, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q[0];
        partials[gridDim.x + blockIdx.x] = q[1];
    }
}

// The cycle's FP32 scalars: r0 = r, p = r on the first iteration; stop at
// the FP64 tolerance over the scale with a margin of 2, or at a reduction
// by kappa, whichever is reached first; the iteration count and cap carry
// over from the FP64 scalars.
__global__ void k_tmix_init(const double* __restrict__ partials, int n, BicgScalars* bs,
                            const BicgScalars* __restrict__ bs64, double kappa) {
    __shared__ double lds[1024 / WAVE];
    __shared__ double tot[2];
    sum_partials(partials, n, 2, tot, lds);
    if (threadIdx.x == 0) {
        bs->maxit = bs64->maxit;
        int all = 1;
        for (int f = 0; f < 2; ++f) {
            bs->iter[f] = bs64->iter[f];
            bs->s_conv[f] = 0;
            bs->alpha[f] = 1.0;
            bs->omega[f] = 1.0;
            bs->beta[f] = 0.0;
            if (bs64->done[f]) {
                bs->done[f] = 1;
                bs->converged[f] = bs64->converged[f];
                bs->rr[f] = 0.0;
                continue;
            }
            const double rr = tot[f];
            const double sc = tmix_scale(bs64->rr[f]);
            const double abs2 = 0.25 * bs64->tol2[f] / (sc * sc);
            bs->rr[f] = rr;
            bs->rho[f] = rr;
            bs->r0r[f] = rr;
            bs->bnorm2[f] = rr;
            bs->tol2[f] = fmax(kappa * kappa * rr, abs2);
            const int ok = !(rr > bs->tol2[f]);
            bs->done[f] = ok;
            bs->converged[f] = ok;
            all &= ok;
        }
        bs->all_done = all;
    }
}

// p = r + beta (p - omega v), FP32.
__global__ __launch_bounds__(EW_BLOCK) void k_tmix_p(TMixVecs V0, TMixVecs V1,
                                                     const BicgScalars* __restrict__ bs,
                                                     uint32_t N) {
    if (bs->all_done) return;
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const float be0 = (float)bs->beta[0], om0 = (float)bs->omega[0];
    const float be1 = (float)bs->beta[1], om1 = (float)bs->omega[1];
    const float *__restrict__ r0 = V0.v[0], *__restrict__ w0 = V0.v[3];
    const float *__restrict__ r1 = V1.v[0], *__restrict__ w1 = V1.v[3];
    float *__restrict__ p0 = V0.v[2], *__restrict__ p1 = V1.v[2];
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N / 4;
         u += gridDim.x * blockDim.x) {
        if (a0) stf4(p0, u, ldf4(r0, u) + be0 * (ldf4(p0, u) - om0 * ldf4(w0, u)));
        if (a1) stf4(p1, u, ldf4(r1, u) + be1 * (ldf4(p1, u) - om1 * ldf4(w1, u)));
    }
}

// s = r - alpha v, FP32; partials s.s
__global__ __launch_bounds__(EW_BLOCK) void k_tmix_s(TMixVecs V0, TMixVecs V1,
                                                     const BicgScalars* __restrict__ bs,
                                                     uint32_t N, double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const float al0 = (float)bs->alpha[0], al1 = (float)bs->alpha[1];
    const float *__restrict__ r0 = V0.v[0], *__restrict__ w0 = V0.v[3];
    const float *__restrict__ r1 = V1.v[0], *__restrict__ w1 = V1.v[3];
    float *__restrict__ s0 = V0.v[4], *__restrict__ s1 = V1.v[4];
    double q0 = 0.0, q1 = 0.0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N / 4;
         u += gridDim.x * blockDim.x) {
        if (a0) {
            const f4 sv = ldf4(r0, u) - al0 * ldf4(w0, u);
            stf4(s0, u, sv);
            q0 += dot4(sv, sv);
        }
        if (a1) {
            const f4 sv = ldf4(r1, u) - al1 * ldf4(w1, u);
            stf4(s1, u, sv);
            q1 += dot4(sv, sv);
        }
    }
    double q[2] = {q0, q1};
    block_reduce<2, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q[0];
        partials[gridDim.x + blockIdx.x] = q[1];
    }
}

// e += alpha p + omega s; r = s - omega t, FP32; partials r.r, rh.r
__global__ __launch_bounds__(EW_BLOCK) void k_tmix_x(TMixVecs V0, TMixVecs V1,
                                                     const BicgScalars* __restrict__ bs,
                                                     uint32_t N, double* __restrict__ partials) {
    if (bs->all_done) return;
    __shared__ double lds[4 * EW_BLOCK / WAVE];
    const bool a0 = !bs->done[0], a1 = !bs->done[1];
    const float al0 = (float)bs->alpha[0], al1 = (float)bs->alpha[1];
    const float om0 = (float)bs->omega[0], om1 = (float)bs->omega[1];
    const float *__restrict__ h0 = V0.v[1], *__restrict__ p0 = V0.v[2],
                *__restrict__ s0 = V0.v[4], *__restrict__ t0 = V0.v[5];
    const float *__restrict__ h1 = V1.v[1], *__restrict__ p1 = V1.v[2],
                *__restrict__ s1 = V1.v[4], *__restrict__ t1 = V1.v[5];
    float *__restrict__ r0 = V0.v[0], *__restrict__ e0 = V0.v[6];
    float *__restrict__ r1 = V1.v[0], *__restrict__ e1 = V1.v[6];
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N / 4;
         u += gridDim.x * blockDim.x) {
        if (a0) {
            const f4 sv = ldf4(s0, u);
            stf4(e0, u, ldf4(e0, u) + (al0 * ldf4(p0, u) + om0 * sv));
            const f4 rv = sv - om0 * ldf4(t0, u);
            stf4(r0, u, rv);
            q0 += dot4(rv, rv);
            q2 += dot4(ldf4(h0, u), rv);
        }
        if (a1) {
            const f4 sv = ldf4(s1, u);
            stf4(e1, u, ldf4(e1, u) + (al1 * ldf4(p1, u) + om1 * sv));
            const f4 rv = sv - om1 * ldf4(t1, u);
            stf4(r1, u, rv);
            q1 += dot4(rv, rv);
            q3 += dot4(ldf4(h1, u), rv);
        }
    }
    double q[4] = {q0, q1, q2, q3};
    block_reduce<4, RED_SUM, RED_SUM, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        partials[blockIdx.x] = q[0];
        partials[nb + blockIdx.x] = q[1];
        partials[2 * nb + blockIdx.x] = q[2];
        partials[3 * nb + blockIdx.x] = q[3];
    }
}

// End of a cycle: y += scale e (FP64) for the systems iterated; the FP64
// scalars take the iteration count.
__global__ __launch_bounds__(EW_BLOCK) void k_tmix_update(
    double* __restrict__ y0, double* __restrict__ y1, const float* __restrict__ e0,
    const float* __restrict__ e1, BicgScalars* __restrict__ bs64,
    const BicgScalars* __restrict__ bs, uint32_t N) {
    const bool a0 = !bs64->done[0], a1 = !bs64->done[1];
    const double sc0 = a0 ? tmix_scale(bs64->rr[0]) : 0.0;
    const double sc1 = a1 ? tmix_scale(bs64->rr[1]) : 0.0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N / 4;
         u += gridDim.x * blockDim.x) {
        if (a0) {
            const f4 e = ldf4(e0, u);
            st2(y0, 2 * u, ld2(y0, 2 * u) + sc0 * d2{(double)e.x, (double)e.y});
            st2(y0, 2 * u + 1, ld2(y0, 2 * u + 1) + sc0 * d2{(double)e.z, (double)e.w});
        }
        if (a1) {
            const f4 e = ldf4(e1, u);
            st2(y1, 2 * u, ld2(y1, 2 * u) + sc1 * d2{(double)e.x, (double)e.y});
            st2(y1, 2 * u + 1, ld2(y1, 2 * u + 1) + sc1 * d2{(double)e.z, (double)e.w});
        }
    }
    // every block reads bs64->rr / done above before any write here: the
    // counts are the only fields written, and no block reads them
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int f = 0; f < 2; ++f) bs64->iter[f] = bs->iter[f];
}

__global__ void k_bicg_rho(const double* __restrict__ partials, int n, BicgScalars* bs) {
    if (bs->all_done) return;
    __shared__ double lds[1024 / WAVE];
    __shared__ double tot[4];
    sum_partials(partials, n, 4, tot, lds);
    if (threadIdx.x == 0) {
        int all = 1;
        for (int f = 0; f < 2; ++f) {
            if (!bs->done[f]) {
                const double rr = tot[f], rho_new = tot[2 + f];
                bs->rr[f] = rr;
                bs->iter[f] += 1;
                if (!isfinite(rr)) {
                    bs->done[f] = 1;
                    bs->converged[f] = 0;
                } else if (rr <= bs->tol2[f]) {
                    bs->done[f] = 1;
                    bs->converged[f] = 1;
                } else if (bs->iter[f] >= bs->maxit) {
                    bs->done[f] = 1;
                    bs->converged[f] = 0;
                } else if (rho_new == 0.0 || bs->omega[f] == 0.0) {
                    bs->done[f] = 1;      // breakdown: the host restarts from x
                    bs->converged[f] = 0;
                } else {
                    bs->beta[f] = (rho_new / bs->rho[f]) * (bs->alpha[f] / bs->omega[f]);
                    bs->rho[f] = rho_new;
                }
            }
            all &= bs->done[f];
        }
        bs->all_done = all;
    }
}

// first iteration uses rho = r0.r of the (re)start
__global__ void k_bicg_first_rho(BicgScalars* bs) {
    for (int f = 0; f < 2; ++f) {
        if (bs->done[f]) continue;
        bs->beta[f] = 0.0;          // p = r
        bs->omega[f] = 1.0;
        bs->rho[f] = bs->r0r[f];
    }
}

// A system with b = 0 is solved by x = 0, which the host has just written:
// it takes no iteration, and the other system keeps its start state.
__global__ void k_bicg_zero_rhs(BicgScalars* bs) {
    for (int f = 0; f < 2; ++f)
        if (bs->bnorm2[f] == 0.0) {
            bs->rr[f] = 0.0;
            bs->done[f] = 1;
            bs->converged[f] = 1;
        }
    bs->all_done = bs->done[0] && bs->done[1];
}

// Refinement cycles of the mixed-precision Schur form before FP64 BiCGStab
// takes over.
#ifndef LS_ADR_MAX_CYCLES
#define LS_ADR_MAX_CYCLES 30
#endif

// Lock-step BiCGStab of the two systems x0, x1 (KIND 0: conc, acid; KIND 1:
// the implicit saturation in x0, x1 idle).  Returns 0 converged, 1 not.
template <int KIND>
static int bicg_pair(ls_ctx* c, double* x0_, double* x1_, const double* sf, const AdrConst& a,
                     int* iters_out, double* relres_out, int* sys_iters = nullptr) {
    TGeo g;
    g.ni = c->ni; g.nj = c->nj; g.nk = c->nk;
    g.sj = (uint32_t)c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    const uint32_t N = (uint32_t)c->N;
    const int nb = ew_blocks(N);
    double* x[2] = {x0_, x1_};
    double *r[2], *rh[2], *pp[2], *vv[2], *s[2], *t[2], *ph[2], *sh[2];
    for (int f = 0; f < 2; ++f) {
        r[f] = c->tv[f][0]; rh[f] = c->tv[f][1]; pp[f] = c->tv[f][2]; vv[f] = c->tv[f][3];
        s[f] = c->tv[f][4]; t[f] = c->tv[f][5]; ph[f] = c->tv[f][6]; sh[f] = c->tv[f][7];
    }
    const int batch = 4;
    int restarts = 0;
    bool first = true;
    c->t_cycles = 0;
    c->t_fallback = 0;
    c->t_rows32_valid = false;
    // stored rows (nk even) and the Schur form, as the pool was cut for:
    // the solver then runs on the black unknowns (N/2 each)
    const dim3 vg = adr_vec_grid(c);
    const int nbm = (int)(vg.x * vg.y * vg.z);
    const bool stored = c->t_stored;
    const bool tschur = c->t_schur;
    // every phase using the pool lays its own buffers over the pressure system
    c->press_assembled = false;
    // system 1 known to be 0 with b = 0 (the idle system of the implicit
    // saturation; acid that has never had a source): the Schur form then
    // builds, starts and recovers system 0 alone
    const bool one = tschur && (KIND == 1 || c->acid_zero);
    // transport zeroes the flux at Rock and Precipitate voxels (the Schur
    // form's row builder does it itself)
    if (KIND == 0 && !tschur) {
        if (int st = zero_flux_rock_precip(c)) return st;
    }
    const uint32_t Nh = N / 2;
    const int nbh = ew_blocks(Nh);
    double *xs[2] = {c->txs[0], c->txs[1]}, *wr[2] = {c->twr[0], c->twr[1]};
    if (tschur) {
        // transport: conc starts from its extrapolation in time through the
        // fields of the last LS_ADR_EXTRAP_ORDER solves
        double* hist[3] = {nullptr, nullptr, nullptr};
        auto history = [&](double** buf) {
            if (*buf == nullptr) {
                if (hipMalloc((void**)buf, Nh * sizeof(double)) == hipSuccess) {
                    c->allocations.push_back(*buf);
                    c->grid_bytes += Nh * sizeof(double);
                } else {        // out of memory: a lower order instead
                    *buf = nullptr;
                    (void)hipGetLastError();
                }
            }
            return *buf;
        };
        int held = 0;           // history buffers held (the extrapolation's order)
        if (KIND == 0 && c->P.adr_extrapolate)
            while (held < LS_ADR_EXTRAP_ORDER && held < 3 && history(&c->thist[held]) != nullptr) {
                hist[held] = c->thist[held];
                ++held;
            }
        // the current field goes into the oldest buffer, which then becomes
        // the newest
        double* const hw = held > 0 ? hist[held - 1] : nullptr;
        const double h = 1.0 / a.inv_dt;
        // Lagrange weights at t = h through the current field (t = 0) and the
        // usable history fields (t = -(dt_0 + ... + dt_k))
        Extrap ex{};
        ex.n = std::min(held, c->tprev_n);
        for (int k = 0; k < ex.n; ++k)
            if (!(c->thist_dt[k] > 0.0)) ex.n = k;
        {
            double tau[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 1; k <= ex.n; ++k) tau[k] = tau[k - 1] - c->thist_dt[k - 1];
            for (int q = 0; q <= ex.n; ++q) {
                double w = 1.0;
                for (int l = 0; l <= ex.n; ++l)
                    if (l != q) w *= (h - tau[l]) / (tau[q] - tau[l]);
                ex.w[q] = w;
            }
        }
        // conc's rows with a voxel pair per lane where the rows fit the grid
        TVGeo pg;
        pg.nhr = (uint32_t)(c->nk / 2);
        pg.nrows = (uint32_t)c->ni * (uint32_t)c->nj;
        pg.plane = (uint32_t)c->nj * pg.nhr;
        pg.ni = c->ni;
        pg.nj = c->nj;
        const dim3 pgrid = row_grid((pg.nhr + WAVE - 1) / WAVE,
                                    (pg.nrows + TSV_TR * TSV_RPW - 1) / (TSV_TR * TSV_RPW));
        if (KIND == 0 && a.theta == 1.0)
            hipLaunchKernelGGL(k_adr_rows_pair<false>, pgrid, dim3(WAVE, TSV_TR), 0, c->stream,
                               c->img, c->qx, c->qy, c->qz, pg, N, a, c->arows, x[0], x[1],
                               c->tb[0], c->tb[1], one ? 1 : 0, xs[0], xs[1], hist[0], hist[1],
                               hist[2], hw, ex);
        else if (KIND == 0)
            hipLaunchKernelGGL(k_adr_rows_pair<true>, pgrid, dim3(WAVE, TSV_TR), 0, c->stream,
                               c->img, c->qx, c->qy, c->qz, pg, N, a, c->arows, x[0], x[1],
                               c->tb[0], c->tb[1], one ? 1 : 0, xs[0], xs[1], hist[0], hist[1],
                               hist[2], hw, ex);
        else
            hipLaunchKernelGGL((k_adr_rows_split<KIND>), dim3(nb), dim3(EW_BLOCK), 0, c->stream,
                               c->img, c->qx, c->qy, c->qz, sf, g, N, a, c->arows, x[0], x[1],
                               c->tb[0], c->tb[1], one ? 1 : 0, xs[0], xs[1], hist[0], hist[1],
                               hist[2], hw, ex);
        if (KIND == 0) {
            if (held == 0) {
                c->tprev_n = 0;
            } else {
                for (int k = held - 1; k > 0; --k) c->thist[k] = c->thist[k - 1];
                c->thist[0] = hw;
                c->thist_dt[2] = c->thist_dt[1];
                c->thist_dt[1] = c->thist_dt[0];
                c->thist_dt[0] = h;
                c->tprev_n = std::min(c->tprev_n + 1, held);
            }
        }
        LS_LAUNCH_CHECK();
    } else if (stored) {       // rows, right-hand sides and D^-1 once per solve
        hipLaunchKernelGGL((k_adr_rows<KIND>), dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->img,
                           c->qx, c->qy, c->qz, sf, g, N, a, c->arows, x[0], x[1], c->tb[0],
                           c->tb[1], c->tinvd);
        LS_LAUNCH_CHECK();
    }
    // the solver's unknowns and length; the Schur form has the Jacobi
    // scaling inside its operator (unknowns y = D_b x_b), the full system
    // applies D^-1 to the directions
    double* xv[2] = {tschur ? xs[0] : x[0], tschur ? xs[1] : x[1]};
    const uint32_t Nv = tschur ? Nh : N;
    const int nbv = tschur ? nbh : nb;
    const double* invd_v = c->tinvd;
    // the vectorised Schur passes: nk % 4 == 0 (16 B aligned pairs per row)
    // and the row count within the grid's y limit
    TVGeo tvg;
    tvg.nhr = (uint32_t)(c->nk / 2);
    tvg.nrows = (uint32_t)c->ni * (uint32_t)c->nj;
    tvg.plane = (uint32_t)c->nj * tvg.nhr;
    tvg.ni = c->ni;
    tvg.nj = c->nj;
    const dim3 tvgrid = row_grid((tvg.nhr / 2 + WAVE - 1) / WAVE,
                                 (tvg.nrows + TSV_TR * TSV_RPW - 1) / (TSV_TR * TSV_RPW));
    const size_t nb_tv = (size_t)tvgrid.x * tvgrid.y * tvgrid.z;
    const bool tvec = tschur && c->nk % 4 == 0 && 4 * nb_tv <= 8 * (size_t)MAX_PARTIALS;
    const int nbs_schur = tvec ? (int)nb_tv : nbh;
    // mixed precision (the pool was cut for it): FP32 cycles, 4 values per
    // lane in the passes, for conc alone (acid known zero: its vectors hold
    // the FP32 rows)
    const bool tm_wide = TMX_VW != 4 && c->nk % (2 * TMX_VW) == 0;
    const dim3 tmgrid = row_grid((tvg.nhr / (tm_wide ? TMX_VW : 4) + WAVE - 1) / WAVE,
                                 (tvg.nrows + TMX_TR * TMX_RPW - 1) / (TMX_TR * TMX_RPW));
    const int nb_mix = (int)((size_t)tmgrid.x * tmgrid.y * tmgrid.z);
    bool mixed = tvec && one && c->t_mixed && c->trows32 != nullptr &&
                 4 * (size_t)nb_mix <= 8 * (size_t)MAX_PARTIALS;
    double* nulls[2] = {nullptr, nullptr};
    // systems still iterating, as the host last saw them (a system that
    // converges inside a batch is masked on the device)
    int sys = 3;
    auto launch_rows = [&](auto MC, auto SC, const double* xa, const double* xb, double* ya,
                           double* yb, const double* wa, const double* wb) {
        constexpr int M = decltype(MC)::value, S = decltype(SC)::value;
        hipLaunchKernelGGL((k_adr_matvec_rows<M, S>), vg, dim3(WAVE, ADR_TJ), 0, c->stream,
                           c->arows, xa, xb, ya, yb, wa, wb, g, N, c->bicg, c->partials, rh[0],
                           rh[1], pp[0], pp[1], vv[0], vv[1]);
    };
    // profiling: the first red pass of every batch is timed with events
    bool sample_red = false, red_pending = false;
    auto launch_schur = [&](auto MC, auto SC, const double* xa, const double* xb, double* ya,
                            double* yb, const double* wa, const double* wb) {
        constexpr int M = decltype(MC)::value, S = decltype(SC)::value;
        const bool sample = sample_red;
        if (sample) {
            (void)hipEventRecord(c->ev[16], c->stream);
            sample_red = false;
        }
        // black mode M: 0, 1 apply; 2 start; 3 final start (fields written)
        constexpr int RM = M == 2 ? 1 : (M == 3 ? 2 : 0);
        // the mixed form: the first start writes the FP32 rows, and no
        // start writes the FP64 rh, p, v its FP32 vectors lie over
        float* const r32 = (mixed && M == 2) ? c->trows32 : nullptr;
        double* const* rhs = mixed ? nulls : rh;
        double* const* pps = mixed ? nulls : pp;
        double* const* vvs = mixed ? nulls : vv;
        if (tvec && mixed && RM != 0)
            // unscaled FP64 red rows, rescaled in registers
            hipLaunchKernelGGL((k_tschur_red_v<S, RM == 1 ? 3 : 4>), tvgrid, dim3(WAVE, TSV_TR),
                               0, c->stream, c->arows, tvg, Nh, xa, xb, wr[0], wr[1], c->bicg,
                               c->tb[0], c->tb[1], x[0], x[1], r32);
        else if (tvec)
            hipLaunchKernelGGL((k_tschur_red_v<S, RM>), tvgrid, dim3(WAVE, TSV_TR), 0, c->stream,
                               c->arows, tvg, Nh, xa, xb, wr[0], wr[1], c->bicg, c->tb[0],
                               c->tb[1], x[0], x[1], r32);
        else
            hipLaunchKernelGGL((k_tschur_red<S, RM>), dim3(nbh), dim3(EW_BLOCK), 0, c->stream,
                               c->arows, g, Nh, xa, xb, wr[0], wr[1], c->bicg, c->tb[0], c->tb[1],
                               x[0], x[1]);
        if (sample) {
            (void)hipEventRecord(c->ev[17], c->stream);
            red_pending = true;
            // red rows (6 coefficients of the red voxels), p at the black
            // voxels and w at the red ones, per system iterated
            c->prof_tred_bytes = (double)c->N * (24.0 + (S == 3 ? 16.0 : 8.0));
        }
        if (tvec)
            hipLaunchKernelGGL((k_tschur_black_v<M, S>), tvgrid, dim3(WAVE, TSV_TR), 0, c->stream,
                               c->arows, tvg, Nh, xa, xb, wr[0], wr[1], ya, yb, wa, wb, c->bicg,
                               c->partials, rhs[0], rhs[1], pps[0], pps[1], vvs[0], vvs[1],
                               c->tb[0], c->tb[1], x[0], x[1], r32);
        else
            hipLaunchKernelGGL((k_tschur_black<M, S>), dim3(nbh), dim3(EW_BLOCK), 0, c->stream,
                               c->arows, g, Nh, xa, xb, wr[0], wr[1], ya, yb, wa, wb, c->bicg,
                               c->partials, rh[0], rh[1], pp[0], pp[1], vv[0], vv[1], c->tb[0],
                               c->tb[1], x[0], x[1]);
    };
    auto matvec = [&](auto MC, const double* xa, const double* xb, double* ya, double* yb,
                      const double* wa, const double* wb) {
        if (tschur) {
            if (sys == 1)
                launch_schur(MC, std::integral_constant<int, 1>{}, xa, xb, ya, yb, wa, wb);
            else if (sys == 2)
                launch_schur(MC, std::integral_constant<int, 2>{}, xa, xb, ya, yb, wa, wb);
            else
                launch_schur(MC, std::integral_constant<int, 3>{}, xa, xb, ya, yb, wa, wb);
            return nbs_schur;
        }
        if (stored) {
            if (sys == 1)
                launch_rows(MC, std::integral_constant<int, 1>{}, xa, xb, ya, yb, wa, wb);
            else if (sys == 2)
                launch_rows(MC, std::integral_constant<int, 2>{}, xa, xb, ya, yb, wa, wb);
            else
                launch_rows(MC, std::integral_constant<int, 3>{}, xa, xb, ya, yb, wa, wb);
        } else {
            constexpr int M = decltype(MC)::value;
            hipLaunchKernelGGL((k_adr_matvec<M, KIND>), dim3(nb), dim3(EW_BLOCK), 0, c->stream,
                               c->img, c->qx, c->qy, c->qz, sf, xa, xb, ya, yb, wa, wb, g, N, a,
                               c->bicg, c->partials);
        }
        return stored ? nbm : nb;
    };
    // the mixed form's FP32 apply z = S' p for conc (red pass into w, then
    // the black pass with the partials of its mode)
    auto mix_apply = [&](auto MC, const float* pa, const float* pb, float* za, float* zb,
                         const float* da, const float* db) {
        constexpr int M = decltype(MC)::value;
        auto go = [&](auto SC) {
            constexpr int S = decltype(SC)::value;
            const bool sample = sample_red;
            if (sample) {
                (void)hipEventRecord(c->ev[16], c->stream);
                sample_red = false;
            }
            if (tm_wide)
                hipLaunchKernelGGL((k_tmix_red<TMX_VW, S>), tmgrid, dim3(WAVE, TMX_TR), 0,
                                   c->stream, c->trows32, tvg, Nh, pa, pb, c->tf[0][7], nullptr,
                                   c->bicg32);
            else
                hipLaunchKernelGGL((k_tmix_red<4, S>), tmgrid, dim3(WAVE, TMX_TR), 0, c->stream,
                                   c->trows32, tvg, Nh, pa, pb, c->tf[0][7], nullptr, c->bicg32);
            if (sample) {
                (void)hipEventRecord(c->ev[17], c->stream);
                red_pending = true;
                // FP32 red rows (6 coefficients at N/2), p at the black and w
                // at the red voxels per system iterated
                c->prof_tred_bytes = (double)c->N * (12.0 + (S == 3 ? 8.0 : 4.0));
            }
            if (tm_wide)
                hipLaunchKernelGGL((k_tmix_black<M, TMX_VW, S>), tmgrid, dim3(WAVE, TMX_TR), 0,
                                   c->stream, c->trows32, tvg, Nh, pa, pb, c->tf[0][7], nullptr,
                                   za, zb, da, db, c->bicg32, c->partials);
            else
                hipLaunchKernelGGL((k_tmix_black<M, 4, S>), tmgrid, dim3(WAVE, TMX_TR), 0,
                                   c->stream, c->trows32, tvg, Nh, pa, pb, c->tf[0][7], nullptr,
                                   za, zb, da, db, c->bicg32, c->partials);
        };
        go(std::integral_constant<int, 1>{});     // conc alone
        return nb_mix;
    };
    // r = b - A x, rh = r, p = v = 0, partials (b.b, r.r); in the Schur form
    // every start after the first also writes the iterate into the fields
    // (the red values from the black ones), so the last one leaves the
    // solution there
    auto start = [&](bool build) {
        if (tschur) {
            sys = one ? 1 : 3;
            if (build) {
                if (one)
                    launch_schur(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{},
                                 xs[0], xs[1], r[0], r[1], nullptr, nullptr);
                else
                    launch_schur(std::integral_constant<int, 2>{}, std::integral_constant<int, 3>{},
                                 xs[0], xs[1], r[0], r[1], nullptr, nullptr);
            } else {
                if (one)
                    launch_schur(std::integral_constant<int, 3>{}, std::integral_constant<int, 1>{},
                                 xs[0], xs[1], r[0], r[1], nullptr, nullptr);
                else
                    launch_schur(std::integral_constant<int, 3>{}, std::integral_constant<int, 3>{},
                                 xs[0], xs[1], r[0], r[1], nullptr, nullptr);
            }
            return nbs_schur;
        }
        if (stored) {
            sys = 3;
            launch_rows(std::integral_constant<int, 2>{}, std::integral_constant<int, 3>{}, x[0],
                        x[1], r[0], r[1], c->tb[0], c->tb[1]);
            return nbm;
        }
        hipLaunchKernelGGL((k_adr_start<KIND>), dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->img,
                           c->qx, c->qy, c->qz, sf, x[0], x[1], c->tb[0], c->tb[1], c->tinvd,
                           r[0], r[1], rh[0], rh[1], pp[0], pp[1], vv[0], vv[1], g, N, a,
                           build ? 1 : 0, c->partials);
        return nb;
    };
    while (true) {
        const int ns = start(first);
        hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1024), 0, c->stream, c->partials, ns, c->bicg,
                           c->P.adr_rtol, c->P.adr_maxit, first ? 1 : 0);
        hipLaunchKernelGGL(k_bicg_first_rho, dim3(1), dim3(1), 0, c->stream, c->bicg);
        LS_LAUNCH_CHECK();
        LS_HIP(hipMemcpyAsync(c->h_bicg, c->bicg, sizeof(BicgScalars), hipMemcpyDeviceToHost,
                              c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        if (first && (c->h_bicg->bnorm2[0] == 0.0 || c->h_bicg->bnorm2[1] == 0.0)) {
            // a system with b = 0 has x = 0 exactly: write it and mark the
            // system solved (the start state of the other one stands)
            for (int f = 0; f < 2; ++f)
                if (c->h_bicg->bnorm2[f] == 0.0 && !(one && f == 1)) {
                    LS_HIP(hipMemsetAsync(x[f], 0, c->N * sizeof(double), c->stream));
                    if (tschur) LS_HIP(hipMemsetAsync(xs[f], 0, Nh * sizeof(double), c->stream));
                }
            hipLaunchKernelGGL(k_bicg_zero_rhs, dim3(1), dim3(1), 0, c->stream, c->bicg);
            LS_LAUNCH_CHECK();
            LS_HIP(hipMemcpyAsync(c->h_bicg, c->bicg, sizeof(BicgScalars), hipMemcpyDeviceToHost,
                                  c->stream));
            LS_HIP(hipStreamSynchronize(c->stream));
        }
        first = false;
        if (mixed) {
            c->t_rows32_valid = true;
            // FP32 cycles on the FP64 residual of the last start pass; the
            // start pass after each cycle writes the fields and is the stop
            // test.  A cycle that does not cut an unconverged system's FP64
            // residual by 16x hands the solve to FP64 BiCGStab.
            const double kappa = c->P.adr_mixed_inner_rtol;
            // system 1 is idle (known zero, done at every start): its
            // vectors are never touched, the kernels get conc's again
            TMixVecs V0, V1;
            for (int k = 0; k < 8; ++k) {
                V0.v[k] = c->tf[0][k];
                V1.v[k] = c->tf[0][k];
            }
            int cycles = 0;
            bool wrote = false, stalled = false;
            while (!c->h_bicg->all_done) {
                const BicgScalars h0 = *c->h_bicg;
                hipLaunchKernelGGL(k_tmix_start, dim3(nbh), dim3(EW_BLOCK), 0, c->stream, r[0],
                                   r[1], V0, V1, c->bicg, Nh, c->partials);
                hipLaunchKernelGGL(k_tmix_init, dim3(1), dim3(1024), 0, c->stream, c->partials,
                                   nbh, c->bicg32, c->bicg, kappa);
                LS_LAUNCH_CHECK();
                c->h_bicg32->all_done = 0;
                c->h_bicg32->done[0] = h0.done[0];
                c->h_bicg32->done[1] = h0.done[1];
                while (!c->h_bicg32->all_done) {
                    sample_red = c->profiling;
                    for (int it = 0; it < batch; ++it) {
                        hipLaunchKernelGGL(k_tmix_p, dim3(nbh), dim3(EW_BLOCK), 0, c->stream, V0,
                                           V1, c->bicg32, Nh);
                        // v = S' p, partials rh.v
                        const int n0 = mix_apply(std::integral_constant<int, 0>{}, V0.v[2],
                                                 V1.v[2], V0.v[3], V1.v[3], V0.v[1], V1.v[1]);
                        hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(1024), 0, c->stream,
                                           c->partials, n0, c->bicg32);
                        hipLaunchKernelGGL(k_tmix_s, dim3(nbh), dim3(EW_BLOCK), 0, c->stream, V0,
                                           V1, c->bicg32, Nh, c->partials);
                        hipLaunchKernelGGL(k_bicg_scheck, dim3(1), dim3(1024), 0, c->stream,
                                           c->partials, nbh, c->bicg32);
                        // t = S' s, partials t.t, t.s
                        const int n1 = mix_apply(std::integral_constant<int, 1>{}, V0.v[4],
                                                 V1.v[4], V0.v[5], V1.v[5], V0.v[4], V1.v[4]);
                        hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(1024), 0, c->stream,
                                           c->partials, n1, c->bicg32);
                        hipLaunchKernelGGL(k_tmix_x, dim3(nbh), dim3(EW_BLOCK), 0, c->stream, V0,
                                           V1, c->bicg32, Nh, c->partials);
                        hipLaunchKernelGGL(k_bicg_rho, dim3(1), dim3(1024), 0, c->stream,
                                           c->partials, nbh, c->bicg32);
                    }
                    LS_LAUNCH_CHECK();
                    LS_HIP(hipMemcpyAsync(c->h_bicg32, c->bicg32, sizeof(BicgScalars),
                                          hipMemcpyDeviceToHost, c->stream));
                    LS_HIP(hipStreamSynchronize(c->stream));
                    if (red_pending) {
                        float ms = 0.0f;
                        LS_HIP(hipEventElapsedTime(&ms, c->ev[16], c->ev[17]));
                        c->prof_ms[3] += ms;
                        c->prof_n[3] += 1;
                        red_pending = false;
                    }
                }
                hipLaunchKernelGGL(k_tmix_update, dim3(nbh), dim3(EW_BLOCK), 0, c->stream, xs[0],
                                   xs[1], V0.v[6], V1.v[6], c->bicg, c->bicg32, Nh);
                const int nc = start(false);
                hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1024), 0, c->stream, c->partials,
                                   nc, c->bicg, c->P.adr_rtol, c->P.adr_maxit, 0);
                LS_LAUNCH_CHECK();
                LS_HIP(hipMemcpyAsync(c->h_bicg, c->bicg, sizeof(BicgScalars),
                                      hipMemcpyDeviceToHost, c->stream));
                LS_HIP(hipStreamSynchronize(c->stream));
                wrote = true;
                c->t_cycles += 1;
                const BicgScalars& h = *c->h_bicg;
                if (h.all_done) break;
                if (h.iter[0] >= c->P.adr_maxit || h.iter[1] >= c->P.adr_maxit) break;
                stalled = ++cycles >= LS_ADR_MAX_CYCLES;
                for (int f = 0; f < 2; ++f)
                    if (!h.done[f] && !(h.rr[f] < 0.0625 * h0.rr[f])) stalled = true;
                if (stalled) break;
            }
            if (!stalled) {
                if (!wrote) {       // converged at the first start: write the fields
                    const int nc = start(false);
                    hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1024), 0, c->stream,
                                       c->partials, nc, c->bicg, c->P.adr_rtol, c->P.adr_maxit, 0);
                    LS_LAUNCH_CHECK();
                    LS_HIP(hipMemcpyAsync(c->h_bicg, c->bicg, sizeof(BicgScalars),
                                          hipMemcpyDeviceToHost, c->stream));
                    LS_HIP(hipStreamSynchronize(c->stream));
                }
                break;
            }
            // FP64 BiCGStab from the current iterate, on the red rows rescaled
            // to R' in place (the mixed form left them unscaled)
            hipLaunchKernelGGL((k_tschur_red_v<1, 1>), tvgrid, dim3(WAVE, TSV_TR), 0, c->stream,
                               c->arows, tvg, Nh, xs[0], xs[1], wr[0], wr[1], c->bicg, c->tb[0],
                               c->tb[1], x[0], x[1], nullptr);
            LS_LAUNCH_CHECK();
            mixed = false;
            c->t_fallback = 1;
            continue;
        }
        while (!c->h_bicg->all_done) {
            sys = (c->h_bicg->done[0] ? 0 : 1) | (c->h_bicg->done[1] ? 0 : 2);
            if (sys == 0) sys = 3;
            sample_red = c->profiling && tschur;
            // the directions the operator is applied to: D^-1 p and D^-1 s
            // (full system), or p and s themselves (Schur form)
            double* const* pd = tschur ? pp : ph;
            double* const* sd = tschur ? s : sh;
            for (int it = 0; it < batch; ++it) {
                if (tschur)
                    hipLaunchKernelGGL(k_bicg_p_v, dim3(nbv), dim3(EW_BLOCK), 0, c->stream,
                                       pp[0], pp[1], r[0], r[1], vv[0], vv[1], c->bicg, Nv);
                else
                    hipLaunchKernelGGL(k_bicg_p<true>, dim3(nbv), dim3(EW_BLOCK), 0, c->stream,
                                       pp[0], pp[1], r[0], r[1], vv[0], vv[1], ph[0], ph[1],
                                       invd_v, c->bicg, Nv);
                const int n0 = matvec(std::integral_constant<int, 0>{}, pd[0], pd[1], vv[0],
                                      vv[1], rh[0], rh[1]);
                hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(1024), 0, c->stream, c->partials,
                                   n0, c->bicg);
                if (tschur)
                    hipLaunchKernelGGL(k_bicg_s_v, dim3(nbv), dim3(EW_BLOCK), 0, c->stream,
                                       r[0], r[1], vv[0], vv[1], s[0], s[1], c->bicg, Nv,
                                       c->partials);
                else
                    hipLaunchKernelGGL(k_bicg_s<true>, dim3(nbv), dim3(EW_BLOCK), 0, c->stream,
                                       r[0], r[1], vv[0], vv[1], s[0], s[1], sh[0], sh[1], invd_v,
                                       c->bicg, Nv, c->partials);
                hipLaunchKernelGGL(k_bicg_scheck, dim3(1), dim3(1024), 0, c->stream, c->partials,
                                   nbv, c->bicg);
                const int n1 = matvec(std::integral_constant<int, 1>{}, sd[0], sd[1], t[0], t[1],
                                      s[0], s[1]);
                hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(1024), 0, c->stream, c->partials,
                                   n1, c->bicg);
                if (tschur)
                    hipLaunchKernelGGL(k_bicg_x_v, dim3(nbv), dim3(EW_BLOCK), 0, c->stream,
                                       xv[0], xv[1], r[0], r[1], pp[0], pp[1], s[0], s[1], t[0],
                                       t[1], rh[0], rh[1], c->bicg, Nv, c->partials);
                else
                    hipLaunchKernelGGL(k_bicg_x<true>, dim3(nbv), dim3(EW_BLOCK), 0, c->stream,
                                       xv[0], xv[1], r[0], r[1], ph[0], ph[1], sh[0], sh[1], s[0],
                                       s[1], t[0], t[1], rh[0], rh[1], c->bicg, Nv, c->partials);
                hipLaunchKernelGGL(k_bicg_rho, dim3(1), dim3(1024), 0, c->stream, c->partials, nbv,
                                   c->bicg);
            }
            LS_LAUNCH_CHECK();
            LS_HIP(hipMemcpyAsync(c->h_bicg, c->bicg, sizeof(BicgScalars), hipMemcpyDeviceToHost,
                                  c->stream));
            LS_HIP(hipStreamSynchronize(c->stream));
            if (red_pending) {
                float ms = 0.0f;
                LS_HIP(hipEventElapsedTime(&ms, c->ev[16], c->ev[17]));
                c->prof_ms[3] += ms;
                c->prof_n[3] += 1;
                red_pending = false;
            }
        }
        // confirm on the recomputed residual (the start kernel with the
        // stored right-hand sides); restart systems that fall short
        const int nc = start(false);
        hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1024), 0, c->stream, c->partials, nc, c->bicg,
                           c->P.adr_rtol, c->P.adr_maxit, 0);
        LS_LAUNCH_CHECK();
        LS_HIP(hipMemcpyAsync(c->h_bicg, c->bicg, sizeof(BicgScalars), hipMemcpyDeviceToHost,
                              c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        const BicgScalars& h = *c->h_bicg;
        const bool ok = h.converged[0] && h.converged[1];
        const bool capped = h.iter[0] >= c->P.adr_maxit || h.iter[1] >= c->P.adr_maxit;
        if (ok || capped || restarts >= 20) break;
        ++restarts;
        hipLaunchKernelGGL(k_bicg_first_rho, dim3(1), dim3(1), 0, c->stream, c->bicg);
        // loop re-enters with the start kernel (rebuild of b skipped)
    }
    const BicgScalars& h = *c->h_bicg;
    c->trans_valid = tschur && KIND == 0;
    if (iters_out) *iters_out = h.iter[0] > h.iter[1] ? h.iter[0] : h.iter[1];
    if (sys_iters) {
        sys_iters[0] = h.iter[0];
        sys_iters[1] = h.iter[1];
    }
    if (relres_out) {
        double r0v = h.bnorm2[0] > 0 ? sqrt(h.rr[0] / h.bnorm2[0]) : 0.0;
        double r1v = h.bnorm2[1] > 0 ? sqrt(h.rr[1] / h.bnorm2[1]) : 0.0;
        *relres_out = r0v > r1v ? r0v : r1v;
    }
    return (h.converged[0] && h.converged[1]) ? 0 : 1;
}

// Re-arm the solver scalars so that apply passes run: system 0 iterating,
// system 1 done (timing only; the next solve starts afresh).
__global__ void k_bicg_rearm(BicgScalars* bs) {
    bs->done[0] = 0;
    bs->done[1] = 1;
    bs->all_done = 0;
}

// Time one Schur pass of the conc system on the rows and vectors the last
// transport solve left in the pool: LS_TIME_TRED w = R' p (the red rows' 6
// rescaled coefficients, p at the black voxels, w at the red ones: 32 B per
// voxel), LS_TIME_TBLACK z = p - A_br w with the dot of z and a third vector
// (the black rows' 6 coefficients, p, w, z and the dot's vector: 40 B per
// voxel).  The launch shapes are bicg_pair's.
int time_transport_pass(ls_ctx* c, int which, int reps, double* ms, double* bytes) {
    if (!c->t_schur || !c->trans_valid) {
        c->err = "no transport system in place (the Schur form's rows are built by a "
                 "transport solve; a pressure phase reuses their memory)";
        return -1;
    }
    TGeo g;
    g.ni = c->ni; g.nj = c->nj; g.nk = c->nk;
    g.sj = (uint32_t)c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    const uint32_t N = (uint32_t)c->N, Nh = N / 2;
    const int nbh = ew_blocks(Nh);
    TVGeo tvg;
    tvg.nhr = (uint32_t)(c->nk / 2);
    tvg.nrows = (uint32_t)c->ni * (uint32_t)c->nj;
    tvg.plane = (uint32_t)c->nj * tvg.nhr;
    tvg.ni = c->ni;
    tvg.nj = c->nj;
    const dim3 tvgrid = row_grid((tvg.nhr / 2 + WAVE - 1) / WAVE,
                                 (tvg.nrows + TSV_TR * TSV_RPW - 1) / (TSV_TR * TSV_RPW));
    const bool tvec = c->nk % 4 == 0 &&
                      4 * (size_t)tvgrid.x * tvgrid.y * tvgrid.z <= 8 * (size_t)MAX_PARTIALS;
    // the iteration's first matvec: p -> (red) w -> (black) v, dot with r0
    const double* xa = c->tv[0][2];
    double* w = c->twr[0];
    double* y = c->tv[0][3];
    const double* wd = c->tv[0][1];
    hipLaunchKernelGGL(k_bicg_rearm, dim3(1), dim3(1), 0, c->stream, c->bicg);
    const bool f32 = which == LS_TIME_TRED32 || which == LS_TIME_TBLACK32;
    if (f32 && !c->t_rows32_valid) {
        c->err = "the FP32 transport passes need a transport solve in the mixed-precision form "
                 "(adr_mixed, nk % 8 == 0, acid zero)";
        return -3;
    }
    const dim3 tmgrid = row_grid((tvg.nhr / 4 + WAVE - 1) / WAVE,
                                 (tvg.nrows + TMX_TR * TMX_RPW - 1) / (TMX_TR * TMX_RPW));
    if (f32) hipLaunchKernelGGL(k_bicg_rearm, dim3(1), dim3(1), 0, c->stream, c->bicg32);
    auto launch = [&]() {
        if (which == LS_TIME_TRED32) {
            hipLaunchKernelGGL((k_tmix_red<4, 1>), tmgrid, dim3(WAVE, TMX_TR), 0, c->stream,
                               c->trows32, tvg, Nh, c->tf[0][2], nullptr, c->tf[0][7], nullptr,
                               c->bicg32);
        } else if (which == LS_TIME_TBLACK32) {
            hipLaunchKernelGGL((k_tmix_black<0, 4, 1>), tmgrid, dim3(WAVE, TMX_TR), 0, c->stream,
                               c->trows32, tvg, Nh, c->tf[0][2], nullptr, c->tf[0][7], nullptr,
                               c->tf[0][3], nullptr, c->tf[0][1], nullptr, c->bicg32,
                               c->partials);
        } else if (which == LS_TIME_TRED) {
            if (tvec)
                hipLaunchKernelGGL((k_tschur_red_v<1, 0>), tvgrid, dim3(WAVE, TSV_TR), 0,
                                   c->stream, c->arows, tvg, Nh, xa, nullptr, w, nullptr, c->bicg,
                                   c->tb[0], c->tb[1], nullptr, nullptr, nullptr);
            else
                hipLaunchKernelGGL((k_tschur_red<1, 0>), dim3(nbh), dim3(EW_BLOCK), 0, c->stream,
                                   c->arows, g, Nh, xa, nullptr, w, nullptr, c->bicg, c->tb[0],
                                   c->tb[1], nullptr, nullptr);
        } else {
            if (tvec)
                hipLaunchKernelGGL((k_tschur_black_v<0, 1>), tvgrid, dim3(WAVE, TSV_TR), 0,
                                   c->stream, c->arows, tvg, Nh, xa, nullptr, w, nullptr, y,
                                   nullptr, wd, nullptr, c->bicg, c->partials, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, c->tb[0], c->tb[1],
                                   nullptr, nullptr, nullptr);
            else
                hipLaunchKernelGGL((k_tschur_black<0, 1>), dim3(nbh), dim3(EW_BLOCK), 0,
                                   c->stream, c->arows, g, Nh, xa, nullptr, w, nullptr, y,
                                   nullptr, wd, nullptr, c->bicg, c->partials, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, c->tb[0], c->tb[1],
                                   nullptr, nullptr);
        }
    };
    launch();                  // warm
    hipEvent_t e0 = c->ev[14], e1 = c->ev[15];
    LS_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) launch();
    LS_LAUNCH_CHECK();
    LS_HIP(hipEventRecord(e1, c->stream));
    LS_HIP(hipEventSynchronize(e1));
    float f = 0.0f;
    LS_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = (double)f / reps;
    *bytes = (double)c->N * (which == LS_TIME_TRED ? 32.0 : which == LS_TIME_TBLACK ? 40.0 :
                             which == LS_TIME_TRED32 ? 16.0 : 20.0);
    return 0;
}

static AdrConst adr_const(const ls_ctx* c, double dt) {
    AdrConst a;
    a.inv_dt = 1.0 / dt;
    a.alpha = c->P.alpha_transfer;
    a.theta = c->P.theta;
    a.D = c->P.D;
    a.Dpf = c->P.Dpore_fac;
    a.dx = c->P.dx;
    a.phi = c->P.porosity;
    a.evap = c->P.evap_flux;
    a.diff_pore = -(a.D * a.Dpf) / (a.dx * a.dx);
    a.diff_other = -a.D / (a.dx * a.dx);
    int e2 = 0;
    a.dx_pow2 = std::frexp(a.dx, &e2) == 0.5 ? 1 : 0;
    a.inv_dx = 1.0 / a.dx;
    return a;
}

int phase_transport(ls_ctx* c, double dt, ls_step_report* rep) {
    int st = 0;
    int iters = 0, sys_iters[2] = {0, 0};
    double relres = 0.0;
    st = bicg_pair<0>(c, c->conc, c->acid, nullptr, adr_const(c, dt), &iters, &relres, sys_iters);
    if (rep) {
        rep->adr_iters = iters;
        rep->adr_relres = relres;
        rep->adr_iters_conc = sys_iters[0];
        rep->adr_iters_acid = sys_iters[1];
        rep->adr_cycles = c->t_cycles;
        rep->adr_fallback = c->t_fallback;
    }
    if (st) c->err = "transport solve did not converge";
    return st;
}

// Raw min / max of the saturation (block partials), then the clamp to
// [0, 1] when the raw solution is physical.
__global__ void k_sat_range(const double* __restrict__ s, uint32_t N, double* __restrict__ partials) {
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    double lo = 1e300, hi = -1e300;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        lo = fmin(lo, s[v]);
        hi = fmax(hi, s[v]);
    }
    double a[2] = {lo, hi};
    block_reduce<2, RED_MIN, RED_MAX>(a, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = a[0];
        partials[gridDim.x + blockIdx.x] = a[1];
    }
}

__global__ void k_sat_range_final(const double* __restrict__ partials, int n, double* out) {
    __shared__ double lds[2 * 1024 / WAVE];
    double lo = 1e300, hi = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        lo = fmin(lo, partials[i]);
        hi = fmax(hi, partials[n + i]);
    }
    double a[2] = {lo, hi};
    block_reduce<2, RED_MIN, RED_MAX>(a, lds);
    if (threadIdx.x == 0) {
        out[0] = a[0];
        out[1] = a[1];
    }
}

__global__ void k_sat_clamp(double* __restrict__ s, uint32_t N) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        s[v] = fmax(0.0, fmin(1.0, s[v]));
}

int phase_saturation_implicit(ls_ctx* c, double dt, ls_step_report* rep) {
    const uint32_t N = (uint32_t)c->N;
    const int nb = ew_blocks(N);
    // the assembly-time saturation is frozen in a transport-phase vector of
    // the pool; the second system idles on another, zeroed
    double* sf = c->taux[0];
    double* idle = c->taux[1];
    if (!sf || !idle) {
        c->err = "implicit saturation: no pool vectors (set implicit_sat through the parameters)";
        return -3;
    }
    LS_HIP(hipMemcpyAsync(sf, c->s, c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    LS_HIP(hipMemsetAsync(idle, 0, c->N * sizeof(double), c->stream));
    int iters = 0;
    double relres = 0.0;
    const int st = bicg_pair<1>(c, c->s, idle, sf, adr_const(c, dt), &iters, &relres);
    hipLaunchKernelGGL(k_sat_range, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->s, N, c->partials);
    hipLaunchKernelGGL(k_sat_range_final, dim3(1), dim3(1024), 0, c->stream, c->partials, nb,
                       &c->ss->sat_lo);
    LS_LAUNCH_CHECK();
    LS_HIP(hipMemcpyAsync(c->h_ss, c->ss, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    const double lo = c->h_ss->sat_lo, hi = c->h_ss->sat_hi;
    const bool valid = !(hi > 1.05 || lo < -0.05);
    if (valid) {
        hipLaunchKernelGGL(k_sat_clamp, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->s, N);
        LS_LAUNCH_CHECK();
    }
    c->sat_valid = valid;
    if (rep) {
        rep->sat_iters = iters;
        rep->sat_valid = valid ? 1 : 0;
        rep->sat_raw_min = lo;
        rep->sat_raw_max = hi;
    }
    if (st) c->err = "implicit saturation solve did not converge";
    return st;
}

}  // namespace ls
