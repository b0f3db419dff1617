/*
 * ls_oracle.cpp — CPU FP64 oracle of the LeachingSim per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see ls_oracle.h): the values the HIP path is
 * checked against, and the timed CPU baseline.  Never linked into the product.
 *
 * Every function restates the reference loop it cites, for one domain
 * (global faces only, no ghost planes).  Loops run over the flat k-fastest
 * index; the reference runs k -> j -> i with i innermost, which gives the
 * same result for every phase except updateFrac, whose in-place sweep is
 * order dependent and is therefore run in the reference's exact order.
 *
 * Build: g++ -O2 -fopenmp -shared -fPIC (see oracle/Makefile).  x86-64
 * without -mfma does not contract a*b+c, so every expression rounds as
 * written, like the GPU library built with -ffp-contract=off.
 */
#include "ls_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint8_t AIR = LS_AIR, PORE = LS_PORE, ROCK = LS_ROCK, SUL = LS_SULPHIDE,
                  PREC = LS_PRECIPITATE;

/* Pressure ramp, src/Simulation.cpp:662-670. */
double effective_top_pressure(double top, double t) {
    const double ramp_duration = 3600.0;
    double eff = top;
    if (t < ramp_duration) eff = top * (t / ramp_duration) + 1e-6;
    return eff;
}

/* Harmonic-mean face conductance of a Pore row, src/Simulation.cpp:737-768. */
inline double pore_link(double perm_self, double k_neighbor, double mu_self) {
    double K_self = std::max(perm_self, 1e-25);
    double K_neigh = std::max(k_neighbor, 1e-25);
    double K_avg = (2.0 * K_self * K_neigh) / (K_self + K_neigh);
    return K_avg / mu_self;
}

struct Csr {
    int n = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
    std::vector<double> diag;
    /* Build from COO, summing duplicates (PETSc ADD_VALUES). */
    void build(int n_, const std::vector<int>& r, const std::vector<int>& c,
               const std::vector<double>& v) {
        n = n_;
        ptr.assign(n + 1, 0);
        for (int x : r) ptr[x + 1]++;
        for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        std::vector<int> cc(r.size());
        std::vector<double> vv(r.size());
        for (size_t e = 0; e < r.size(); ++e) {
            int p = pos[r[e]]++;
            cc[p] = c[e];
            vv[p] = v[e];
        }
        /* merge duplicates within each row, keeping insertion order of sums */
        col.clear();
        val.clear();
        std::vector<int> nptr(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            int start = (int)col.size();
            for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
                int found = -1;
                for (int q = start; q < (int)col.size(); ++q)
                    if (col[q] == cc[p]) { found = q; break; }
                if (found >= 0) val[found] += vv[p];
                else { col.push_back(cc[p]); val.push_back(vv[p]); }
            }
            nptr[i + 1] = (int)col.size();
        }
        ptr = nptr;
        diag.assign(n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int p = ptr[i]; p < ptr[i + 1]; ++p)
                if (col[p] == i) diag[i] += val[p];
    }
    void apply(const double* x, double* y) const {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int p = ptr[i]; p < ptr[i + 1]; ++p) s += val[p] * x[col[p]];
            y[i] = s;
        }
    }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    const long long n = (long long)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

}  // namespace

struct lso_ctx {
    int ni = 0, nj = 0, nk = 0;
    size_t N = 0;
    ls_params P{};
    std::string err;

    std::vector<uint8_t> img;
    std::vector<float> frac;
    std::vector<double> inv, s, p, conc, acid, qx, qy, qz;
    std::vector<double> krel, pc, perm, mu;

    std::vector<uint8_t> b_img;
    std::vector<float> b_frac;
    std::vector<double> b_inv, b_s, b_p, b_conc, b_acid, b_qx, b_qy, b_qz;

    /* Reduced pressure operator of the last pressure phase. */
    std::vector<int8_t> cls;          /* 0 Dirichlet, 1 solved, 2 floored   */
    std::vector<double> cf[3];        /* +i/+j/+k face conductance at v     */
    std::vector<double> cin;          /* virtual inlet face, plane i = 0    */
    std::vector<double> d;            /* sum of couplings of a Pore row     */
    double p_top_prev = 0.0;
    bool have_prev = false;
    bool sat_valid = true;      /* physics_valid of the last implicit saturation */
    /* linear solvers: 0 Jacobi-PCG on the reduced SPD pressure system (stop
     * test ||b - A x|| / ||b|| <= press_rtol) and BiCGStab for transport;
     * 1 the reference's KSP for both: GMRES(30) with left Jacobi, pressure
     * on the full reference-form system, stop test on the preconditioned
     * residual norm <= rtol ||D^-1 b||, each KSPSolve issued twice */
    int ksp = 0;

    size_t idx(int i, int j, int k) const {
        return (size_t)k + (size_t)nk * ((size_t)j + (size_t)nj * (size_t)i);
    }
    void coords(size_t v, int& i, int& j, int& k) const {
        k = (int)(v % nk);
        j = (int)((v / nk) % nj);
        i = (int)(v / ((size_t)nk * nj));
    }
    /* Reference neighbour order: +i, -i, +j, -j, +k, -k
     * (src/Simulation.cpp:771-776). Returns false if outside the grid. */
    bool nb(int i, int j, int k, int f, size_t& n) const {
        static const int di[6] = {1, -1, 0, 0, 0, 0};
        static const int dj[6] = {0, 0, 1, -1, 0, 0};
        static const int dk[6] = {0, 0, 0, 0, 1, -1};
        int a = i + di[f], b = j + dj[f], c = k + dk[f];
        if (a < 0 || a >= ni || b < 0 || b >= nj || c < 0 || c >= nk) return false;
        n = idx(a, b, c);
        return true;
    }
};

/* ------------------------------------------------------------------------ */
/* Properties: src/Simulation.cpp:1634-1664, 1596-1628, 187-248             */
/* ------------------------------------------------------------------------ */

static void update_rel_perm(lso_ctx& c) {
    const double m = 1.0 - (1.0 / c.P.vg_n);
    const double s_res = c.P.s_res;
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < (long long)c.N; ++v) {
        uint8_t t = c.img[v];
        if (t == PORE || t == SUL) {
            double Se = (c.s[v] - s_res) / (1.0 - s_res);
            Se = std::max(0.0, std::min(1.0, Se));
            if (Se <= 0.0) {
                c.krel[v] = 0.0;
            } else {
                double term1 = 1.0 - std::pow(1.0 - std::pow(Se, 1.0 / m), m);
                c.krel[v] = std::sqrt(Se) * term1 * term1;
            }
        } else {
            c.krel[v] = 0.0;
        }
    }
}

static void update_cap_pressure(lso_ctx& c) {
    const double m = 1.0 - (1.0 / c.P.vg_n);
    const double s_res = c.P.s_res;
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < (long long)c.N; ++v) {
        uint8_t t = c.img[v];
        if (t == PORE || t == SUL) {
            double Se = (c.s[v] - s_res) / (1.0 - s_res);
            Se = std::max(1e-6, std::min(1.0, Se));
            if (Se >= 1.0) {
                c.pc[v] = 0.0;
            } else {
                double Se_inv_m = std::pow(Se, -1.0 / m);
                c.pc[v] = (1.0 / c.P.vg_alpha) * std::pow(Se_inv_m - 1.0, 1.0 / c.P.vg_n);
            }
        } else {
            c.pc[v] = 0.0;
        }
    }
}

static void init_properties(lso_ctx& c) {
    const double pore_permeability = 1.0e-12;
    const double rock_permeability = 1.0e-18;
    const double log_k_rock = std::log10(rock_permeability);
    const double log_k_pore = std::log10(pore_permeability);
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < (long long)c.N; ++v) {
        uint8_t type = c.img[v];
        double K_intrinsic = rock_permeability;
        if (type == AIR) {
            K_intrinsic = pore_permeability;
        } else if (type == PORE) {
            double p_frac = c.inv[v];
            p_frac = std::max(0.0, std::min(1.0, p_frac));
            double log_k_eff = (p_frac * log_k_rock) + ((1.0 - p_frac) * log_k_pore);
            K_intrinsic = std::pow(10.0, log_k_eff);
        } else if (type == SUL) {
            double f = c.frac[v];
            double log_k_eff = (f * log_k_rock) + ((1.0 - f) * log_k_pore);
            K_intrinsic = std::pow(10.0, log_k_eff);
        }
        c.mu[v] = (type == AIR) ? 1.8e-5 : 1.0e-3;
        double K_rel = std::max(c.krel[v], 1e-15);
        c.perm[v] = K_intrinsic * K_rel;
    }
}

static void phase_properties(lso_ctx& c) {
    update_rel_perm(c);
    update_cap_pressure(c);
    init_properties(c);
}

/* ------------------------------------------------------------------------ */
/* Pressure: src/Simulation.cpp:658-798 (assembly), 804-855 (solve)          */
/* ------------------------------------------------------------------------ */

/* Surface Air voxel: any in-grid neighbour is not Air (:696-711). */
static bool air_is_surface(const lso_ctx& c, int i, int j, int k) {
    for (int f = 0; f < 6; ++f) {
        size_t n;
        if (c.nb(i, j, k, f, n) && c.img[n] != AIR) return true;
    }
    return false;
}

/* Build the reduced operator: for each Pore row the six (plus virtual)
 * couplings exactly as set_neighbor_link computes them, the diagonal sum, the
 * floored-row class, and the Dirichlet values of every non-solved node. */
static void assemble_pressure(lso_ctx& c, double t, std::vector<double>& pD,
                              std::vector<double>& b) {
    const double ptop = effective_top_pressure(c.P.top_pressure, t);
    const bool zf = c.P.bc_mode == LS_BC_ZFACES;
    for (int a = 0; a < 3; ++a) c.cf[a].assign(c.N, 0.0);
    c.cin.assign((size_t)c.nj * c.nk, 0.0);
    c.d.assign(c.N, 0.0);
    c.cls.assign(c.N, 0);
    pD.assign(c.N, 0.0);
    b.assign(c.N, 0.0);
    std::vector<double> src(c.N, 0.0);

    for (size_t v = 0; v < c.N; ++v) {
        int i, j, k;
        c.coords(v, i, j, k);
        uint8_t type = c.img[v];
        if (type == PORE) {
            double diagonal_term = 0.0;
            double source_term = 0.0;
            for (int f = 0; f < 6; ++f) {
                size_t n;
                if (!c.nb(i, j, k, f, n)) continue;
                double K_neighbor;
                if (c.img[n] == AIR) {
                    source_term -= c.P.evap_flux * c.P.dx;
                    K_neighbor = 1.0e-9;
                } else {
                    K_neighbor = c.perm[n];
                }
                double coeff = pore_link(c.perm[v], K_neighbor, c.mu[v]);
                diagonal_term -= coeff;
                /* store on the face: +faces at v, -faces at the neighbour */
                if (f == 0) c.cf[0][v] = coeff;
                else if (f == 1) c.cf[0][n] = coeff;
                else if (f == 2) c.cf[1][v] = coeff;
                else if (f == 3) c.cf[1][n] = coeff;
                else if (f == 4) c.cf[2][v] = coeff;
                else c.cf[2][n] = coeff;
            }
            if (zf && i == 0) {
                double coeff = pore_link(c.perm[v], 1.0e-9, c.mu[v]);
                diagonal_term -= coeff;
                c.cin[(size_t)j * c.nk + k] = coeff;
            }
            if (zf && i == c.ni - 1) {
                double coeff = pore_link(c.perm[v], 1.0e-9, c.mu[v]);
                diagonal_term -= coeff;
                c.cf[0][v] = coeff; /* outlet face, virtual node value 0 */
            }
            c.d[v] = -diagonal_term;
            src[v] = source_term;
            c.cls[v] = (std::fabs(diagonal_term) < 1e-16) ? 2 : 1;
        } else if (type == AIR) {
            pD[v] = air_is_surface(c, i, j, k) ? ptop : 0.0;
        } else {
            pD[v] = 0.0;
        }
    }
    /* floored rows keep their current (lagged) value as Dirichlet data */
    for (size_t v = 0; v < c.N; ++v)
        if (c.cls[v] == 2) pD[v] = c.p[v];

    /* right-hand side of the solved rows: -source + Dirichlet couplings */
    for (size_t v = 0; v < c.N; ++v) {
        if (c.cls[v] != 1) continue;
        int i, j, k;
        c.coords(v, i, j, k);
        double rhs = -src[v];
        for (int f = 0; f < 6; ++f) {
            size_t n;
            if (!c.nb(i, j, k, f, n)) continue;
            if (c.cls[n] == 1) continue;
            double cc = (f == 0) ? c.cf[0][v] : (f == 1) ? c.cf[0][n] : (f == 2) ? c.cf[1][v]
                      : (f == 3) ? c.cf[1][n] : (f == 4) ? c.cf[2][v] : c.cf[2][n];
            rhs += cc * pD[n];
        }
        if (zf && i == 0) rhs += c.cin[(size_t)j * c.nk + k] * ptop;
        b[v] = rhs;
    }
    (void)src;
}

/* y = A x over solved rows: d_v x_v - sum over solved neighbours c x_n. */
static void apply_reduced(const lso_ctx& c, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < (long long)c.N; ++v) {
        if (c.cls[v] != 1) { y[v] = 0.0; continue; }
        int i, j, k;
        c.coords((size_t)v, i, j, k);
        double acc = c.d[v] * x[v];
        for (int f = 0; f < 6; ++f) {
            size_t n;
            if (!c.nb(i, j, k, f, n)) continue;
            if (c.cls[n] != 1) continue;
            double cc = (f == 0) ? c.cf[0][v] : (f == 1) ? c.cf[0][n] : (f == 2) ? c.cf[1][v]
                      : (f == 3) ? c.cf[1][n] : (f == 4) ? c.cf[2][v] : c.cf[2][n];
            acc -= cc * x[n];
        }
        y[v] = acc;
    }
}

/* Jacobi-PCG on the reduced SPD system, stopping on the true relative
 * residual ||b - A x|| / ||b|| <= rtol (checked on the recurrence, then
 * confirmed on the recomputed residual). */
static int pcg_reduced(lso_ctx& c, const std::vector<double>& b, std::vector<double>& x,
                       double rtol, int maxit, int& iters, double& relres, double& bnorm) {
    const size_t N = c.N;
    bnorm = std::sqrt(dot(b, b));
    iters = 0;
    if (bnorm == 0.0) {
        for (size_t v = 0; v < N; ++v) if (c.cls[v] == 1) x[v] = 0.0;
        relres = 0.0;
        return 0;
    }
    std::vector<double> r(N), z(N), pvec(N), Ap(N);
    apply_reduced(c, x.data(), Ap.data());
    for (size_t v = 0; v < N; ++v) r[v] = (c.cls[v] == 1) ? b[v] - Ap[v] : 0.0;
    double rr = dot(r, r);
    relres = std::sqrt(rr) / bnorm;
    if (relres <= rtol) return 0;
    for (size_t v = 0; v < N; ++v) z[v] = (c.cls[v] == 1) ? r[v] / c.d[v] : 0.0;
    pvec = z;
    double rz = dot(r, z);
    while (iters < maxit) {
        apply_reduced(c, pvec.data(), Ap.data());
        double pAp = dot(pvec, Ap);
        double alpha = rz / pAp;
#pragma omp parallel for schedule(static)
        for (long long v = 0; v < (long long)N; ++v) {
            x[v] += alpha * pvec[v];
            r[v] -= alpha * Ap[v];
        }
        ++iters;
        rr = dot(r, r);
        relres = std::sqrt(rr) / bnorm;
        if (relres <= rtol) {
            apply_reduced(c, x.data(), Ap.data());
            double tr = 0.0;
            for (size_t v = 0; v < N; ++v)
                if (c.cls[v] == 1) { double e = b[v] - Ap[v]; tr += e * e; }
            relres = std::sqrt(tr) / bnorm;
            if (relres <= rtol) return 0;
            for (size_t v = 0; v < N; ++v) r[v] = (c.cls[v] == 1) ? b[v] - Ap[v] : 0.0;
        }
#pragma omp parallel for schedule(static)
        for (long long v = 0; v < (long long)N; ++v)
            z[v] = (c.cls[v] == 1) ? r[v] / c.d[v] : 0.0;
        double rz_new = dot(r, z);
        double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static)
        for (long long v = 0; v < (long long)N; ++v) pvec[v] = z[v] + beta * pvec[v];
    }
    return 1;
}

long long lso_pressure_system(lso_ctx* c, double t, int* rows, int* cols, double* vals,
                              double* rhs);

/* Restarted GMRES(m) with left Jacobi preconditioning: the algorithm of
 * PETSc's KSPGMRES + PCJACOBI that the reference configures
 * (src/Simulation.cpp:146-157; PETSc 3.20 defaults: restart 30, left
 * preconditioning, convergence when the preconditioned residual norm of the
 * Arnoldi recurrence falls to rtol ||D^-1 b||, maxit iterations).  The
 * Arnoldi vectors are orthogonalised by modified Gram-Schmidt (PETSc's
 * default is classical Gram-Schmidt; in exact arithmetic both give the same
 * iterates).  Warm start from x.  Returns 0 converged, 1 not. */
static int gmres_left_jacobi(const Csr& A, const std::vector<double>& b, std::vector<double>& x,
                             double rtol, int maxit, int m, int& iters, double& prec_relres) {
    const size_t N = b.size();
    std::vector<double> dinv(N);
    for (size_t i = 0; i < N; ++i) dinv[i] = A.diag[i] != 0.0 ? 1.0 / A.diag[i] : 1.0;
    std::vector<double> z(N), Ax(N);
    for (size_t i = 0; i < N; ++i) z[i] = dinv[i] * b[i];
    const double bnorm = std::sqrt(dot(z, z));
    const double ttol = std::max(rtol * bnorm, 1e-50);
    iters = 0;
    prec_relres = 0.0;
    if (bnorm == 0.0) {
        std::fill(x.begin(), x.end(), 0.0);
        return 0;
    }
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(N));
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    auto h = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
    while (true) {
        A.apply(x.data(), Ax.data());
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)N; ++i) V[0][i] = dinv[i] * (b[i] - Ax[i]);
        const double beta = std::sqrt(dot(V[0], V[0]));
        prec_relres = beta / bnorm;
        if (beta <= ttol) return 0;
        if (iters >= maxit) return 1;
        for (size_t i = 0; i < N; ++i) V[0][i] /= beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        bool done = false;
        for (; k < m && iters < maxit; ++k) {
            A.apply(V[k].data(), Ax.data());
            std::vector<double>& w = V[k + 1];
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)N; ++i) w[i] = dinv[i] * Ax[i];
            for (int i = 0; i <= k; ++i) {
                const double hij = dot(w, V[i]);
                h(i, k) = hij;
#pragma omp parallel for schedule(static)
                for (long long q = 0; q < (long long)N; ++q) w[q] -= hij * V[i][q];
            }
            const double hn = std::sqrt(dot(w, w));
            h(k + 1, k) = hn;
            if (hn > 0.0)
                for (size_t q = 0; q < N; ++q) w[q] /= hn;
            for (int i = 0; i < k; ++i) {          /* previous rotations */
                const double a = h(i, k), bb = h(i + 1, k);
                h(i, k) = cs[i] * a + sn[i] * bb;
                h(i + 1, k) = -sn[i] * a + cs[i] * bb;
            }
            const double a = h(k, k), bb = h(k + 1, k);
            const double r = std::hypot(a, bb);
            cs[k] = r > 0.0 ? a / r : 1.0;
            sn[k] = r > 0.0 ? bb / r : 0.0;
            h(k, k) = r;
            h(k + 1, k) = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            ++iters;
            prec_relres = std::fabs(g[k + 1]) / bnorm;
            if (std::fabs(g[k + 1]) <= ttol || hn == 0.0) {
                ++k;
                done = true;
                break;
            }
        }
        for (int i = k - 1; i >= 0; --i) {       /* back substitution */
            double acc = g[i];
            for (int j = i + 1; j < k; ++j) acc -= h(i, j) * y[j];
            y[i] = acc / h(i, i);
        }
        for (int j = 0; j < k; ++j) {
#pragma omp parallel for schedule(static)
            for (long long q = 0; q < (long long)N; ++q) x[q] += y[j] * V[j][q];
        }
        if (done) return 0;
        if (iters >= maxit) return 1;
    }
}

/* The reference issues each KSPSolve twice in a row (pressure
 * src/Simulation.cpp:823-825, conc :1262-1264): iterations summed, relres of
 * the last solve. */
static int ksp_twice(const Csr& A, const std::vector<double>& b, std::vector<double>& x,
                     double rtol, int maxit, int& iters, double& relres) {
    int it1 = 0, it2 = 0;
    double rr1 = 0.0, rr2 = 0.0;
    int st = gmres_left_jacobi(A, b, x, rtol, maxit, 30, it1, rr1);
    if (st == 0) st = gmres_left_jacobi(A, b, x, rtol, maxit, 30, it2, rr2);
    iters = it1 + it2;
    relres = st == 0 ? rr2 : rr1;
    return st;
}

/* solvePressure with the reference's KSP (src/Simulation.cpp:804-855): the
 * full reference-form system of every voxel, warm start = previous pressure,
 * KSPSolve called twice (:823, :825). */
static int phase_pressure_reference(lso_ctx& c, double t, ls_step_report* rep) {
    const long long nnz = lso_pressure_system(&c, t, nullptr, nullptr, nullptr, nullptr);
    std::vector<int> R(nnz), C(nnz);
    std::vector<double> V(nnz), b(c.N);
    lso_pressure_system(&c, t, R.data(), C.data(), V.data(), b.data());
    Csr A;
    A.build((int)c.N, R, C, V);
    std::vector<double> x = c.p;
    int its = 0;
    double rr = 0.0;
    const int st = ksp_twice(A, b, x, c.P.press_rtol, c.P.press_maxit, its, rr);
    c.p = x;
    const double ptop = effective_top_pressure(c.P.top_pressure, t);
    c.p_top_prev = ptop;
    c.have_prev = true;
    if (rep) {
        rep->press_iters = its;
        rep->press_outer = 2;
        rep->press_relres = rr;
        rep->top_pressure_eff = ptop;
        double mn = 1e300, mx = -1e300;
        for (size_t v = 0; v < c.N; ++v) { mn = std::min(mn, c.p[v]); mx = std::max(mx, c.p[v]); }
        rep->p_min = mn;
        rep->p_max = mx;
    }
    if (st != 0) {
        c.err = "pressure solve did not converge";
        return 1;
    }
    return 0;
}

static int phase_pressure(lso_ctx& c, double t, ls_step_report* rep) {
    if (c.ksp == 1) return phase_pressure_reference(c, t, rep);
    std::vector<double> pD, b;
    const double ptop = effective_top_pressure(c.P.top_pressure, t);
    assemble_pressure(c, t, pD, b);
    /* warm start: previous pressure (src/Simulation.cpp:157,808), whatever
     * warm_start asks (it changes only the iteration count, not the answer
     * beyond the tolerance) */
    std::vector<double> x(c.N, 0.0);
    for (size_t v = 0; v < c.N; ++v) if (c.cls[v] == 1) x[v] = c.p[v];
    int iters = 0;
    double relres = 0.0, bnorm = 0.0;
    int st = pcg_reduced(c, b, x, c.P.press_rtol, c.P.press_maxit, iters, relres, bnorm);
    long long nfl = 0;
    /* floored rows from their own reference row:
     * 1e-16 p_f + sum_n c p_n = source  (src/Simulation.cpp:783-797) */
    std::vector<double> pnew(c.N);
    for (size_t v = 0; v < c.N; ++v) pnew[v] = (c.cls[v] == 1) ? x[v] : pD[v];
    const bool zf = c.P.bc_mode == LS_BC_ZFACES;
    for (size_t v = 0; v < c.N; ++v) {
        if (c.cls[v] != 2) continue;
        ++nfl;
        int i, j, k;
        c.coords(v, i, j, k);
        double source_term = 0.0, sum = 0.0;
        for (int f = 0; f < 6; ++f) {
            size_t n;
            if (!c.nb(i, j, k, f, n)) continue;
            if (c.img[n] == AIR) source_term -= c.P.evap_flux * c.P.dx;
            double cc = (f == 0) ? c.cf[0][v] : (f == 1) ? c.cf[0][n] : (f == 2) ? c.cf[1][v]
                      : (f == 3) ? c.cf[1][n] : (f == 4) ? c.cf[2][v] : c.cf[2][n];
            sum += cc * pnew[n];
        }
        if (zf && i == 0) sum += c.cin[(size_t)j * c.nk + k] * ptop;
        pD[v] = (source_term - sum) / 1e-16;
    }
    for (size_t v = 0; v < c.N; ++v) c.p[v] = (c.cls[v] == 1) ? x[v] : pD[v];
    c.p_top_prev = ptop;
    c.have_prev = true;
    if (rep) {
        rep->press_iters = iters;
        rep->press_outer = 1;
        rep->press_relres = relres;
        rep->press_bnorm = bnorm;
        rep->n_floored = nfl;
        rep->top_pressure_eff = ptop;
        double mn = 1e300, mx = -1e300;
        for (size_t v = 0; v < c.N; ++v) { mn = std::min(mn, c.p[v]); mx = std::max(mx, c.p[v]); }
        rep->p_min = mn;
        rep->p_max = mx;
    }
    if (st != 0) {
        c.err = "pressure solve did not converge";
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Darcy flux: src/Simulation.cpp:866-953 (one-sided only at global faces)  */
/* ------------------------------------------------------------------------ */

static void phase_flux(lso_ctx& c) {
    const double dx = c.P.dx;
    const double mcg = c.P.max_cap_grad;
    const double rho_g = 1000.0 * 9.81;
    const int ni = c.ni, nj = c.nj, nk = c.nk;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                size_t v = c.idx(i, j, k);
                const double* P = c.p.data();
                const double* C = c.pc.data();
                auto grad = [&](const double* F, int pos, int ext, size_t stride) {
                    if (pos == 0) return (F[v + stride] - F[v]) / dx;
                    if (pos == ext - 1) return (F[v] - F[v - stride]) / dx;
                    return (F[v + stride] - F[v - stride]) / (2.0 * dx);
                };
                const size_t si = (size_t)nj * nk, sj = (size_t)nk;
                double gradP_x = grad(P, i, ni, si);
                double gradP_y = grad(P, j, nj, sj);
                double gradP_z = grad(P, k, nk, 1);
                double gradPc_x = grad(C, i, ni, si);
                double gradPc_y = grad(C, j, nj, sj);
                double gradPc_z = grad(C, k, nk, 1);
                gradPc_x = std::max(-mcg, std::min(mcg, gradPc_x));
                gradPc_y = std::max(-mcg, std::min(mcg, gradPc_y));
                gradPc_z = std::max(-mcg, std::min(mcg, gradPc_z));
                double transmissivity = c.perm[v] / c.mu[v];
                double liq_grad_x = gradP_x - gradPc_x;
                double liq_grad_y = gradP_y - gradPc_y;
                double liq_grad_z = gradP_z - gradPc_z;
                c.qx[v] = -transmissivity * (liq_grad_x - rho_g);
                c.qy[v] = -transmissivity * liq_grad_y;
                c.qz[v] = -transmissivity * liq_grad_z;
            }
}

/* ------------------------------------------------------------------------ */
/* Explicit saturation: src/Simulation.cpp:1672-1856                        */
/* ------------------------------------------------------------------------ */

static inline bool sat_inactive(uint8_t t) {
    return t == AIR || t == ROCK || t == SUL || t == PREC;
}

static void sat_rates(const lso_ctx& c, size_t v, double& div_q, double& evap_sink) {
    int i, j, k;
    c.coords(v, i, j, k);
    const double dx = c.P.dx;
    auto face = [&](const std::vector<double>& q, int f) {
        size_t n;
        return c.nb(i, j, k, f, n) ? (q[v] + q[n]) / 2.0 : 0.0;
    };
    double q_face_xp = face(c.qx, 0), q_face_xm = face(c.qx, 1);
    double q_face_yp = face(c.qy, 2), q_face_ym = face(c.qy, 3);
    double q_face_zp = face(c.qz, 4), q_face_zm = face(c.qz, 5);
    div_q = ((q_face_xp - q_face_xm) + (q_face_yp - q_face_ym) + (q_face_zp - q_face_zm)) / dx;
    evap_sink = 0.0;
    if (c.img[v] == PORE) {
        int air_faces = 0;
        for (int f = 0; f < 6; ++f) {
            size_t n;
            if (c.nb(i, j, k, f, n) && c.img[n] == AIR) air_faces++;
        }
        evap_sink = (air_faces * c.P.evap_flux) / dx;
    }
}

static void phase_saturation(lso_ctx& c, double dt, ls_step_report* rep) {
    double max_div_q = 0.0;
#pragma omp parallel for reduction(max : max_div_q) schedule(static)
    for (long long v = 0; v < (long long)c.N; ++v) {
        if (sat_inactive(c.img[v])) continue;
        double div_q, evap;
        sat_rates(c, (size_t)v, div_q, evap);
        double total_rate = std::abs(div_q) + evap;
        if (total_rate > max_div_q) max_div_q = total_rate;
    }
    double max_dsat = 0.1;
    double dt_safe = dt;
    if (max_div_q > 1e-20) dt_safe = (max_dsat * c.P.porosity) / max_div_q;
    double calculated_steps = std::ceil(dt / dt_safe);
    int num_steps = 1;
    if (std::isinf(calculated_steps) || std::isnan(calculated_steps)) num_steps = 1000;
    else if (calculated_steps > 1000) num_steps = 1000;
    else if (calculated_steps > 1.0) num_steps = (int)calculated_steps;
    double dt_sub = dt / num_steps;
    double dt_sub_over_phi = dt_sub / c.P.porosity;

    std::vector<double> s_new;
    for (int step = 0; step < num_steps; ++step) {
        s_new = c.s; /* inactive voxels keep their value */
#pragma omp parallel for schedule(static)
        for (long long v = 0; v < (long long)c.N; ++v) {
            if (sat_inactive(c.img[v])) continue;
            double div_q, evap_sink;
            sat_rates(c, (size_t)v, div_q, evap_sink);
            double new_sat = c.s[v] - dt_sub_over_phi * (div_q + evap_sink);
            s_new[v] = std::max(0.0, std::min(1.0, new_sat));
        }
        c.s.swap(s_new);
    }
    if (rep) {
        rep->sat_substeps = num_steps;
        rep->sat_max_divq = max_div_q;
    }
}

/* ------------------------------------------------------------------------ */
/* Surface effects: src/Simulation.cpp:2063-2103                            */
/* ------------------------------------------------------------------------ */

static void phase_surface(lso_ctx& c) {
    /* Order independent: only Pore -> Precipitate, tested against Air. */
    for (size_t v = 0; v < c.N; ++v) {
        if (c.img[v] == PORE && c.s[v] > 0.0) {
            int i, j, k;
            c.coords(v, i, j, k);
            bool is_surface_voxel = false;
            for (int f = 0; f < 6; ++f) {
                size_t n;
                if (c.nb(i, j, k, f, n) && c.img[n] == AIR) { is_surface_voxel = true; break; }
            }
            if (is_surface_voxel) {
                c.conc[v] = 1.0;
                c.img[v] = PREC;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Transport: src/Simulation.cpp:966-1237 (assembly), 1243-1295, 2167-2224   */
/* ------------------------------------------------------------------------ */

static void assemble_transport(lso_ctx& c, double dt, const std::vector<double>& field,
                               std::vector<int>& R, std::vector<int>& C,
                               std::vector<double>& V, std::vector<double>& b) {
    R.clear();
    C.clear();
    V.clear();
    b.assign(c.N, 0.0);
    const double inv_dt = 1.0 / dt;
    const double alpha = c.P.alpha_transfer;
    const double theta = c.P.theta;
    const double dx = c.P.dx;
    auto add = [&](size_t r, size_t col, double val) {
        R.push_back((int)r);
        C.push_back((int)col);
        V.push_back(val);
    };
    /* the reference loops k -> j -> i; the flux zeroing at Rock/Precipitate
     * voxels is visible to later voxels' -face reads, which only ever read a
     * Rock/Precipitate neighbour's flux through a link that ignores it
     * (Rock: exchange only; Precipitate: q := 0), so the order is immaterial */
    for (int kk = 0; kk < c.nk; ++kk)
        for (int jj = 0; jj < c.nj; ++jj)
            for (int ii = 0; ii < c.ni; ++ii) {
                size_t v = c.idx(ii, jj, kk);
                uint8_t voxel_type = c.img[v];
                if (voxel_type == PREC) {
                    c.qx[v] = 0.0;
                    c.qy[v] = 0.0;
                    c.qz[v] = 0.0;
                    add(v, v, 1.0);
                    b[v] += field[v];
                    continue;
                }
                if (voxel_type == ROCK) {
                    c.qx[v] = 0.0;
                    c.qy[v] = 0.0;
                    c.qz[v] = 0.0;
                    add(v, v, inv_dt);
                    b[v] += field[v] * inv_dt;
                    for (int f = 0; f < 6; ++f) {
                        size_t n;
                        if (c.nb(ii, jj, kk, f, n) && c.img[n] == PORE) {
                            add(v, v, alpha);
                            add(v, n, -alpha);
                        }
                    }
                    continue;
                }
                add(v, v, inv_dt);
                b[v] += field[v] * inv_dt;
                const size_t si = (size_t)c.nj * c.nk, sj = (size_t)c.nk;
                for (int f = 0; f < 6; ++f) {
                    size_t n;
                    if (!c.nb(ii, jj, kk, f, n)) continue;
                    double q_face;
                    switch (f) {
                        case 0: q_face = c.qx[v]; break;
                        case 1: q_face = -c.qx[v - si]; break;
                        case 2: q_face = c.qy[v]; break;
                        case 3: q_face = -c.qy[v - sj]; break;
                        case 4: q_face = c.qz[v]; break;
                        default: q_face = -c.qz[v - 1]; break;
                    }
                    uint8_t neigh_type = c.img[n];
                    if (neigh_type == ROCK) {
                        add(v, v, alpha);
                        add(v, n, -alpha);
                        continue;
                    }
                    double actual_q_face = q_face;
                    if (neigh_type == PREC || neigh_type == SUL) actual_q_face = 0.0;
                    double D_eff = (neigh_type == PORE) ? (c.P.D * c.P.Dpore_fac) : c.P.D;
                    double diffusion_coeff = -D_eff / (dx * dx);
                    double coeff_neighbor = diffusion_coeff;
                    double coeff_self = -diffusion_coeff;
                    double adv_coeff = actual_q_face / dx;
                    if (adv_coeff > 0) coeff_self += adv_coeff;
                    else coeff_neighbor -= adv_coeff;
                    add(v, n, theta * coeff_neighbor);
                    add(v, v, theta * coeff_self);
                    double explicit_rhs = -(coeff_self * field[v] + coeff_neighbor * field[n]);
                    b[v] += (1.0 - theta) * explicit_rhs;
                }
            }
}

/* BiCGStab with Jacobi (right) preconditioning; stops on the true relative
 * residual ||b - A x|| / ||b|| <= rtol. Warm start from x. */
static int bicgstab(const Csr& A, const std::vector<double>& b, std::vector<double>& x,
                    double rtol, int maxit, int& iters, double& relres) {
    const size_t N = b.size();
    iters = 0;
    double bnorm = std::sqrt(dot(b, b));
    if (bnorm == 0.0) {
        std::fill(x.begin(), x.end(), 0.0);
        relres = 0.0;
        return 0;
    }
    std::vector<double> r(N), r0(N), pv(N, 0.0), v(N, 0.0), s(N), t(N), ph(N), sh(N), tmp(N);
    auto true_res = [&]() {
        A.apply(x.data(), tmp.data());
        for (size_t i = 0; i < N; ++i) r[i] = b[i] - tmp[i];
        return std::sqrt(dot(r, r)) / bnorm;
    };
    relres = true_res();
    if (relres <= rtol) return 0;
    r0 = r;
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    while (iters < maxit) {
        double rho_new = dot(r0, r);
        if (rho_new == 0.0) { r0 = r; rho_new = dot(r0, r); pv.assign(N, 0.0); v.assign(N, 0.0); rho = alpha = omega = 1.0; }
        double beta = (rho_new / rho) * (alpha / omega);
        rho = rho_new;
        for (size_t i = 0; i < N; ++i) pv[i] = r[i] + beta * (pv[i] - omega * v[i]);
        for (size_t i = 0; i < N; ++i) ph[i] = pv[i] / A.diag[i];
        A.apply(ph.data(), v.data());
        alpha = rho / dot(r0, v);
        for (size_t i = 0; i < N; ++i) s[i] = r[i] - alpha * v[i];
        ++iters;
        if (std::sqrt(dot(s, s)) / bnorm <= rtol) {
            for (size_t i = 0; i < N; ++i) x[i] += alpha * ph[i];
            relres = true_res();
            if (relres <= rtol) return 0;
            continue;
        }
        for (size_t i = 0; i < N; ++i) sh[i] = s[i] / A.diag[i];
        A.apply(sh.data(), t.data());
        double tt = dot(t, t);
        omega = (tt > 0.0) ? dot(t, s) / tt : 0.0;
        for (size_t i = 0; i < N; ++i) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = s[i] - omega * t[i];
        }
        if (std::sqrt(dot(r, r)) / bnorm <= rtol) {
            relres = true_res();
            if (relres <= rtol) return 0;
        }
        if (omega == 0.0) { relres = true_res(); r0 = r; rho = alpha = omega = 1.0; pv.assign(N, 0.0); v.assign(N, 0.0); }
    }
    relres = true_res();
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Implicit saturation: setupSaturationEqns + solveSaturation               */
/* (src/Simulation.cpp:1858-2053).  Inactive voxels keep their value; Pore  */
/* rows: phi/dt s_new + upwinded mobility advection = phi/dt s - evap sink, */
/* the outward face flux being the mean of the two cells' flux and the      */
/* upwind divisor the assembly-time saturation (floored at 1e-6).           */
/* ------------------------------------------------------------------------ */

static void assemble_saturation(lso_ctx& c, double dt, std::vector<int>& R, std::vector<int>& C,
                                std::vector<double>& V, std::vector<double>& b) {
    R.clear();
    C.clear();
    V.clear();
    b.assign(c.N, 0.0);
    const double inv_dt = 1.0 / dt;
    const double dx = c.P.dx;
    auto add = [&](size_t r, size_t col, double val) {
        R.push_back((int)r);
        C.push_back((int)col);
        V.push_back(val);
    };
    for (int kk = 0; kk < c.nk; ++kk)
        for (int jj = 0; jj < c.nj; ++jj)
            for (int ii = 0; ii < c.ni; ++ii) {
                const size_t v = c.idx(ii, jj, kk);
                const uint8_t t = c.img[v];
                if (t == AIR || t == ROCK || t == SUL || t == PREC) {
                    add(v, v, 1.0);
                    b[v] += c.s[v];
                    continue;
                }
                int air_faces = 0;
                for (int f = 0; f < 6; ++f) {
                    size_t n;
                    if (c.nb(ii, jj, kk, f, n) && c.img[n] == AIR) ++air_faces;
                }
                const double evap_sink = (air_faces * c.P.evap_flux) / dx;
                double coeff_self = c.P.porosity * inv_dt;
                const double explicit_rhs = (c.P.porosity * inv_dt * c.s[v]) - evap_sink;
                for (int f = 0; f < 6; ++f) {      /* xp, xm, yp, ym, zp, zm */
                    size_t n;
                    if (!c.nb(ii, jj, kk, f, n)) continue;
                    const std::vector<double>& q = f < 2 ? c.qx : (f < 4 ? c.qy : c.qz);
                    const double q_out = (f % 2 == 0) ? (q[v] + q[n]) / 2.0 : -(q[v] + q[n]) / 2.0;
                    if (q_out > 0) {
                        const double v_eff = q_out / std::max(c.s[v], 1e-6);
                        coeff_self += v_eff / dx;
                    } else {
                        const double v_eff = q_out / std::max(c.s[n], 1e-6);
                        add(v, n, v_eff / dx);
                    }
                }
                add(v, v, coeff_self);
                b[v] += explicit_rhs;
            }
}

static int phase_saturation_implicit(lso_ctx& c, double dt, ls_step_report* rep) {
    std::vector<int> R, C;
    std::vector<double> V, b;
    assemble_saturation(c, dt, R, C, V, b);
    Csr A;
    A.build((int)c.N, R, C, V);
    int it = 0;
    double rr = 0.0;
    const int st = bicgstab(A, b, c.s, c.P.adr_rtol, c.P.adr_maxit, it, rr);  /* warm start: s */
    double lo = 1e300, hi = -1e300;
    for (size_t v = 0; v < c.N; ++v) {
        lo = std::min(lo, c.s[v]);
        hi = std::max(hi, c.s[v]);
    }
    const bool valid = !(hi > 1.05 || lo < -0.05);
    if (valid)
        for (size_t v = 0; v < c.N; ++v) c.s[v] = std::max(0.0, std::min(1.0, c.s[v]));
    c.sat_valid = valid;
    if (rep) {
        rep->sat_iters = it;
        rep->sat_valid = valid ? 1 : 0;
        rep->sat_raw_min = lo;
        rep->sat_raw_max = hi;
    }
    return st;
}

static int phase_transport(lso_ctx& c, double dt, ls_step_report* rep) {
    std::vector<int> R, C;
    std::vector<double> V, b;
    int it1 = 0, it2 = 0;
    double rr1 = 0.0, rr2 = 0.0;
    /* conc: setupConcentrationEqns(dt, conc, 0, 1) + solveConc */
    assemble_transport(c, dt, c.conc, R, C, V, b);
    Csr A;
    A.build((int)c.N, R, C, V);
    int st = c.ksp == 1 ? ksp_twice(A, b, c.conc, c.P.adr_rtol, c.P.adr_maxit, it1, rr1)
                        : bicgstab(A, b, c.conc, c.P.adr_rtol, c.P.adr_maxit, it1, rr1);
    if (st == 0) {
        /* acid: setupConcentrationEqns(dt, conc_acid, 1, 0) + solveAcid */
        assemble_transport(c, dt, c.acid, R, C, V, b);
        Csr A2;
        A2.build((int)c.N, R, C, V);
        st = c.ksp == 1 ? ksp_twice(A2, b, c.acid, c.P.adr_rtol, c.P.adr_maxit, it2, rr2)
                        : bicgstab(A2, b, c.acid, c.P.adr_rtol, c.P.adr_maxit, it2, rr2);
    }
    if (rep) {
        rep->adr_iters = std::max(it1, it2);
        rep->adr_iters_conc = it1;
        rep->adr_iters_acid = it2;
        rep->adr_relres = std::max(rr1, rr2);
        double a = 1e300, bb = -1e300, e = 1e300, f = -1e300;
        for (size_t v = 0; v < c.N; ++v) {
            a = std::min(a, c.conc[v]); bb = std::max(bb, c.conc[v]);
            e = std::min(e, c.acid[v]); f = std::max(f, c.acid[v]);
        }
        rep->conc_min = a; rep->conc_max = bb; rep->acid_min = e; rep->acid_max = f;
    }
    if (st != 0) {
        c.err = "transport solve did not converge";
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Precipitation: src/Simulation.cpp:2114-2165                              */
/* ------------------------------------------------------------------------ */

static void phase_precipitation(lso_ctx& c) {
    for (size_t v = 0; v < c.N; ++v) {
        if (c.img[v] == PORE && c.conc[v] > c.P.c_sat) {
            double excess_conc = c.conc[v] - c.P.c_sat;
            if (c.P.instant_precip) {
                c.img[v] = PREC;
            } else {
                double fluid_vol_fraction = c.P.porosity * c.s[v];
                double precipitated_vol_fraction = (excess_conc * fluid_vol_fraction) / c.P.solid_density;
                c.inv[v] += (precipitated_vol_fraction * c.P.time_scalar);
                double critical_limit = 0.90;
                if (c.inv[v] > critical_limit) c.img[v] = PREC;
            }
            c.conc[v] = c.P.c_sat;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Leaching: src/Simulation.cpp:315-617 (in place, reference loop order)   */
/* ------------------------------------------------------------------------ */

static int phase_leaching(lso_ctx& c, double dt, ls_step_report* rep) {
    const double dx = c.P.dx;
    const double k_trans = c.P.D / dx;
    const double k_eff = 1.0 / ((1.0 / k_trans) + (1.0 / c.P.kreac));
    auto surface_flux = [&](int i, int j, int k, int f) {
        size_t n;
        if (c.nb(i, j, k, f, n)) {
            uint8_t n_type = c.img[n];
            if (n_type == PORE || n_type == ROCK) return k_eff * c.acid[n];
        }
        return 0.0;
    };
    double max_rate = 0.0;
    for (size_t v = 0; v < c.N; ++v) {
        if (c.img[v] != SUL) continue;
        int i, j, k;
        c.coords(v, i, j, k);
        double flux_sum = 0.0;
        for (int f = 0; f < 6; ++f) flux_sum += surface_flux(i, j, k, f);
        double area_per_face = dx * dx;
        double voxel_volume = dx * dx * dx;
        double vol_rate = flux_sum * area_per_face;
        double fractional_rate = vol_rate / voxel_volume;
        if (fractional_rate > max_rate) max_rate = fractional_rate;
    }
    double effective_fractional_rate = max_rate * c.P.time_scalar;
    double dt_safe = dt;
    if (effective_fractional_rate > 1e-20) dt_safe = 0.1 / effective_fractional_rate;
    double calculated_steps = std::ceil(dt / dt_safe);
    int num_steps = 1;
    if (calculated_steps > 1000) num_steps = 1000;
    else if (calculated_steps > 1.0) num_steps = (int)calculated_steps;
    double dt_sub = dt / num_steps;

    int leached_locally = 0;
    for (int step = 0; step < num_steps; ++step) {
        for (int k = 0; k < c.nk; ++k)
            for (int j = 0; j < c.nj; ++j)
                for (int i = 0; i < c.ni; ++i) {
                    size_t v = c.idx(i, j, k);
                    if (!(c.img[v] >= SUL && c.frac[v] > 0.0)) continue;
                    double f_xp = surface_flux(i, j, k, 0);
                    double f_xm = surface_flux(i, j, k, 1);
                    double f_yp = surface_flux(i, j, k, 2);
                    double f_ym = surface_flux(i, j, k, 3);
                    double f_zp = surface_flux(i, j, k, 4);
                    double f_zm = surface_flux(i, j, k, 5);
                    double flux_sum = f_xp + f_xm + f_yp + f_ym + f_zp + f_zm;
                    double area_per_face = dx * dx;
                    double voxel_volume = dx * dx * dx;
                    double mass_loss = flux_sum * area_per_face * dt_sub;
                    double fractional_loss = mass_loss / voxel_volume;
                    c.frac[v] -= (fractional_loss * c.P.time_scalar);
                    if (flux_sum > 0.0) {
                        const double fs[6] = {f_xp, f_xm, f_yp, f_ym, f_zp, f_zm};
                        for (int f = 0; f < 6; ++f) {
                            size_t n;
                            if (c.nb(i, j, k, f, n)) c.acid[n] -= fractional_loss * (fs[f] / flux_sum);
                        }
                    }
                    if (c.frac[v] <= 1e-4) {
                        c.frac[v] = 0.0;
                        if (c.img[v] == SUL) {
                            c.img[v] = PORE;
                            c.inv[v] = 0.0;
                            c.conc[v] = c.P.c_sat;
                            leached_locally++;
                        }
                    }
                }
    }
    if (rep) {
        rep->leach_substeps = num_steps;
        rep->leach_max_rate = max_rate;
        rep->leached = leached_locally;
    }
    return leached_locally;
}

/* ------------------------------------------------------------------------ */
/* Snapshot / rollback: src/Simulation.cpp:2312-2353                        */
/* ------------------------------------------------------------------------ */

static void snapshot(lso_ctx& c) {
    c.b_img = c.img; c.b_frac = c.frac; c.b_inv = c.inv; c.b_s = c.s; c.b_p = c.p;
    c.b_conc = c.conc; c.b_acid = c.acid; c.b_qx = c.qx; c.b_qy = c.qy; c.b_qz = c.qz;
}

static void rollback(lso_ctx& c) {
    c.img = c.b_img; c.frac = c.b_frac; c.inv = c.b_inv; c.s = c.b_s; c.p = c.b_p;
    c.conc = c.b_conc; c.acid = c.b_acid; c.qx = c.b_qx; c.qy = c.b_qy; c.qz = c.b_qz;
    phase_properties(c);
}

static void summarize(lso_ctx& c, ls_step_report* rep) {
    if (!rep) return;
    double mc = 0.0, ma = 0.0, mw = 0.0;
    for (size_t v = 0; v < c.N; ++v) { mc += c.conc[v]; ma += c.acid[v]; mw += c.s[v]; }
    rep->mass_conc = mc;
    rep->mass_acid = ma;
    rep->mass_water = mw * c.P.porosity;
}

/* ------------------------------------------------------------------------ */
/* C API                                                                     */
/* ------------------------------------------------------------------------ */

extern "C" {

int lso_set_ksp(lso_ctx* c, int which) {
    if (which != 0 && which != 1) return -1;
    c->ksp = which;
    return 0;
}

void lso_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int lso_create(lso_ctx** out, int ni, int nj, int nk, const ls_params* params) {
    if (!out || !params || ni < 2 || nj < 2 || nk < 2) return -1;
    lso_ctx* c = new lso_ctx();
    c->ni = ni; c->nj = nj; c->nk = nk;
    c->N = (size_t)ni * nj * nk;
    c->P = *params;
    const size_t N = c->N;
    c->img.assign(N, ROCK);
    c->frac.assign(N, 0.0f);
    for (auto* f : {&c->inv, &c->s, &c->p, &c->conc, &c->acid, &c->qx, &c->qy, &c->qz,
                    &c->krel, &c->pc, &c->perm, &c->mu})
        f->assign(N, 0.0);
    *out = c;
    return 0;
}

void lso_destroy(lso_ctx* c) { delete c; }
const char* lso_last_error(const lso_ctx* c) { return c ? c->err.c_str() : "null context"; }
int lso_set_params(lso_ctx* c, const ls_params* p) { c->P = *p; return 0; }

static std::vector<double>* dfield(lso_ctx* c, int f) {
    switch (f) {
        case LS_F_INVENTORY: return &c->inv;
        case LS_F_SATURATION: return &c->s;
        case LS_F_PRESSURE: return &c->p;
        case LS_F_CONC: return &c->conc;
        case LS_F_ACID: return &c->acid;
        case LS_F_QX: return &c->qx;
        case LS_F_QY: return &c->qy;
        case LS_F_QZ: return &c->qz;
        case LS_F_KREL: return &c->krel;
        case LS_F_PC: return &c->pc;
        case LS_F_PERM: return &c->perm;
        case LS_F_MU: return &c->mu;
        default: return nullptr;
    }
}

int lso_upload(lso_ctx* c, int field, const void* host, size_t eb) {
    if (field == LS_F_IMG) {
        for (size_t v = 0; v < c->N; ++v) {
            unsigned lab = (eb == 1) ? ((const uint8_t*)host)[v] : ((const uint16_t*)host)[v];
            if (lab < 1 || lab > 5) { c->err = "label outside 1..5"; return -2; }
            c->img[v] = (uint8_t)lab;
        }
        return 0;
    }
    if (field == LS_F_FRAC) {
        if (eb != 4) return -3;
        std::memcpy(c->frac.data(), host, c->N * 4);
        return 0;
    }
    auto* f = dfield(c, field);
    if (!f || eb != 8 || field >= LS_F_KREL) return -3;
    std::memcpy(f->data(), host, c->N * 8);
    return 0;
}

int lso_download(lso_ctx* c, int field, void* host, size_t eb) {
    if (field == LS_F_IMG) {
        for (size_t v = 0; v < c->N; ++v) {
            if (eb == 1) ((uint8_t*)host)[v] = c->img[v];
            else ((uint16_t*)host)[v] = c->img[v];
        }
        return 0;
    }
    if (field == LS_F_FRAC) {
        std::memcpy(host, c->frac.data(), c->N * 4);
        return 0;
    }
    if (field >= LS_F_KREL) phase_properties(*c); /* derived: from current state */
    auto* f = dfield(c, field);
    if (!f || eb != 8) return -3;
    std::memcpy(host, f->data(), c->N * 8);
    return 0;
}

int lso_init_state(lso_ctx* c, long long* nsul) {
    long long cnt = 0;
    for (size_t v = 0; v < c->N; ++v) {
        uint8_t t = c->img[v];
        c->p[v] = c->conc[v] = c->acid[v] = 0.0;
        c->qx[v] = c->qy[v] = c->qz[v] = 0.0;
        c->inv[v] = 0.0;
        c->krel[v] = c->pc[v] = 0.0;
        c->frac[v] = (t == SUL) ? 1.0f : 0.0f;          /* initFrac   :254-281 */
        if (t == SUL) cnt++;
        c->s[v] = (t == PORE || t == SUL) ? 0.1 : 0.0;  /* initSaturation :1573-1590 */
    }
    c->have_prev = false;
    phase_properties(*c);
    if (nsul) *nsul = cnt;
    return 0;
}

int lso_phase(lso_ctx* c, int phase, double t, double dt, ls_step_report* rep) {
    switch (phase) {
        case LS_PH_PROPERTIES: phase_properties(*c); return 0;
        case LS_PH_PRESSURE: return phase_pressure(*c, t, rep);
        case LS_PH_FLUX: phase_flux(*c); return 0;
        case LS_PH_SATURATION:
            if (c->P.implicit_sat) return phase_saturation_implicit(*c, dt, rep);
            phase_saturation(*c, dt, rep);
            return 0;
        case LS_PH_SURFACE: phase_surface(*c); return 0;
        case LS_PH_TRANSPORT: return phase_transport(*c, dt, rep);
        case LS_PH_PRECIPITATION: phase_precipitation(*c); return 0;
        case LS_PH_LEACHING: phase_leaching(*c, dt, rep); return 0;
        default: return -1;
    }
}

int lso_step(lso_ctx* c, double t, double dt, ls_step_report* rep) {
    if (rep) std::memset(rep, 0, sizeof(*rep));
    snapshot(*c);
    int st = 0;
    /* hydrodynamics: 1 Picard pass, or 3 with the implicit saturation
     * (src/main.cpp:312-370); an unphysical final pass forces a retry */
    const int passes = c->P.implicit_sat ? 3 : 1;
    for (int pass = 0; pass < passes && st == 0; ++pass) {
        phase_properties(*c);
        st = phase_pressure(*c, t, rep);
        if (st != 0) break;
        phase_flux(*c);
        if (c->P.implicit_sat) st = phase_saturation_implicit(*c, dt, rep);
        else phase_saturation(*c, dt, rep);
    }
    if (st == 0 && c->P.implicit_sat && !c->sat_valid) {
        c->err = "implicit saturation: unphysical state after the Picard passes";
        st = 1;
    }
    if (st == 0) {
        phase_surface(*c);
        st = phase_transport(*c, dt, rep);
    }
    if (st == 0) {
        phase_precipitation(*c);
        phase_leaching(*c, dt, rep);
    }
    summarize(*c, rep);
    if (rep) rep->status = st;
    return st;
}

int lso_snapshot(lso_ctx* c) { snapshot(*c); return 0; }
int lso_rollback(lso_ctx* c) { rollback(*c); return 0; }

long long lso_pressure_system(lso_ctx* c, double t, int* rows, int* cols, double* vals,
                              double* rhs) {
    /* the reference's rows, src/Simulation.cpp:683-797 */
    const double ptop = effective_top_pressure(c->P.top_pressure, t);
    const bool zf = c->P.bc_mode == LS_BC_ZFACES;
    long long e = 0;
    auto put = [&](size_t r, size_t col, double v) {
        if (rows) { rows[e] = (int)r; cols[e] = (int)col; vals[e] = v; }
        ++e;
    };
    for (size_t v = 0; v < c->N; ++v) {
        int i, j, k;
        c->coords(v, i, j, k);
        uint8_t type = c->img[v];
        if (type != PORE) {
            put(v, v, 1.0);
            if (rhs) rhs[v] = (type == AIR && air_is_surface(*c, i, j, k)) ? ptop : 0.0;
            continue;
        }
        double diagonal_term = 0.0, source_term = 0.0;
        for (int f = 0; f < 6; ++f) {
            size_t n;
            if (!c->nb(i, j, k, f, n)) continue;
            double K_neighbor;
            if (c->img[n] == AIR) { source_term -= c->P.evap_flux * c->P.dx; K_neighbor = 1.0e-9; }
            else K_neighbor = c->perm[n];
            double coeff = pore_link(c->perm[v], K_neighbor, c->mu[v]);
            put(v, n, coeff);
            diagonal_term -= coeff;
        }
        if (zf && i == 0) {
            double coeff = pore_link(c->perm[v], 1.0e-9, c->mu[v]);
            diagonal_term -= coeff;
            source_term -= coeff * ptop;
        }
        if (zf && i == c->ni - 1) diagonal_term -= pore_link(c->perm[v], 1.0e-9, c->mu[v]);
        if (std::fabs(diagonal_term) < 1e-16) diagonal_term = 1e-16;
        put(v, v, diagonal_term);
        if (rhs) rhs[v] = source_term;
    }
    return e;
}

long long lso_saturation_system(lso_ctx* c, double dt, int* rows, int* cols, double* vals,
                                double* rhs) {
    std::vector<int> R, C;
    std::vector<double> V, b;
    assemble_saturation(*c, dt, R, C, V, b);
    if (rows) {
        for (size_t e = 0; e < R.size(); ++e) { rows[e] = R[e]; cols[e] = C[e]; vals[e] = V[e]; }
        std::memcpy(rhs, b.data(), c->N * 8);
    }
    return (long long)R.size();
}

long long lso_transport_system(lso_ctx* c, double dt, int field, int* rows, int* cols,
                               double* vals, double* rhs) {
    std::vector<int> R, C;
    std::vector<double> V, b;
    assemble_transport(*c, dt, field == LS_F_ACID ? c->acid : c->conc, R, C, V, b);
    if (rows) {
        for (size_t e = 0; e < R.size(); ++e) { rows[e] = R[e]; cols[e] = C[e]; vals[e] = V[e]; }
        std::memcpy(rhs, b.data(), c->N * 8);
    }
    return (long long)R.size();
}

int lso_apply_pressure(lso_ctx* c, const double* x, double* y) {
    if (c->cls.size() != c->N) return -1;
    apply_reduced(*c, x, y);
    return 0;
}

int lso_pressure_classes(lso_ctx* c, int8_t* out) {
    if (c->cls.size() != c->N) return -1;
    std::memcpy(out, c->cls.data(), c->N);
    return 0;
}

}  // extern "C"
