// k_pressure.hip — the Darcy pressure solve (replaces KSPSolve in
// src/Simulation.cpp:804-855) on the reduced SPD system.
//
// The Pore rows with |diagonal| >= 1e-16 form the solved set R.  Jacobi-PCG
// on A is run as CG on the symmetrically scaled operator
//     A^ = D^-1/2 A D^-1/2,  A^ x = x - sum_faces c^_f x_n,
// whose faces c^ are zero wherever either side is not in R.  Vectors are zero
// outside R, so the operator needs no mask byte: y = A^ x reads x and the
// three +face arrays and writes y — 2*s_v + 3*s_c bytes per voxel.
//
// Kernels per CG iteration: stencil + p.Ap, finalize(alpha), update x, r +
// r.r, finalize(beta, stop test), direction update.  All scalars stay on the
// device; every kernel returns at once when the `done` flag is set, so the
// host launches iterations in batches and polls once per batch.
#include <cmath>

#include "ls_context.h"

namespace ls {

struct SGeo {
    int ni, nj, nk;
    uint32_t si, sj;
};

static SGeo sgeo(const ls_ctx* c) {
    SGeo g;
    g.ni = c->ni; g.nj = c->nj; g.nk = c->nk;
    g.sj = (uint32_t)c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    return g;
}

static dim3 stencil_grid(const ls_ctx* c) {
    return dim3((c->nk + TK - 1) / TK, (c->nj + TJ - 1) / TJ, (c->ni + ICHUNK - 1) / ICHUNK);
}

// y = A^ x, marching ICHUNK planes along i per thread.  Optionally also
// y += 0 and the block partial of x.y (DOT) and the residual form
// r = b - A^ x (RES).  MODE 0: y = A^x, dot x.y.   MODE 1: r = b - A^x, p = r,
// partials (r.r, sum (sd r)^2, sum (sd b)^2).
template <typename T, int MODE>
__global__ __launch_bounds__(TK * TJ) void k_stencil(
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ cx,
    const T* __restrict__ cy, const T* __restrict__ cz, SGeo g, double* __restrict__ partials,
    const int* __restrict__ done, const double* __restrict__ b, T* __restrict__ p2,
    const double* __restrict__ sd) {
    if (done != nullptr && *done) return;
    __shared__ double lds[3 * (TK * TJ) / WAVE];
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nb = nbx * nby * gridDim.z;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int t = xcd_remap(hw, nb);
    const int bx = t % nbx;
    const int rest = t / nbx;
    const int by = rest % nby;
    const int bz = rest / nby;
    const int k = bx * TK + threadIdx.x;
    const int j = by * TJ + threadIdx.y;
    const int i0 = bz * ICHUNK;
    const int i1 = min(i0 + ICHUNK, g.ni);
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (k < g.nk && j < g.nj) {
        uint32_t v = (uint32_t)i0 * g.si + (uint32_t)j * g.sj + (uint32_t)k;
        const bool hj = j > 0, hk = k > 0, pj = j + 1 < g.nj, pk = k + 1 < g.nk;
        T xm = (i0 > 0) ? x[v - g.si] : T(0);
        T cxm = (i0 > 0) ? cx[v - g.si] : T(0);
        T xc = x[v];
        for (int i = i0; i < i1; ++i, v += g.si) {
            const T xp = (i + 1 < g.ni) ? x[v + g.si] : T(0);
            const T cxp = cx[v];
            const T cyp = cy[v];
            const T czp = cz[v];
            const T cym = hj ? cy[v - g.sj] : T(0);
            const T czm = hk ? cz[v - 1] : T(0);
            const T xjp = pj ? x[v + g.sj] : T(0);
            const T xjm = hj ? x[v - g.sj] : T(0);
            const T xkp = pk ? x[v + 1] : T(0);
            const T xkm = hk ? x[v - 1] : T(0);
            const T off = cxp * xp + cxm * xm + cyp * xjp + cym * xjm + czp * xkp + czm * xkm;
            const T ax = xc - off;
            if (MODE == 0) {
                y[v] = ax;
                acc0 += (double)xc * (double)ax;
            } else {
                const double bv = b[v];
                const double r = bv - (double)ax;
                y[v] = (T)r;
                p2[v] = (T)r;
                const double s = sd[v];
                acc0 += r * r;
                acc1 += (s * r) * (s * r);
                acc2 += (s * bv) * (s * bv);
            }
            xm = xc;
            xc = xp;
            cxm = cxp;
        }
    }
    if (MODE == 0) {
        double a[1] = {acc0};
        block_reduce<1, RED_SUM>(a, lds);
        if (flat_tid() == 0) partials[t] = a[0];
    } else {
        double a[3] = {acc0, acc1, acc2};
        block_reduce<3, RED_SUM, RED_SUM, RED_SUM>(a, lds);
        if (flat_tid() == 0) {
            partials[t] = a[0];
            partials[nb + t] = a[1];
            partials[2 * nb + t] = a[2];
        }
    }
}


// ---------------------------------------------------------------------------
// Vectorised kernels (used when nk is a multiple of the vector width W:
// 4 floats / 2 doubles = 16 B per lane and array).  Each lane owns W
// consecutive k values of one (j) row and marches ICH planes along i; k
// neighbours across lanes come from wave shuffles, the lanes at the wave
// edges load their single halo value.  Faces are read with non-temporal
// loads and y is written with non-temporal stores (streamed once).
// ---------------------------------------------------------------------------
template <typename T, int W> struct VecOf;
template <> struct VecOf<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecOf<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename T> struct StencilShape;
template <> struct StencilShape<float> { static constexpr int W = 4, ICH = 4; };
template <> struct StencilShape<double> { static constexpr int W = 2, ICH = 8; };

template <typename T>
__global__ __launch_bounds__(WAVE * TJ) void k_stencil_vec(
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ cx,
    const T* __restrict__ cy, const T* __restrict__ cz, SGeo g, double* __restrict__ partials,
    const int* __restrict__ done) {
    constexpr int W = StencilShape<T>::W, ICH = StencilShape<T>::ICH;
    typedef typename VecOf<T, W>::type V;
    if (done != nullptr && *done) return;
    __shared__ double lds[TJ];
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nb = nbx * nby * gridDim.z;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int t = xcd_remap(hw, nb);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int lane = threadIdx.x;
    const int k0 = (bx * WAVE + lane) * W;
    const int j = by * TJ + threadIdx.y;
    const int i0 = bz * ICH, i1 = min(i0 + ICH, g.ni);
    double acc = 0.0;
    // rows past nj idle as whole waves (threadIdx.y is uniform per wave)
    if (j < g.nj) {
        const bool ok = k0 < g.nk;
        const bool hj = j > 0, pj = j + 1 < g.nj;
        uint32_t v = (uint32_t)i0 * g.si + (uint32_t)j * g.sj + (ok ? (uint32_t)k0 : 0u);
        const V zero = V(0);
        V xm = (ok && i0 > 0) ? *(const V*)(x + v - g.si) : zero;
        V cxm = (ok && i0 > 0) ? *(const V*)(cx + v - g.si) : zero;
        V xc = ok ? *(const V*)(x + v) : zero;
        for (int i = i0; i < i1; ++i, v += g.si) {
            V xp = zero, cxp = zero, cyp = zero, czp = zero, cym = zero, xjp = zero, xjm = zero;
            if (ok) {
                if (i + 1 < g.ni) xp = *(const V*)(x + v + g.si);
                cxp = __builtin_nontemporal_load((const V*)(cx + v));
                cyp = __builtin_nontemporal_load((const V*)(cy + v));
                czp = __builtin_nontemporal_load((const V*)(cz + v));
                if (hj) {
                    cym = *(const V*)(cy + v - g.sj);
                    xjm = *(const V*)(x + v - g.sj);
                }
                if (pj) xjp = *(const V*)(x + v + g.sj);
            }
            T left = __shfl_up(xc[W - 1], 1, WAVE);
            T czl = __shfl_up(czp[W - 1], 1, WAVE);
            T right = __shfl_down(xc[0], 1, WAVE);
            if (lane == 0) {
                left = (ok && k0 > 0) ? x[v - 1] : T(0);
                czl = (ok && k0 > 0) ? cz[v - 1] : T(0);
            }
            if (lane == WAVE - 1) right = (ok && k0 + W < g.nk) ? x[v + W] : T(0);
            V ax;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const T xkm = (w == 0) ? left : xc[w - 1];
                const T xkp = (w == W - 1) ? right : xc[w + 1];
                const T czm = (w == 0) ? czl : czp[w - 1];
                ax[w] = xc[w] - (cxp[w] * xp[w] + cxm[w] * xm[w] + cyp[w] * xjp[w] +
                                 cym[w] * xjm[w] + czp[w] * xkp + czm * xkm);
                acc += (double)xc[w] * (double)ax[w];
            }
            if (ok) __builtin_nontemporal_store(ax, (V*)(y + v));
            xm = xc;
            xc = xp;
            cxm = cxp;
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) lds[threadIdx.y] = acc;
    __syncthreads();
    if (flat_tid() == 0) {
        double s = 0.0;
        for (int q = 0; q < TJ; ++q) s += lds[q];
        partials[t] = s;
    }
}

// Vectorised CG vector updates: U vectors of 16 B per thread, all loads
// issued before the stores, non-temporal (each vector is streamed once per
// kernel).  The scalar tail (N % W) is handled by block 0.
constexpr int VU = 4;

template <typename T>
__global__ __launch_bounds__(EW_BLOCK) void k_cg_update_vec(
    T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ ap,
    const double* __restrict__ sd, const CgScalars* __restrict__ cs, uint32_t N,
    double* __restrict__ partials) {
    if (cs->done) return;
    constexpr int W = StencilShape<T>::W;
    typedef typename VecOf<T, W>::type V;
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    const T a = (T)cs->alpha;
    const uint32_t nv = N / W;
    const uint32_t base = blockIdx.x * (uint32_t)(EW_BLOCK * VU) + threadIdx.x;
    V xv[VU], rv[VU], pv[VU], av[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) {
            xv[u] = __builtin_nontemporal_load((const V*)x + q);
            rv[u] = __builtin_nontemporal_load((const V*)r + q);
            pv[u] = __builtin_nontemporal_load((const V*)p + q);
            av[u] = __builtin_nontemporal_load((const V*)ap + q);
        }
    }
    double rr = 0.0, wrr = 0.0;
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) {
            const V rn = rv[u] - a * av[u];
            const V xn = xv[u] + a * pv[u];
            __builtin_nontemporal_store(xn, (V*)x + q);
            __builtin_nontemporal_store(rn, (V*)r + q);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const double rd = (double)rn[w];
                rr += rd * rd;
                if (sd != nullptr) {
                    const double s = sd[q * W + w] * rd;
                    wrr += s * s;
                }
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < N - nv * W) {  // scalar tail
        const uint32_t v = nv * W + threadIdx.x;
        const T rn = r[v] - a * ap[v];
        x[v] = x[v] + a * p[v];
        r[v] = rn;
        const double rd = (double)rn;
        rr += rd * rd;
        if (sd != nullptr) { const double s = sd[v] * rd; wrr += s * s; }
    }
    double q2[2] = {rr, wrr};
    block_reduce<2, RED_SUM, RED_SUM>(q2, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q2[0];
        partials[gridDim.x + blockIdx.x] = q2[1];
    }
}

template <typename T>
__global__ __launch_bounds__(EW_BLOCK) void k_cg_direction_vec(
    T* __restrict__ p, const T* __restrict__ r, const CgScalars* __restrict__ cs, uint32_t N) {
    if (cs->done) return;
    constexpr int W = StencilShape<T>::W;
    typedef typename VecOf<T, W>::type V;
    const T beta = (T)cs->beta;
    const uint32_t nv = N / W;
    const uint32_t base = blockIdx.x * (uint32_t)(EW_BLOCK * VU) + threadIdx.x;
    V pv[VU], rv[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) {
            pv[u] = __builtin_nontemporal_load((const V*)p + q);
            rv[u] = __builtin_nontemporal_load((const V*)r + q);
        }
    }
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) __builtin_nontemporal_store(rv[u] + beta * pv[u], (V*)p + q);
    }
    if (blockIdx.x == 0 && threadIdx.x < N - nv * W) {
        const uint32_t v = nv * W + threadIdx.x;
        p[v] = r[v] + beta * p[v];
    }
}

template <typename T>
static int vec_blocks(uint32_t N) {
    const uint32_t nv = N / StencilShape<T>::W;
    uint32_t b = (nv + EW_BLOCK * VU - 1) / (EW_BLOCK * VU);
    return (int)(b < 1 ? 1 : b);
}

template <typename T>
static dim3 vec_grid(const ls_ctx* c) {
    constexpr int W = StencilShape<T>::W, ICH = StencilShape<T>::ICH;
    return dim3((c->nk / W + WAVE - 1) / WAVE, (c->nj + TJ - 1) / TJ, (c->ni + ICH - 1) / ICH);
}

// The vectorised path needs nk % W == 0 and its block partials to fit.
template <typename T>
static bool use_vec(const ls_ctx* c) {
    if (c->nk % StencilShape<T>::W != 0) return false;
    const dim3 g = vec_grid<T>(c);
    return (size_t)g.x * g.y * g.z <= (size_t)MAX_PARTIALS &&
           2 * (size_t)vec_blocks<T>((uint32_t)c->N) <= (size_t)MAX_PARTIALS;
}

// alpha = r.r / p.Ap
__global__ void k_cg_alpha(const double* __restrict__ partials, int n, CgScalars* cs) {
    if (cs->done) return;
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        double pAp = a[0];
        cs->pAp = pAp;
        if (!(pAp > 0.0) || !isfinite(pAp)) {  // breakdown: not SPD or NaN
            cs->done = 1;
            cs->converged = 0;
            cs->alpha = 0.0;
        } else {
            cs->alpha = cs->rr / pAp;
        }
    }
}

// x += alpha p, r -= alpha Ap, partials of r.r and (when sd != null) the
// unscaled residual norm sum (sd r)^2.
template <typename T>
__global__ void k_cg_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                            const T* __restrict__ ap, const double* __restrict__ sd,
                            const CgScalars* __restrict__ cs, uint32_t N,
                            double* __restrict__ partials) {
    if (cs->done) return;
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    const double alpha = cs->alpha;
    const T a = (T)alpha;
    double rr = 0.0, wrr = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const T pv = p[v];
        const T rv = r[v] - a * ap[v];
        x[v] = x[v] + a * pv;
        r[v] = rv;
        const double rd = (double)rv;
        rr += rd * rd;
        if (sd != nullptr) {
            const double w = sd[v] * rd;
            wrr += w * w;
        }
    }
    double q[2] = {rr, wrr};
    block_reduce<2, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q[0];
        partials[gridDim.x + blockIdx.x] = q[1];
    }
}

// beta = r.r(new) / r.r(old); stop test on the unscaled residual (outer FP64
// solves, use_w = 1) or on the scaled residual (inner FP32 solves).
__global__ void k_cg_beta(const double* __restrict__ partials, int n, CgScalars* cs, int use_w) {
    if (cs->done) return;
    __shared__ double lds[2 * 1024 / WAVE];
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        a0 += partials[i];
        a1 += partials[n + i];
    }
    double a[2] = {a0, a1};
    block_reduce<2, RED_SUM, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        const double rr_new = a[0];
        const double wrr = a[1];
        cs->iter += 1;
        const double test = use_w ? wrr : rr_new;
        cs->wrr = wrr;
        cs->relres = use_w ? sqrt(wrr / cs->bnorm2) : sqrt(rr_new / cs->bnorm2);
        cs->beta = rr_new / cs->rr;
        cs->rr_prev = cs->rr;
        cs->rr = rr_new;
        if (!isfinite(rr_new)) {
            cs->done = 1;
            cs->converged = 0;
        } else if (test <= cs->tol2) {
            cs->done = 1;
            cs->converged = 1;
        } else if (cs->iter >= cs->maxit) {
            cs->done = 1;
            cs->converged = 0;
        }
    }
}

template <typename T>
__global__ void k_cg_direction(T* __restrict__ p, const T* __restrict__ r,
                               const CgScalars* __restrict__ cs, uint32_t N) {
    if (cs->done) return;
    const T beta = (T)cs->beta;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        p[v] = r[v] + beta * p[v];
}

// After the residual kernel (MODE 1): r.r, ||r||_w^2, ||b||^2.
__global__ void k_cg_start(const double* __restrict__ partials, int n, CgScalars* cs, double rtol,
                           int maxit) {
    __shared__ double lds[3 * 1024 / WAVE];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        a0 += partials[i];
        a1 += partials[n + i];
        a2 += partials[2 * n + i];
    }
    double a[3] = {a0, a1, a2};
    block_reduce<3, RED_SUM, RED_SUM, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        cs->rr = a[0];
        cs->wrr = a[1];
        cs->bnorm2 = a[2];
        cs->tol2 = (rtol * rtol) * a[2];
        cs->relres = (a[2] > 0.0) ? sqrt(a[1] / a[2]) : 0.0;
        cs->iter = 0;
        cs->maxit = maxit;
        cs->alpha = cs->beta = cs->pAp = 0.0;
        const bool ok = a[1] <= cs->tol2;
        cs->done = ok ? 1 : 0;
        cs->converged = ok ? 1 : 0;
    }
}

// FP32 residual replacement: r32 = r64 / scale, x32 = 0 (and p32 = r32 on
// the first replacement of a solve), partial r32.r32.
__global__ void k_inner_start(const double* __restrict__ r64, float* __restrict__ r32,
                              float* __restrict__ p32, float* __restrict__ x32, uint32_t N,
                              double inv_scale, int reset_p, double* __restrict__ partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double rr = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const float rv = (float)(r64[v] * inv_scale);
        r32[v] = rv;
        if (reset_p) p32[v] = rv;
        x32[v] = 0.0f;
        rr += (double)rv * (double)rv;
    }
    double a[1] = {rr};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

// Start (first = 1) or continue after a residual replacement (first = 0):
// r.r of the new FP32 residual; beta = r.r(new) / r.r(before the last
// iteration) so the search direction carries over; next trigger for the
// FP64 replacement.
__global__ void k_inner_init(const double* __restrict__ partials, int n, CgScalars* cs,
                             double trigger2, int maxit, int first) {
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        if (first) {
            cs->iter = 0;
            cs->bnorm2 = a[0];
            cs->beta = 0.0;
        } else {
            cs->beta = a[0] / cs->rr_prev;
        }
        cs->rr = a[0];
        cs->tol2 = trigger2;
        cs->wrr = 0.0;
        cs->relres = 1.0;
        cs->maxit = maxit;
        cs->done = (a[0] > 0.0 && cs->iter < maxit) ? 0 : 1;
        cs->converged = 0;
    }
}

__global__ void k_set_iter(CgScalars* cs, int iter) { cs->iter = iter; }

// x64 += scale * x32  (the FP64 refinement update)
__global__ void k_refine_add(double* __restrict__ x64, const float* __restrict__ x32, uint32_t N,
                             double scale) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        x64[v] += scale * (double)x32[v];
}

__global__ void k_scale_vec(const double* __restrict__ a, const double* __restrict__ s,
                            double* __restrict__ out, uint32_t N) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        out[v] = a[v] * s[v];
}

static int nblocks_stencil(const ls_ctx* c) {
    dim3 g = stencil_grid(c);
    return (int)(g.x * g.y * g.z);
}

// r = b^ - A^ x^ (FP64), p = r; fills cs (rr, wrr, bnorm2, tol2, done).
static int cg64_residual(ls_ctx* c, double rtol, int maxit) {
    SGeo g = sgeo(c);
    dim3 grd = stencil_grid(c), blk(TK, TJ);
    int nb = nblocks_stencil(c);
    hipLaunchKernelGGL((k_stencil<double, 1>), grd, blk, 0, c->stream, c->x64, c->r64, c->c64[0],
                       c->c64[1], c->c64[2], g, c->partials, nullptr, c->bhat, c->p64, c->sd);
    LS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(1024), 0, c->stream, c->partials, nb, c->cg, rtol,
                       maxit);
    LS_LAUNCH_CHECK();
    return 0;
}

// One stencil launch y = A^ x with the block partials of x.y; returns the
// number of partials.
template <typename T>
static int launch_stencil(ls_ctx* c, const T* x, T* y, T* const* faces, const int* done) {
    SGeo g = sgeo(c);
    if (use_vec<T>(c)) {
        dim3 grd = vec_grid<T>(c);
        hipLaunchKernelGGL(k_stencil_vec<T>, grd, dim3(WAVE, TJ), 0, c->stream, x, y, faces[0],
                           faces[1], faces[2], g, c->partials, done);
        return (int)(grd.x * grd.y * grd.z);
    }
    dim3 grd = stencil_grid(c);
    hipLaunchKernelGGL((k_stencil<T, 0>), grd, dim3(TK, TJ), 0, c->stream, x, y, faces[0],
                       faces[1], faces[2], g, c->partials, done, nullptr, nullptr, nullptr);
    return (int)(grd.x * grd.y * grd.z);
}

template <typename T>
static int cg_iterations(ls_ctx* c, T* x, T* r, T* p, T* ap, T* const* faces, const double* sdw,
                         int batch) {
    uint32_t N = (uint32_t)c->N;
    const bool vec = use_vec<T>(c);
    const int nbe = vec ? vec_blocks<T>(N) : ew_blocks(N);
    for (int b = 0; b < batch; ++b) {
        const bool sample = c->profiling && b == 0;
        if (sample) LS_HIP(hipEventRecord(c->ev[12], c->stream));
        const int nbs = launch_stencil<T>(c, p, ap, faces, &c->cg->done);
        if (sample) {
            LS_HIP(hipEventRecord(c->ev[13], c->stream));
            c->prof_pending = sizeof(T) == 8 ? 64 : 32;
        }
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1024), 0, c->stream, c->partials, nbs, c->cg);
        if (vec) {
            hipLaunchKernelGGL(k_cg_update_vec<T>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, x, r,
                               p, ap, sdw, c->cg, N, c->partials);
        } else {
            hipLaunchKernelGGL(k_cg_update<T>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, x, r, p,
                               ap, sdw, c->cg, N, c->partials);
        }
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe, c->cg,
                           sdw != nullptr ? 1 : 0);
        if (vec)
            hipLaunchKernelGGL(k_cg_direction_vec<T>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, p,
                               r, c->cg, N);
        else
            hipLaunchKernelGGL(k_cg_direction<T>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, p, r,
                               c->cg, N);
    }
    LS_LAUNCH_CHECK();
    return 0;
}

static int poll_cg(ls_ctx* c) {
    LS_HIP(hipMemcpyAsync(c->h_cg, c->cg, sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    if (c->prof_pending) {
        float ms = 0.0f;
        LS_HIP(hipEventElapsedTime(&ms, c->ev[12], c->ev[13]));
        const int q = c->prof_pending == 64 ? 0 : 1;
        c->prof_ms[q] += ms;
        c->prof_n[q] += 1;
        c->prof_pending = 0;
    }
    return 0;
}

static int batch_size(const ls_ctx* c) {
    return c->N >= (size_t)64 * 64 * 64 * 8 ? 16 : 32;
}

// Plain FP64 Jacobi-PCG (as CG on A^).  Returns 0 converged, 1 not.
static int solve_plain64(ls_ctx* c, ls_step_report* rep) {
    int st = cg64_residual(c, c->P.press_rtol, c->P.press_maxit);
    if (st) return st;
    if ((st = poll_cg(c))) return st;
    if (c->h_cg->bnorm2 == 0.0) {
        LS_HIP(hipMemsetAsync(c->x64, 0, c->N * sizeof(double), c->stream));
        rep->press_relres = 0.0;
        rep->press_bnorm = 0.0;
        return 0;
    }
    int total = 0, restarts = 0;
    const int bs = batch_size(c);
    while (true) {
        while (!c->h_cg->done) {
            if ((st = cg_iterations<double>(c, c->x64, c->r64, c->p64, c->ap64, c->c64, c->sd, bs)))
                return st;
            if ((st = poll_cg(c))) return st;
        }
        total += c->h_cg->iter;
        const bool conv = c->h_cg->converged;
        // confirm on the recomputed residual
        if ((st = cg64_residual(c, c->P.press_rtol, c->P.press_maxit - total))) return st;
        if ((st = poll_cg(c))) return st;
        if (c->h_cg->converged || !conv || total >= c->P.press_maxit || restarts > 5) break;
        ++restarts;
    }
    rep->press_iters += total;
    rep->press_outer += 1;
    rep->press_relres = c->h_cg->relres;
    rep->press_bnorm = sqrt(c->h_cg->bnorm2);
    return c->h_cg->converged ? 0 : 1;
}

// Mixed precision with reliable updates (Sleijpen & van der Vorst 1996, as
// used for mixed-precision CG by Clark et al. 2010): one CG run on A^ whose
// stencil, vectors and direction updates are FP32 (dot products FP64).
// Whenever the FP32 recurrence residual has fallen by `delta` since the last
// replacement, or its estimated true norm reaches the tolerance, the FP32
// solution is added into the FP64 iterate, the FP64 residual b - A x is
// recomputed with the FP64 operator and replaces the FP32 one.  The search
// direction is kept, so the Krylov convergence is not restarted.  The stop
// test is always on the FP64 residual.
static int solve_mixed(ls_ctx* c, ls_step_report* rep) {
    uint32_t N = (uint32_t)c->N;
    int nbe = ew_blocks(N);
    const int bs = batch_size(c);
    const double delta = c->P.mixed_inner_rtol > 0 ? c->P.mixed_inner_rtol : 0.1;
    int st;
    if ((st = cg64_residual(c, c->P.press_rtol, 1))) return st;
    if ((st = poll_cg(c))) return st;
    const double bnorm2 = c->h_cg->bnorm2;
    const double tol2 = c->h_cg->tol2;
    rep->press_bnorm = sqrt(bnorm2);
    rep->press_relres = c->h_cg->relres;
    if (bnorm2 == 0.0) {
        LS_HIP(hipMemsetAsync(c->x64, 0, c->N * sizeof(double), c->stream));
        rep->press_relres = 0.0;
        return 0;
    }
    if (c->h_cg->converged) return 0;
    // fixed power-of-two scale of the FP32 vectors for this solve
    const double scale = std::ldexp(1.0, std::ilogb(sqrt(c->h_cg->rr)));
    double rr64 = c->h_cg->rr, wrr = c->h_cg->wrr;
    double last_wrr = wrr;
    int stalls = 0;
    bool first = true;
    const int maxit = c->P.press_maxit;
    while (true) {
        // trigger: FP32 r.r fell by delta^2, or the estimated true norm hits tol
        const double rr32 = rr64 / (scale * scale);
        const double w2 = wrr / rr64;
        double trig = delta * delta * rr32;
        const double t2 = 0.81 * tol2 / (w2 * scale * scale);
        if (t2 > trig) trig = t2;
        hipLaunchKernelGGL(k_inner_start, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->r64, c->r32,
                           c->p32, c->x32, N, 1.0 / scale, first ? 1 : 0, c->partials);
        hipLaunchKernelGGL(k_inner_init, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe,
                           c->cg, trig, maxit, first ? 1 : 0);
        if (!first) {
            if (use_vec<float>(c))
                hipLaunchKernelGGL(k_cg_direction_vec<float>, dim3(vec_blocks<float>(N)),
                                   dim3(EW_BLOCK), 0, c->stream, c->p32, c->r32, c->cg, N);
            else
                hipLaunchKernelGGL(k_cg_direction<float>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream,
                                   c->p32, c->r32, c->cg, N);
        }
        LS_LAUNCH_CHECK();
        first = false;
        if ((st = poll_cg(c))) return st;
        while (!c->h_cg->done) {
            if ((st = cg_iterations<float>(c, c->x32, c->r32, c->p32, c->ap32, c->c32, nullptr, bs)))
                return st;
            if ((st = poll_cg(c))) return st;
        }
        const int iters = c->h_cg->iter;
        // stopped without reaching the trigger or the cap: breakdown / non-finite
        const bool broke = !c->h_cg->converged && iters < maxit;
        hipLaunchKernelGGL(k_refine_add, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->x64, c->x32,
                           N, scale);
        LS_LAUNCH_CHECK();
        rep->press_outer += 1;
        // FP64 residual replacement (keeps the device iteration counter)
        if ((st = cg64_residual(c, c->P.press_rtol, 1))) return st;
        if ((st = poll_cg(c))) return st;
        rr64 = c->h_cg->rr;
        wrr = c->h_cg->wrr;
        rep->press_relres = c->h_cg->relres;
        rep->press_iters = iters;
        if (c->h_cg->converged) return 0;
        if (iters >= maxit) return 1;
        // FP32 stagnation guard: the true residual must keep falling
        stalls = (wrr > 0.25 * last_wrr) ? stalls + 1 : 0;
        last_wrr = wrr;
        if (broke || stalls >= 3) {
            rep->press_fallback = 1;
            return solve_plain64(c, rep);   // FP64 PCG from the current iterate
        }
        // restore the device iteration counter reset by the residual kernel
        hipLaunchKernelGGL(k_set_iter, dim3(1), dim3(1), 0, c->stream, c->cg, iters);
        LS_LAUNCH_CHECK();
    }
}

static double effective_top_pressure(double top, double t) {
    const double ramp_duration = 3600.0;
    double eff = top;
    if (t < ramp_duration) eff = top * (t / ramp_duration) + 1e-6;
    return eff;
}

int solve_pressure(ls_ctx* c, double t, ls_step_report* rep) {
    const double ptop = effective_top_pressure(c->P.top_pressure, t);
    int st = assemble_pressure(c, t);
    if (st) return st;
    if (c->P.solver == LS_SOLVER_PLAIN64) st = solve_plain64(c, rep);
    else if (c->P.solver == LS_SOLVER_MIXED) st = solve_mixed(c, rep);
    else {
        c->err = "learned solver not available in this build";
        return -20;
    }
    if (st < 0) return st;
    int st2 = finalize_pressure(c, ptop);
    if (st2) return st2;
    c->p_top_prev = ptop;
    c->have_prev = true;
    rep->top_pressure_eff = ptop;
    LS_HIP(hipMemcpyAsync(c->h_ss, c->ss, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    rep->n_floored = (long long)c->h_ss->n_floored;
    if (st > 0) c->err = "pressure solve did not converge";
    return st;
}

int apply_pressure_unscaled(ls_ctx* c, const double* x_dev, double* y_dev) {
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    uint32_t N = (uint32_t)c->N;
    int nbe = ew_blocks(N);
    SGeo g = sgeo(c);
    hipLaunchKernelGGL(k_scale_vec, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, x_dev, c->sd, c->p64,
                       N);
    launch_stencil<double>(c, c->p64, c->ap64, c->c64, nullptr);
    hipLaunchKernelGGL(k_scale_vec, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->ap64, c->sd, y_dev,
                       N);
    LS_LAUNCH_CHECK();
    return 0;
}

int time_stencil(ls_ctx* c, int precision, int reps, double* ms, double* bytes) {
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    hipEvent_t e0 = c->ev[14], e1 = c->ev[15];
    for (int pass = 0; pass < 2; ++pass) {
        int n = pass == 0 ? 1 : reps;
        if (pass == 1) LS_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < n; ++r) {
            if (precision == 32) launch_stencil<float>(c, c->p32, c->ap32, c->c32, nullptr);
            else launch_stencil<double>(c, c->p64, c->ap64, c->c64, nullptr);
        }
        LS_LAUNCH_CHECK();
    }
    LS_HIP(hipEventRecord(e1, c->stream));
    LS_HIP(hipEventSynchronize(e1));
    float f = 0.0f;
    LS_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = (double)f / reps;
    const double sv = precision == 32 ? 4.0 : 8.0;
    *bytes = (double)c->N * (2.0 * sv + 3.0 * sv);
    return 0;
}

}  // namespace ls
