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
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>

#include "ls_context.h"
#include "ls_unet.h"

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
                if (p2 != nullptr) p2[v] = (T)r;
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
// edges load their single halo value.  y is written with non-temporal
// stores (streamed once); the faces are read with ordinary loads (the -j
// face row is re-read by the next row's wave).  Tile shape from sweeps at
// 512^3 on gfx950 (ICH 2..32, TJ 4..16, load/store flavours, 2 or 4 doubles
// per lane): ICH 4 x TJ 8 was fastest in both precisions (FP64: 1.03 ms
// against 1.07 ms for ICH 8 x TJ 4, 1.10-1.15 ms with 4 doubles per lane).
// ---------------------------------------------------------------------------
template <typename T, int W> struct VecOf;
template <> struct VecOf<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecOf<float, 2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct VecOf<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct VecOf<double, 4> { typedef double type __attribute__((ext_vector_type(4))); };

template <typename T> struct StencilShape;
template <> struct StencilShape<float> { static constexpr int W = 4, ICH = 4, TJV = 8; };
template <> struct StencilShape<double> { static constexpr int W = 2, ICH = 4, TJV = 8; };

// MODE 0: y = A^ x with the block partials of x.y.  MODE 1 (FP64, the
// residual of k_stencil<double, 1>): y = p2 = b - A^ x with the partials of
// r.r, |sd r|^2 and |sd b|^2 (three arrays of nb).  MODE 2: y = A^ x with
// the partials of x.y and x.b (two arrays of nb; the one-vector projection).
template <typename T, int MODE = 0>
__global__ __launch_bounds__(WAVE * StencilShape<T>::TJV) void k_stencil_vec(
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ cx,
    const T* __restrict__ cy, const T* __restrict__ cz, SGeo g, double* __restrict__ partials,
    const int* __restrict__ done, const T* __restrict__ b = nullptr, T* __restrict__ p2 = nullptr,
    const T* __restrict__ sd = nullptr) {
    constexpr int W = StencilShape<T>::W, ICH = StencilShape<T>::ICH;
    constexpr int TJV = StencilShape<T>::TJV;
    typedef typename VecOf<T, W>::type V;
    if (done != nullptr && *done) return;
    __shared__ double lds[3 * TJV];
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nb = nbx * nby * gridDim.z;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int t = xcd_remap(hw, nb);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int lane = threadIdx.x;
    const int k0 = (bx * WAVE + lane) * W;
    const int j = by * TJV + threadIdx.y;
    const int i0 = bz * ICH, i1 = min(i0 + ICH, g.ni);
    double acc = 0.0, acc1 = 0.0, acc2 = 0.0;
    // rows past nj idle as whole waves (threadIdx.y is uniform per wave)
    if (j < g.nj) {
        const bool ok = k0 < g.nk;
        const bool hj = j > 0, pj = j + 1 < g.nj;
        uint32_t v = (uint32_t)i0 * g.si + (uint32_t)j * g.sj + (ok ? (uint32_t)k0 : 0u);
        const V zero = V(0);
        V xm = (ok && i0 > 0) ? *(const V*)(x + v - g.si) : zero;
        V cxm = (ok && i0 > 0) ? *(const V*)(cx + v - g.si) : zero;
        V xc = ok ? *(const V*)(x + v) : zero;
        auto plane = [&](int i) {
            V xp = zero, cxp = zero, cyp = zero, czp = zero, cym = zero, xjp = zero, xjm = zero;
            if (ok) {
                if (i + 1 < g.ni) xp = *(const V*)(x + v + g.si);
                cxp = *(const V*)(cx + v);
                cyp = *(const V*)(cy + v);
                czp = *(const V*)(cz + v);
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
            V bv = zero, sv = zero;
            if (MODE == 1 && ok) {
                bv = *(const V*)(b + v);
                sv = *(const V*)(sd + v);
            }
            if (MODE == 2 && ok) bv = *(const V*)(b + v);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const T xkm = (w == 0) ? left : xc[w - 1];
                const T xkp = (w == W - 1) ? right : xc[w + 1];
                const T czm = (w == 0) ? czl : czp[w - 1];
                ax[w] = xc[w] - (cxp[w] * xp[w] + cxm[w] * xm[w] + cyp[w] * xjp[w] +
                                 cym[w] * xjm[w] + czp[w] * xkp + czm * xkm);
                if (MODE == 0) {
                    acc += (double)xc[w] * (double)ax[w];
                } else if (MODE == 2) {
                    acc += (double)xc[w] * (double)ax[w];
                    acc1 += (double)xc[w] * (double)bv[w];
                } else {
                    const double r = (double)bv[w] - (double)ax[w];
                    ax[w] = (T)r;
                    const double sr = (double)sv[w] * r, sb = (double)sv[w] * (double)bv[w];
                    acc += r * r;
                    acc1 += sr * sr;
                    acc2 += sb * sb;
                }
            }
            if (ok) {
                __builtin_nontemporal_store(ax, (V*)(y + v));
                if (MODE == 1 && p2 != nullptr) __builtin_nontemporal_store(ax, (V*)(p2 + v));
            }
            xm = xc;
            xc = xp;
            cxm = cxp;
            v += g.si;
        };
        if (i1 - i0 == ICH) {
            // full chunk: unrolled, so the next plane's loads overlap
#pragma unroll
            for (int q = 0; q < ICH; ++q) plane(i0 + q);
        } else {
            for (int i = i0; i < i1; ++i) plane(i);
        }
    }
    acc = wave_sum(acc);
    if (MODE != 0) {
        acc1 = wave_sum(acc1);
        acc2 = wave_sum(acc2);
    }
    if (lane == 0) {
        lds[threadIdx.y] = acc;
        if (MODE != 0) {
            lds[TJV + threadIdx.y] = acc1;
            lds[2 * TJV + threadIdx.y] = acc2;
        }
    }
    __syncthreads();
    if (flat_tid() == 0) {
        double s = 0.0;
        for (int q = 0; q < TJV; ++q) s += lds[q];
        partials[t] = s;
        if (MODE != 0) {
            double s1 = 0.0, s2 = 0.0;
            for (int q = 0; q < TJV; ++q) {
                s1 += lds[TJV + q];
                s2 += lds[2 * TJV + q];
            }
            partials[nb + t] = s1;
            if (MODE == 1) partials[2 * nb + t] = s2;
        }
    }
}

// ---------------------------------------------------------------------------
// Red-black Schur complement.  The 7-point graph is bipartite, so with the
// colour c = (i + j + k) & 1 (black 0, red 1) the scaled operator
// A^ = I - N has no coupling inside a colour, and eliminating the red
// unknowns leaves S = I - N_br N_rb on the black ones.  Since the spectrum of
// N is symmetric, cond(S) ~ cond(A^) / 4 and CG on S needs about half the
// iterations; red values follow from x_r = b_r + N_rb x_b.  One pass kernel,
// with n = sum over the 6 faces of c^ xn[neighbour]:
//   MODE 1 (gather)      y = red ? n : 0                      (w = N_rb p)
//   MODE 2 (apply)       y = black ? xc - n : 0, dot xc.y     (q = S p)
//   MODE 3 (rhs)         y = black ? xc + n : 0, dot y.y      (b_b + N_br b_r)
//   MODE 4 (reconstruct) y = red ? xc + n : xn               (x_r = r_r + N_rb x_b)
//   MODE 5 (rhs, packed) as MODE 3, y written as the compacted black vector
//                        (layout below k_schur_apply's header)
//   MODE 6 (reconstruct and add) as MODE 4, then xacc += scale * y (FP64)
//                        instead of storing y
// Vectorised like k_stencil_vec: W floats of one k-row per lane (4 when
// nk % 4 == 0, else 2 when nk is even), marching ICH planes along i.
// ---------------------------------------------------------------------------
template <int MODE, int W = 4>
__global__ __launch_bounds__(WAVE * StencilShape<float>::TJV) void k_rb_vec(
    const float* __restrict__ xn, const float* __restrict__ xc, float* __restrict__ y,
    const float* __restrict__ cx, const float* __restrict__ cy, const float* __restrict__ cz,
    SGeo g, double* __restrict__ partials, const int* __restrict__ done,
    double* __restrict__ xacc = nullptr, double scale = 0.0) {
    static_assert(W == 4 || (W == 2 && MODE != 5), "the packed output needs 4 values per lane");
    constexpr int ICH = StencilShape<float>::ICH, TJV = StencilShape<float>::TJV;
    typedef typename VecOf<float, W>::type V;
    if (done != nullptr && *done) return;
    __shared__ double lds[TJV];
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nb = nbx * nby * gridDim.z;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int t = xcd_remap(hw, nb);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int lane = threadIdx.x;
    const int k0 = (bx * WAVE + lane) * W;
    const int j = by * TJV + threadIdx.y;
    const int i0 = bz * ICH, i1 = min(i0 + ICH, g.ni);
    double acc = 0.0;
    if (j < g.nj) {
        const bool ok = k0 < g.nk;
        const bool hj = j > 0, pj = j + 1 < g.nj;
        uint32_t v = (uint32_t)i0 * g.si + (uint32_t)j * g.sj + (ok ? (uint32_t)k0 : 0u);
        const V zero = V(0);
        V xm = (ok && i0 > 0) ? *(const V*)(xn + v - g.si) : zero;
        V cxm = (ok && i0 > 0) ? *(const V*)(cx + v - g.si) : zero;
        V xcur = ok ? *(const V*)(xn + v) : zero;
        auto plane = [&](int i) {
            V xp = zero, cxp = zero, cyp = zero, czp = zero, cym = zero, xjp = zero, xjm = zero;
            V cen = zero;
            if (ok) {
                if (i + 1 < g.ni) xp = *(const V*)(xn + v + g.si);
                cxp = *(const V*)(cx + v);
                cyp = *(const V*)(cy + v);
                czp = *(const V*)(cz + v);
                if (hj) {
                    cym = *(const V*)(cy + v - g.sj);
                    xjm = *(const V*)(xn + v - g.sj);
                }
                if (pj) xjp = *(const V*)(xn + v + g.sj);
                if (MODE != 1) cen = *(const V*)(xc + v);
            }
            float left = __shfl_up(xcur[W - 1], 1, WAVE);
            float czl = __shfl_up(czp[W - 1], 1, WAVE);
            float right = __shfl_down(xcur[0], 1, WAVE);
            if (lane == 0) {
                left = (ok && k0 > 0) ? xn[v - 1] : 0.0f;
                czl = (ok && k0 > 0) ? cz[v - 1] : 0.0f;
            }
            if (lane == WAVE - 1) right = (ok && k0 + W < g.nk) ? xn[v + W] : 0.0f;
            const int par = (i + j + k0) & 1;      // colour of the lane's first element
            V out;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const float xkm = (w == 0) ? left : xcur[w - 1];
                const float xkp = (w == W - 1) ? right : xcur[w + 1];
                const float czm = (w == 0) ? czl : czp[w - 1];
                const float n = cxp[w] * xp[w] + cxm[w] * xm[w] + cyp[w] * xjp[w] +
                                cym[w] * xjm[w] + czp[w] * xkp + czm * xkm;
                const bool red = ((par + w) & 1) != 0;
                float o;
                if (MODE == 1) o = red ? n : 0.0f;
                else if (MODE == 2) o = red ? 0.0f : cen[w] - n;
                else if (MODE == 3 || MODE == 5) o = red ? 0.0f : cen[w] + n;
                else o = red ? cen[w] + n : xcur[w];
                out[w] = o;
                if (MODE == 2) acc += (double)cen[w] * (double)o;
                if (MODE == 3 || MODE == 5) acc += (double)o * (double)o;
            }
            if constexpr (MODE == 5) {
                // the lane's two black values, black first element at par
                typedef float F2 __attribute__((ext_vector_type(2)));
                const int pb = (i + j) & 1;
                const F2 o2 = pb ? F2{out[1], out[3]} : F2{out[0], out[2]};
                if (ok) *(F2*)(y + (v >> 1)) = o2;
            } else if constexpr (MODE == 6) {
                // as k_refine_add: x += scale * (double) y
                typedef typename VecOf<double, W>::type VD;
                if (ok) {
                    VD xv = *(const VD*)(xacc + v);
#pragma unroll
                    for (int w = 0; w < W; ++w) xv[w] += scale * (double)out[w];
                    *(VD*)(xacc + v) = xv;
                }
            } else if (ok) {
                __builtin_nontemporal_store(out, (V*)(y + v));
            }
            xm = xcur;
            xcur = xp;
            cxm = cxp;
            v += g.si;
        };
        if (i1 - i0 == ICH) {
#pragma unroll
            for (int q = 0; q < ICH; ++q) plane(i0 + q);
        } else {
            for (int i = i0; i < i1; ++i) plane(i);
        }
    }
    if (MODE == 2 || MODE == 3 || MODE == 5) {
        acc = wave_sum(acc);
        if (lane == 0) lds[threadIdx.y] = acc;
        __syncthreads();
        if (flat_tid() == 0) {
            double s = 0.0;
            for (int q = 0; q < TJV; ++q) s += lds[q];
            partials[t] = s;
        }
    }
}

// The red-black pass for any nk (modes 1-4, full-length vectors): one voxel
// per thread, marching ICHUNK planes along i on k_stencil's grid, with the
// neighbour sum n in k_rb_vec's order.  One FP64 block partial per block.
template <int MODE>
__global__ __launch_bounds__(TK * TJ) void k_rb_scalar(
    const float* __restrict__ xn, const float* __restrict__ xc, float* __restrict__ y,
    const float* __restrict__ cx, const float* __restrict__ cy, const float* __restrict__ cz,
    SGeo g, double* __restrict__ partials, const int* __restrict__ done) {
    static_assert(MODE >= 1 && MODE <= 4, "compacted output needs k_rb_vec");
    if (done != nullptr && *done) return;
    __shared__ double lds[(TK * TJ) / WAVE];
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nb = nbx * nby * gridDim.z;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int t = xcd_remap(hw, nb);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int k = bx * TK + threadIdx.x;
    const int j = by * TJ + threadIdx.y;
    const int i0 = bz * ICHUNK;
    const int i1 = min(i0 + ICHUNK, g.ni);
    double acc = 0.0;
    if (k < g.nk && j < g.nj) {
        uint32_t v = (uint32_t)i0 * g.si + (uint32_t)j * g.sj + (uint32_t)k;
        const bool hj = j > 0, hk = k > 0, pj = j + 1 < g.nj, pk = k + 1 < g.nk;
        float xm = (i0 > 0) ? xn[v - g.si] : 0.0f;
        float cxm = (i0 > 0) ? cx[v - g.si] : 0.0f;
        float xcur = xn[v];
        for (int i = i0; i < i1; ++i, v += g.si) {
            const float xp = (i + 1 < g.ni) ? xn[v + g.si] : 0.0f;
            const float cxp = cx[v];
            const float cyp = cy[v];
            const float czp = cz[v];
            const float cym = hj ? cy[v - g.sj] : 0.0f;
            const float czm = hk ? cz[v - 1] : 0.0f;
            const float xjp = pj ? xn[v + g.sj] : 0.0f;
            const float xjm = hj ? xn[v - g.sj] : 0.0f;
            const float xkp = pk ? xn[v + 1] : 0.0f;
            const float xkm = hk ? xn[v - 1] : 0.0f;
            const float n = cxp * xp + cxm * xm + cyp * xjp + cym * xjm + czp * xkp + czm * xkm;
            const bool red = ((i + j + k) & 1) != 0;
            float o;
            if (MODE == 1) {
                o = red ? n : 0.0f;
            } else {
                const float cen = xc[v];
                if (MODE == 2) o = red ? 0.0f : cen - n;
                else if (MODE == 3) o = red ? 0.0f : cen + n;
                else o = red ? cen + n : xcur;
                if (MODE == 2) acc += (double)cen * (double)o;
                if (MODE == 3) acc += (double)o * (double)o;
            }
            __builtin_nontemporal_store(o, y + v);
            xm = xcur;
            xcur = xp;
            cxm = cxp;
        }
    }
    if (MODE == 2 || MODE == 3) {
        double a[1] = {acc};
        block_reduce<1, RED_SUM>(a, lds);
        if (flat_tid() == 0) partials[t] = a[0];
    }
}

// ---------------------------------------------------------------------------
// The Schur complement on colour-compacted vectors.
//
// With nk even, voxel v = (i nj + j) nk + k of either colour is stored at
// v >> 1 of its colour's half-length array: row (i, j) holds its nk / 2
// black voxels k = par + 2m (par = (i + j) & 1) at m = 0 .. nk/2 - 1, and its
// red voxels k = 1 - par + 2m likewise.  The i and j neighbours of a voxel
// (other colour, other row parity) then sit at the same m; the k neighbours
// of red m at black m - par and m + 1 - par, those of black m at red
// m - 1 + par and m + par.
//
// k_schur_apply: q = S p = p - N_br (N_rb p) on compacted black vectors in
// one pass.  A block owns TJ rows of full k length and marches ICH planes
// along i; wave (y, z) covers 256 compacted values (1024 B) of row
// j0 - 1 + z, the first and last row being halo rows that only compute w.
// At plane i each wave computes w = N_rb p at the red voxels of its row,
// keeps its own row's w of planes i - 2 .. i in registers and publishes
// plane i in LDS (three rotating buffers, one barrier per plane); q of plane
// i - 1 then takes w of the own row (planes i - 2, i), of rows j +- 1 (LDS)
// and of the k neighbours (shuffles; LDS at the wave edges).  A face serves
// w at plane i and q at the black end one plane later from registers, so
// HBM sees p, the three face arrays and q once: 16 B per voxel of the grid
// (the two-pass form on full-length vectors moves 44 B).  Each product is
// summed in the order of k_rb_vec (+i, -i, +j, -j, +k, -k).
// ---------------------------------------------------------------------------
struct RbGeo {
    int ni, nj, nh;            // nh = nk / 2
    uint32_t si, sj, hsi, hsj; // full and compacted strides
};

static RbGeo rbgeo(const ls_ctx* c) {
    RbGeo g;
    g.ni = c->ni; g.nj = c->nj; g.nh = c->nk / 2;
    g.sj = (uint32_t)c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    g.hsj = (uint32_t)g.nh;
    g.hsi = (uint32_t)g.nh * (uint32_t)c->nj;
    return g;
}

#ifndef LS_SCHUR_ROWS
#define LS_SCHUR_ROWS 6        // TJ: rows per block (plus 2 halo rows)
#endif
#ifndef LS_SCHUR_ICH
#define LS_SCHUR_ICH 16        // planes per block
#endif
#ifndef LS_SCHUR_THREADS
#define LS_SCHUR_THREADS 512   // launch bound: 64 x WR x (TJ + 2) at most
#endif
constexpr int SCHUR_ICH = LS_SCHUR_ICH;

typedef float F4 __attribute__((ext_vector_type(4)));
template <int N> struct FVec { typedef float type __attribute__((ext_vector_type(N))); };

#ifndef LS_SCHUR_VW
#define LS_SCHUR_VW 4          // compacted values per lane (2 or 4)
#endif
constexpr int SCHUR_VW = LS_SCHUR_VW;
// Rows wider than one wave's 4 x 64 compacted values (nk > 512) take
// 1024-thread blocks, so that a block still holds LS_SCHUR_ROWS output rows.
constexpr int SCHUR_WIDE_THREADS = 1024;
// Widest row: a block holds WR waves per row and at least 3 rows.
constexpr int SCHUR_MAX_NK = 2 * SCHUR_VW * WAVE * ((SCHUR_WIDE_THREADS / WAVE) / 3);

// Colour split of the 2 VW consecutive values a of one row whose first
// element has colour par (0 black): b[x] = a[2x + par], r[x] = a[2x + 1 - par].
template <int VW>
__device__ __forceinline__ void csplit(const typename FVec<2 * VW>::type& a, int par,
                                       typename FVec<VW>::type& b, typename FVec<VW>::type& r) {
    typename FVec<VW>::type e, o;
#pragma unroll
    for (int x = 0; x < VW; ++x) {
        e[x] = a[2 * x];
        o[x] = a[2 * x + 1];
    }
    b = par ? o : e;
    r = par ? e : o;
}

template <int ICH, int VW, int NT = LS_SCHUR_THREADS>
__global__ __launch_bounds__(NT) void k_schur_apply(
    const float* __restrict__ p, float* __restrict__ q, const float* __restrict__ cx,
    const float* __restrict__ cy, const float* __restrict__ cz, RbGeo g,
    double* __restrict__ partials, const int* __restrict__ done) {
    typedef typename FVec<VW>::type V;          // compacted values of one lane
    typedef typename FVec<2 * VW>::type VF;     // the lane's full-row values (faces)
    if (done != nullptr && *done) return;
    extern __shared__ float wsh[];                   // [3][NS][rowlen]
    __shared__ double red[NT / WAVE];
    const int lane = threadIdx.x, wr = threadIdx.y, s = threadIdx.z;
    const int NS = blockDim.z, TJ = NS - 2;
    const int rowlen = blockDim.y * (WAVE * VW);
    const int bstride = NS * rowlen;
    const int nbj = gridDim.x;
    const int t = xcd_remap(blockIdx.x + nbj * blockIdx.y, nbj * gridDim.y);
    const int j = (t % nbj) * TJ - 1 + s;
    const int i0 = (t / nbj) * ICH, i1 = min(i0 + ICH, g.ni);
    const int mloc = (wr * WAVE + lane) * VW;
    const bool ok = j >= 0 && j < g.nj && mloc < g.nh;
    const bool halo = s == 0 || s == NS - 1;
    const bool hj = j > 0, pj = j + 1 < g.nj;
    const bool hl = mloc > 0, hr = mloc + VW < g.nh;
    const uint32_t vrow = ok ? (uint32_t)j * g.sj + 2u * (uint32_t)mloc : 0u;
    const uint32_t hrow = ok ? (uint32_t)j * g.hsj + (uint32_t)mloc : 0u;
    float* const lrow = wsh + s * rowlen + mloc;
    const V zero = V(0.0f);
    const VF zf = VF(0.0f);
    V Pm = zero, Pc = zero;                 // p of planes i - 1, i (own row)
    V Wmm = zero, Wm = zero;                // w of planes i - 2, i - 1 (own row)
    V cxB1 = zero, cxR1 = zero, cxR2 = zero;    // cx: plane i - 1 (black, red), i - 2 (red)
    V cyB1 = zero, cymB1 = zero, czB1 = zero, czR1 = zero;   // plane i - 1
    float czL1 = 0.0f;
    double acc = 0.0;
    if (ok && i0 >= 2) Pc = *(const V*)(p + (uint32_t)(i0 - 2) * g.hsi + hrow);
#pragma unroll
    for (int st = 0; st < ICH + 3; ++st) {
        const int i = i0 - 2 + st;
        if (i > i1) break;
        const bool inp = ok && i >= 0 && i < g.ni;
        const int par = (i + j) & 1;
        const uint32_t v = (uint32_t)i * g.si + vrow;
        const uint32_t h = (uint32_t)i * g.hsi + hrow;
        // h and v wrap for the planes i < 0 (never dereferenced there);
        // plane i + 1 is indexed from its own unsigned offset
        const uint32_t hn = h + g.hsi;
        V Pn = zero;
        if (ok && i + 1 >= 0 && i + 1 < g.ni) Pn = *(const V*)(p + hn);
        VF cxv = zf;
        if (inp) cxv = *(const VF*)(cx + v);
        V cxB, cxR;
        csplit<VW>(cxv, par, cxB, cxR);
        V cyB = zero, cymB = zero, czB = zero, czR = zero;
        float czL = 0.0f;
        if (st >= 1) {
            VF cyv = zf, czv = zf, cmv = zf;
            V Pjm = zero, Pjp = zero;
            float pL = 0.0f, pR = 0.0f, czl = 0.0f;
            if (inp) {
                cyv = *(const VF*)(cy + v);
                czv = *(const VF*)(cz + v);
                if (hj) {
                    cmv = *(const VF*)(cy + v - g.sj);
                    Pjm = *(const V*)(p + h - g.hsj);
                }
                if (pj) Pjp = *(const V*)(p + h + g.hsj);
                if (lane == 0 && hl) {
                    pL = p[h - 1];
                    czl = cz[v - 1];
                }
                if (lane == WAVE - 1 && hr) pR = p[h + VW];
            }
            const float sl = __shfl_up(Pc[VW - 1], 1, WAVE);
            const float sr = __shfl_down(Pc[0], 1, WAVE);
            const float sz = __shfl_up(czv[2 * VW - 1], 1, WAVE);
            if (lane != 0) { pL = sl; czl = sz; }
            if (lane != WAVE - 1) pR = sr;
            czL = czl;
            V cyR, cymR;
            csplit<VW>(cyv, par, cyB, cyR);
            csplit<VW>(cmv, par, cymB, cymR);
            csplit<VW>(czv, par, czB, czR);
            // w at the red voxels of plane i
            V Wc = zero;
            if (i >= 0 && i < g.ni) {
#pragma unroll
                for (int x = 0; x < VW; ++x) {
                    const float pkp = par ? Pc[x] : (x < VW - 1 ? Pc[x + 1] : pR);
                    const float pkm = par ? (x > 0 ? Pc[x - 1] : pL) : Pc[x];
                    const float czm = par ? (x > 0 ? czB[x - 1] : czL) : czB[x];
                    Wc[x] = cxR[x] * Pn[x] + cxB1[x] * Pm[x] + cyR[x] * Pjp[x] +
                            cymR[x] * Pjm[x] + czR[x] * pkp + czm * pkm;
                }
            }
            if (st >= 2) *(V*)(lrow + (st % 3) * bstride) = Wc;
            if (st >= 3) {
                __syncthreads();
                // q at the black voxels of plane i - 1 (parity 1 - par)
                const float* lb = lrow + ((st - 1) % 3) * bstride;
                const float sl2 = __shfl_up(Wm[VW - 1], 1, WAVE);
                const float sr2 = __shfl_down(Wm[0], 1, WAVE);
                if (!halo) {
                    const V Wjm = *(const V*)(lb - rowlen);
                    const V Wjp = *(const V*)(lb + rowlen);
                    float wL = sl2, wR = sr2;
                    if (lane == 0) wL = hl ? lb[-1] : 0.0f;
                    if (lane == WAVE - 1) wR = hr ? lb[VW] : 0.0f;
                    const int pb = 1 - par;
                    V Q;
#pragma unroll
                    for (int x = 0; x < VW; ++x) {
                        const float wkp = pb ? (x < VW - 1 ? Wm[x + 1] : wR) : Wm[x];
                        const float wkm = pb ? Wm[x] : (x > 0 ? Wm[x - 1] : wL);
                        const float czm = pb ? czR1[x] : (x > 0 ? czR1[x - 1] : czL1);
                        const float n = cxB1[x] * Wc[x] + cxR2[x] * Wmm[x] + cyB1[x] * Wjp[x] +
                                        cymB1[x] * Wjm[x] + czB1[x] * wkp + czm * wkm;
                        Q[x] = Pm[x] - n;
                        acc += (double)Pm[x] * (double)Q[x];
                    }
                    if (ok) __builtin_nontemporal_store(Q, (V*)(q + h - g.hsi));
                }
            }
            Wmm = Wm;
            Wm = Wc;
        }
        Pm = Pc;
        Pc = Pn;
        cxR2 = cxR1;
        cxR1 = cxR;
        cxB1 = cxB;
        cyB1 = cyB;
        cymB1 = cymB;
        czB1 = czB;
        czR1 = czR;
        czL1 = czL;
    }
    acc = wave_sum(acc);
    const int wid = threadIdx.y + blockDim.y * threadIdx.z;
    const int nw = blockDim.y * blockDim.z;
    if (lane == 0) red[wid] = acc;
    __syncthreads();
    if (wid == 0 && lane == 0) {
        double a = 0.0;
        for (int w = 0; w < nw; ++w) a += red[w];
        partials[t] = a;
    }
}

// Compacted black vector -> full-length vector with zero red entries.
__global__ __launch_bounds__(EW_BLOCK) void k_rb_unpack(const float* __restrict__ hc,
                                                       float* __restrict__ y, SGeo g,
                                                       uint32_t N) {
    const uint32_t v = 4u * (blockIdx.x * (uint32_t)blockDim.x + threadIdx.x);
    if (v >= N) return;
    const uint32_t row = v / g.sj;                        // i nj + j
    const int par = (int)((row / (uint32_t)g.nj + row % (uint32_t)g.nj) & 1u);
    typedef float F2 __attribute__((ext_vector_type(2)));
    const F2 e = *(const F2*)(hc + (v >> 1));
    const F4 o = par ? F4{0.0f, e[0], 0.0f, e[1]} : F4{e[0], 0.0f, e[1], 0.0f};
    *(F4*)(y + v) = o;
}

// Vectorised CG vector updates: U vectors of 16 B per thread, all loads
// issued before the stores, non-temporal (each vector is streamed once per
// kernel).  The scalar tail (N % W) is handled by block 0.
constexpr int VU = 4;

// WITHX: also x += alpha p (flexible CG); plain CG leaves it to the
// direction kernel, which reads p anyway.
template <typename T, bool WITHX>
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
            if (WITHX) {
                xv[u] = __builtin_nontemporal_load((const V*)x + q);
                pv[u] = __builtin_nontemporal_load((const V*)p + q);
            }
            rv[u] = __builtin_nontemporal_load((const V*)r + q);
            av[u] = __builtin_nontemporal_load((const V*)ap + q);
        }
    }
    double rr = 0.0, wrr = 0.0;
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) {
            const V rn = rv[u] - a * av[u];
            if (WITHX) __builtin_nontemporal_store(xv[u] + a * pv[u], (V*)x + q);
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
        if (WITHX) x[v] = x[v] + a * p[v];
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

// x += alpha p (when this iteration's update is owed: xlive), then
// p = r + beta p (unless the solve stopped).  After a residual replacement
// (xlive = 0, done = 0) only the direction is rebuilt.
template <typename T>
__global__ __launch_bounds__(EW_BLOCK) void k_cg_direction_vec(
    T* __restrict__ p, const T* __restrict__ r, T* __restrict__ x,
    const CgScalars* __restrict__ cs, uint32_t N) {
    const bool xl = cs->xlive != 0, pl = cs->done == 0;
    if (!xl && !pl) return;
    constexpr int W = StencilShape<T>::W;
    typedef typename VecOf<T, W>::type V;
    const T beta = (T)cs->beta;
    const T a = (T)cs->alpha;
    const uint32_t nv = N / W;
    const uint32_t base = blockIdx.x * (uint32_t)(EW_BLOCK * VU) + threadIdx.x;
    // beta == 0 (a fresh direction): p = r, and p is read only for an owed
    // x update, so a direction vector that holds another phase's bytes is
    // never read
    const bool rp = xl || beta != (T)0;
    V pv[VU], rv[VU], xv[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) {
            pv[u] = rp ? __builtin_nontemporal_load((const V*)p + q) : V{};
            if (pl) rv[u] = __builtin_nontemporal_load((const V*)r + q);
            if (xl) xv[u] = __builtin_nontemporal_load((const V*)x + q);
        }
    }
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const uint32_t q = base + u * EW_BLOCK;
        if (q < nv) {
            if (xl) __builtin_nontemporal_store(xv[u] + a * pv[u], (V*)x + q);
            if (pl) __builtin_nontemporal_store(beta != (T)0 ? rv[u] + beta * pv[u] : rv[u],
                                                (V*)p + q);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < N - nv * W) {
        const uint32_t v = nv * W + threadIdx.x;
        const T pv0 = rp ? p[v] : (T)0;
        if (xl) x[v] = x[v] + a * pv0;
        if (pl) p[v] = beta != (T)0 ? r[v] + beta * pv0 : r[v];
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
    constexpr int TJV = StencilShape<T>::TJV;
    return dim3((c->nk / W + WAVE - 1) / WAVE, (c->nj + TJV - 1) / TJV, (c->ni + ICH - 1) / ICH);
}

// The vectorised path needs nk % W == 0 and its block partials to fit.
template <typename T>
static bool use_vec(const ls_ctx* c) {
    if (c->nk % StencilShape<T>::W != 0) return false;
    const dim3 g = vec_grid<T>(c);
    return (size_t)g.x * g.y * g.z <= (size_t)MAX_PARTIALS &&
           2 * (size_t)vec_blocks<T>((uint32_t)c->N) <= (size_t)MAX_PARTIALS;
}

// The vectorised elementwise CG kernels take any length (block 0 does the
// scalar tail); they need their block partials to fit.
template <typename T>
static bool ew_vec(uint32_t N) {
    return 2 * (size_t)vec_blocks<T>(N) <= (size_t)MAX_PARTIALS;
}

// Width of the red-black passes over full-length FP32 vectors: 4 floats per
// lane (nk % 4 == 0), 2 (nk % 4 == 2), or 0 for one voxel per thread
// (k_rb_scalar, odd nk).
static dim3 rb_grid(const ls_ctx* c, int w) {
    constexpr int ICH = StencilShape<float>::ICH, TJV = StencilShape<float>::TJV;
    return dim3((c->nk / w + WAVE - 1) / WAVE, (c->nj + TJV - 1) / TJV, (c->ni + ICH - 1) / ICH);
}

static int rb_width(const ls_ctx* c) {
    if (use_vec<float>(c)) return 4;
    if (c->nk % 2 != 0) return 0;
    const dim3 g = rb_grid(c, 2);
    return (size_t)g.x * g.y * g.z <= (size_t)MAX_PARTIALS ? 2 : 0;
}

// alpha = r.r / p.Ap
__global__ void k_cg_alpha(const double* __restrict__ partials, int n, CgScalars* cs) {
    if (cs->done) {
        if (threadIdx.x == 0) cs->xlive = 0;
        return;
    }
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
            cs->xlive = 0;
        } else {
            cs->alpha = cs->rr / pAp;
            cs->xlive = 1;
        }
    }
}

// r -= alpha Ap (and x += alpha p when WITHX), partials of r.r and (when
// sd != null) the unscaled residual norm sum (sd r)^2.
template <typename T, bool WITHX>
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
        const T rv = r[v] - a * ap[v];
        if (WITHX) x[v] = x[v] + a * p[v];
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
    if (use_w) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            a0 += partials[i];
            a1 += partials[n + i];
        }
    } else {   // the update kernel wrote zeros there (no weights): not read
        for (int i = threadIdx.x; i < n; i += blockDim.x) a0 += partials[i];
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
__global__ void k_cg_direction(T* __restrict__ p, const T* __restrict__ r, T* __restrict__ x,
                               const CgScalars* __restrict__ cs, uint32_t N) {
    const bool xl = cs->xlive != 0, pl = cs->done == 0;
    if (!xl && !pl) return;
    const T beta = (T)cs->beta;
    const T a = (T)cs->alpha;
    const bool rp = xl || beta != (T)0;     // as k_cg_direction_vec
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const T pv = rp ? p[v] : (T)0;
        if (xl) x[v] = x[v] + a * pv;
        if (pl) p[v] = beta != (T)0 ? r[v] + beta * pv : r[v];
    }
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
        cs->xlive = 0;
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
        cs->xlive = 0;
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
// The FP64 residual r64 = b^ - A^ x64 and the CG scalars of its norms;
// with_p also p64 = r64 (FP64 CG's first direction; the mixed solvers start
// their FP32 iterations from r64 alone).
static int cg64_residual(ls_ctx* c, double rtol, int maxit, bool with_p = true) {
    if (c->resid_ready > 0) {
        // the projected start left r64 (and p64) and the partials in place
        hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(1024), 0, c->stream, c->partials,
                           c->resid_ready, c->cg, rtol, maxit);
        c->resid_ready = 0;
        LS_LAUNCH_CHECK();
        return 0;
    }
    SGeo g = sgeo(c);
    int nb;
    if (use_vec<double>(c)) {
        const dim3 grd = vec_grid<double>(c);
        nb = (int)(grd.x * grd.y * grd.z);
        hipLaunchKernelGGL((k_stencil_vec<double, 1>), grd,
                           dim3(WAVE, StencilShape<double>::TJV), 0, c->stream, c->x64, c->r64,
                           c->c64[0], c->c64[1], c->c64[2], g, c->partials, nullptr, c->bhat,
                           with_p ? c->p64 : nullptr, c->sd);
    } else {
        dim3 grd = stencil_grid(c), blk(TK, TJ);
        nb = nblocks_stencil(c);
        hipLaunchKernelGGL((k_stencil<double, 1>), grd, blk, 0, c->stream, c->x64, c->r64,
                           c->c64[0], c->c64[1], c->c64[2], g, c->partials, nullptr, c->bhat,
                           with_p ? c->p64 : nullptr, c->sd);
    }
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
        hipLaunchKernelGGL(k_stencil_vec<T>, grd, dim3(WAVE, StencilShape<T>::TJV), 0, c->stream, x, y, faces[0],
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
    const bool vec = ew_vec<T>(N);
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
            hipLaunchKernelGGL((k_cg_update_vec<T, false>), dim3(nbe), dim3(EW_BLOCK), 0,
                               c->stream, x, r, p, ap, sdw, c->cg, N, c->partials);
        } else {
            hipLaunchKernelGGL((k_cg_update<T, false>), dim3(nbe), dim3(EW_BLOCK), 0, c->stream, x,
                               r, p, ap, sdw, c->cg, N, c->partials);
        }
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe, c->cg,
                           sdw != nullptr ? 1 : 0);
        if (vec)
            hipLaunchKernelGGL(k_cg_direction_vec<T>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, p,
                               r, x, c->cg, N);
        else
            hipLaunchKernelGGL(k_cg_direction<T>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, p, r,
                               x, c->cg, N);
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
        const int q = c->prof_pending == 64 ? 0 : (c->prof_pending == 32 ? 1 : 2);
        c->prof_ms[q] += ms;
        c->prof_n[q] += 1;
        c->prof_pending = 0;
    }
    return 0;
}

static int batch_size(const ls_ctx* c) {
    return c->N >= (size_t)64 * 64 * 64 * 8 ? 16 : 32;
}

// Batch sizing for the inner FP32 CG of the mixed solver.  The device stops at
// the trigger by itself, so the launches queued past it return at once, but
// each still costs a dispatch: at 64^3 about as much as a working kernel, and
// a reliable-update cycle of ~50 iterations wasted a third of its launches in
// batches of 32.  The next batch is sized from the residual's decay per
// iteration over the last batch (kept across cycles) to reach the trigger
// with a 25 % margin, within [4, bs].  Where the host polls changes no
// iterate: every stop test runs on the device.
struct InnerBatch {
    int bs;
    double rate = 0.0;       // ln(rr) decrease per iteration over the last batch
    double last_rr = 0.0;
    int last_iter = 0;
    explicit InnerBatch(int b) : bs(b) {}
    void start(double rr, int iter) { last_rr = rr; last_iter = iter; }
    int next(double rr, int iter, double trig) {
        if (last_rr > 0.0 && rr > 0.0 && rr < last_rr && iter > last_iter)
            rate = std::log(last_rr / rr) / (iter - last_iter);
        last_rr = rr;
        last_iter = iter;
        if (!(rate > 0.0) || !(rr > trig) || !(trig > 0.0)) return bs;
        const double need = std::log(rr / trig) / rate;
        if (!(need < (double)bs)) return bs;
        return std::max(4, std::min(bs, (int)std::ceil(1.25 * need) + 2));
    }
};

// Plain FP64 Jacobi-PCG (as CG on A^).  Returns 0 converged, 1 not.
static int solve_plain64(ls_ctx* c, ls_step_report* rep, int maxit) {
    int st = cg64_residual(c, c->P.press_rtol, maxit);
    if (st) return st;
    if ((st = poll_cg(c))) return st;
    if (!rep->press_outer) rep->press_relres0 = c->h_cg->relres;
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
        if ((st = cg64_residual(c, c->P.press_rtol, maxit - total))) return st;
        if ((st = poll_cg(c))) return st;
        if (c->h_cg->converged || !conv || total >= maxit || restarts > 5) break;
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
    InnerBatch ib(batch_size(c));
    const double delta = c->P.mixed_inner_rtol > 0 ? c->P.mixed_inner_rtol : 0.1;
    int st;
    if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
    if ((st = poll_cg(c))) return st;
    const double bnorm2 = c->h_cg->bnorm2;
    const double tol2 = c->h_cg->tol2;
    rep->press_bnorm = sqrt(bnorm2);
    rep->press_relres = c->h_cg->relres;
    if (!rep->press_outer) rep->press_relres0 = c->h_cg->relres;
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
    // residual curve of this solve (the learned solver races against it)
    const auto t_start = std::chrono::steady_clock::now();
    const double rr_start32 = rr64 / (scale * scale);
    c->race_tmp.clear();
    auto record = [&](double rr32) {
        if (!(rr32 > 0.0)) return;
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t_start).count();
        c->race_tmp.emplace_back(ms, 0.5 * std::log10(rr_start32 / rr32));
    };
    // FP32 budget: an error concentrated in the slowest modes (eigenvalues
    // near the FP32 rounding of the operator) makes the FP32 recurrence
    // crawl; past this many iterations FP64 PCG finishes from the iterate
    const int maxit = std::min(c->P.press_maxit, std::max(2000, 3 * (c->P.press_maxit / 4)));
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
            // xlive = 0 here: only the direction carries over
            if (ew_vec<float>(N))
                hipLaunchKernelGGL(k_cg_direction_vec<float>, dim3(vec_blocks<float>(N)),
                                   dim3(EW_BLOCK), 0, c->stream, c->p32, c->r32, c->x32, c->cg, N);
            else
                hipLaunchKernelGGL(k_cg_direction<float>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream,
                                   c->p32, c->r32, c->x32, c->cg, N);
        }
        LS_LAUNCH_CHECK();
        first = false;
        if ((st = poll_cg(c))) return st;
        ib.start(c->h_cg->rr, c->h_cg->iter);
        int nb = ib.next(c->h_cg->rr, c->h_cg->iter, trig);
        while (!c->h_cg->done) {
            if ((st = cg_iterations<float>(c, c->x32, c->r32, c->p32, c->ap32, c->c32, nullptr, nb)))
                return st;
            if ((st = poll_cg(c))) return st;
            record(c->h_cg->rr);
            nb = ib.next(c->h_cg->rr, c->h_cg->iter, trig);
        }
        const int iters = c->h_cg->iter;
        // stopped without reaching the trigger or the cap: breakdown / non-finite
        const bool broke = !c->h_cg->converged && iters < maxit;
        hipLaunchKernelGGL(k_refine_add, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->x64, c->x32,
                           N, scale);
        LS_LAUNCH_CHECK();
        rep->press_outer += 1;
        // FP64 residual replacement (keeps the device iteration counter)
        if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
        if ((st = poll_cg(c))) return st;
        rr64 = c->h_cg->rr;
        wrr = c->h_cg->wrr;
        rep->press_relres = c->h_cg->relres;
        rep->press_iters = iters;
        record(rr64 / (scale * scale));
        if (c->h_cg->converged) {
            c->race.swap(c->race_tmp);
            c->have_rate = true;
            return 0;
        }
        // FP32 stagnation guard: the true residual must keep falling
        stalls = (wrr > 0.25 * last_wrr) ? stalls + 1 : 0;
        last_wrr = wrr;
        if (broke || stalls >= 3 || iters >= maxit) {
            rep->press_fallback = 1;
            // FP64 PCG from the current iterate, with the remaining budget
            const int left = c->P.press_maxit - iters;
            if (left <= 0) return 1;
            return solve_plain64(c, rep, left);
        }
        // restore the device iteration counter reset by the residual kernel
        hipLaunchKernelGGL(k_set_iter, dim3(1), dim3(1), 0, c->stream, c->cg, iters);
        LS_LAUNCH_CHECK();
    }
}

// Launch of one red-black pass over the FP32 faces; returns the number of
// block partials (MODE 2 and 3 write them).
// x64 += scale * (red reconstruction of e): k_rb_vec mode 6 in one pass, or
// mode 4 into the scratch w and k_refine_add (one voxel per thread).
static void launch_reconstruct_add(ls_ctx* c, const float* e, const float* r, float* w,
                                   double scale);

template <int MODE>
static int launch_rb(ls_ctx* c, const float* xn, const float* xc, float* y, const int* done) {
    const int w = MODE == 5 ? 4 : rb_width(c);
    if (w == 0) {
        const dim3 grd = stencil_grid(c);
        hipLaunchKernelGGL(k_rb_scalar<(MODE == 5 ? 1 : MODE)>, grd, dim3(TK, TJ), 0, c->stream,
                           xn, xc, y, c->c32[0], c->c32[1], c->c32[2], sgeo(c), c->partials,
                           done);
        return (int)(grd.x * grd.y * grd.z);
    }
    const dim3 grd = rb_grid(c, w);
    if constexpr (MODE != 5) {
        if (w == 2) {
            hipLaunchKernelGGL((k_rb_vec<MODE, 2>), grd, dim3(WAVE, StencilShape<float>::TJV), 0,
                               c->stream, xn, xc, y, c->c32[0], c->c32[1], c->c32[2], sgeo(c),
                               c->partials, done);
            return (int)(grd.x * grd.y * grd.z);
        }
    }
    hipLaunchKernelGGL((k_rb_vec<MODE, 4>), grd, dim3(WAVE, StencilShape<float>::TJV), 0,
                       c->stream, xn, xc, y, c->c32[0], c->c32[1], c->c32[2], sgeo(c), c->partials,
                       done);
    return (int)(grd.x * grd.y * grd.z);
}

static void launch_reconstruct_add(ls_ctx* c, const float* e, const float* r, float* w,
                                   double scale) {
    const int wd = rb_width(c);
    if (wd == 0) {
        launch_rb<4>(c, e, r, w, nullptr);
        hipLaunchKernelGGL(k_refine_add, dim3(ew_blocks((uint32_t)c->N)), dim3(EW_BLOCK), 0,
                           c->stream, c->x64, w, (uint32_t)c->N, scale);
        return;
    }
    const dim3 grd = rb_grid(c, wd);
    if (wd == 2)
        hipLaunchKernelGGL((k_rb_vec<6, 2>), grd, dim3(WAVE, StencilShape<float>::TJV), 0,
                           c->stream, e, r, nullptr, c->c32[0], c->c32[1], c->c32[2], sgeo(c),
                           c->partials, nullptr, c->x64, scale);
    else
        hipLaunchKernelGGL((k_rb_vec<6, 4>), grd, dim3(WAVE, StencilShape<float>::TJV), 0,
                           c->stream, e, r, nullptr, c->c32[0], c->c32[1], c->c32[2], sgeo(c),
                           c->partials, nullptr, c->x64, scale);
}

// Fused Schur apply on compacted vectors: block shape 64 x WR x (TJ + 2)
// with WR waves per row (256 compacted values each); needs nk % 8 == 0 so a
// lane's 4 compacted values are 16 B aligned.
struct SchurShape {
    int wr, tj, nt;
    dim3 grid, block;
    size_t lds;
};

static bool schur_shape(ls_ctx* c, SchurShape* sh) {
    if (c->nk % (2 * SCHUR_VW) != 0) return false;
    const int nh = c->nk / 2;
    const int wr = (nh + SCHUR_VW * WAVE - 1) / (SCHUR_VW * WAVE);
    const int nt = wr == 1 ? LS_SCHUR_THREADS : SCHUR_WIDE_THREADS;
    const int ns_max = (nt / WAVE) / wr;
    const int tj = std::min(LS_SCHUR_ROWS, ns_max - 2);
    if (tj < 1) return false;
    sh->wr = wr;
    sh->tj = tj;
    sh->nt = nt;
    sh->block = dim3(WAVE, wr, tj + 2);
    sh->lds = (size_t)3 * (tj + 2) * wr * SCHUR_VW * WAVE * sizeof(float);
    sh->grid = dim3((c->nj + tj - 1) / tj, (c->ni + SCHUR_ICH - 1) / SCHUR_ICH, 1);
    return (size_t)sh->grid.x * sh->grid.y <= (size_t)MAX_PARTIALS &&
           2 * (size_t)vec_blocks<float>((uint32_t)(c->N / 2)) <= (size_t)MAX_PARTIALS;
}

// The mixed solver's Schur form: 0 none, 1 fused on compacted vectors,
// 2 two passes on full-length vectors.
static int schur_form(ls_ctx* c) {
    if (!c->P.mixed_schur) return 0;
    if (!use_vec<float>(c)) return 2;     // one voxel per thread (k_rb_scalar)
    SchurShape sh;
    if ((c->P.mixed_schur == 1 || c->P.mixed_schur == 3) && schur_shape(c, &sh)) {
        // 1 takes the fused form only where a row fills at least 3/8 of the
        // lanes of its waves (each covers SCHUR_VW * WAVE compacted values):
        // cold solves on smooth cubes measured two-pass 14-32 % faster at
        // 64^3-128^3 (rows 1/8-1/4 full), fused 27-38 % faster at 192^3 and
        // 256^3 (3/8 and 1/2); 3 takes it wherever it applies
        const int lanes = sh.wr * SCHUR_VW * WAVE;
        if (c->P.mixed_schur == 3 || 8 * (c->nk / 2) >= 3 * lanes) return 1;
    }
    return 2;
}

// q = S p on compacted black vectors; returns the number of block partials
// of p.q.
static int launch_schur(ls_ctx* c, const float* p, float* q, const int* done) {
    SchurShape sh;
    schur_shape(c, &sh);
    if (sh.nt == SCHUR_WIDE_THREADS)
        hipLaunchKernelGGL((k_schur_apply<SCHUR_ICH, SCHUR_VW, SCHUR_WIDE_THREADS>), sh.grid,
                           sh.block, sh.lds, c->stream, p, q, c->c32[0], c->c32[1], c->c32[2],
                           rbgeo(c), c->partials, done);
    else
        hipLaunchKernelGGL((k_schur_apply<SCHUR_ICH, SCHUR_VW>), sh.grid, sh.block, sh.lds,
                           c->stream, p, q, c->c32[0], c->c32[1], c->c32[2], rbgeo(c),
                           c->partials, done);
    return (int)(sh.grid.x * sh.grid.y);
}

static void launch_unpack(ls_ctx* c, const float* hc, float* y) {
    const uint32_t N = (uint32_t)c->N;
    const uint32_t nt = N / 4;
    hipLaunchKernelGGL(k_rb_unpack, dim3((nt + EW_BLOCK - 1) / EW_BLOCK), dim3(EW_BLOCK), 0,
                       c->stream, hc, y, sgeo(c), N);
}

__global__ void k_set_tol2(CgScalars* cs, double tol2) { cs->tol2 = tol2; }

// CG iterations on the Schur complement S = I - N_br N_rb.  fused: the
// vectors are compacted black vectors of N/2 values and q = S p is one
// k_schur_apply; else they are full-length with zero red entries and S p
// takes the gather (into w) and apply passes.  e += alpha p rides in the
// direction kernel, as in cg_iterations.
static int cg_iterations_rb(ls_ctx* c, bool fused, float* e, float* rs, float* p, float* q,
                            float* w, int batch) {
    const uint32_t Nv = fused ? (uint32_t)(c->N / 2) : (uint32_t)c->N;
    const bool vec = ew_vec<float>(Nv);
    const int nbe = vec ? vec_blocks<float>(Nv) : ew_blocks(Nv);
    for (int b = 0; b < batch; ++b) {
        const bool sample = c->profiling && b == 0;
        if (sample) LS_HIP(hipEventRecord(c->ev[12], c->stream));
        int nbs;
        if (fused) {
            nbs = launch_schur(c, p, q, &c->cg->done);               // q = S p, p.q
        } else {
            launch_rb<1>(c, p, nullptr, w, &c->cg->done);             // w = N_rb p
        }
        if (sample) {
            LS_HIP(hipEventRecord(c->ev[13], c->stream));
            c->prof_pending = 16;
            c->prof_schur_bytes = (double)c->N * (fused ? 16.0 : 20.0);
        }
        if (!fused) nbs = launch_rb<2>(c, w, p, q, &c->cg->done);    // q = S p, p.q
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1024), 0, c->stream, c->partials, nbs, c->cg);
        if (vec)
            hipLaunchKernelGGL((k_cg_update_vec<float, false>), dim3(nbe), dim3(EW_BLOCK), 0,
                               c->stream, e, rs, p, q, nullptr, c->cg, Nv, c->partials);
        else
            hipLaunchKernelGGL((k_cg_update<float, false>), dim3(nbe), dim3(EW_BLOCK), 0,
                               c->stream, e, rs, p, q, nullptr, c->cg, Nv, c->partials);
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe, c->cg, 0);
        if (vec)
            hipLaunchKernelGGL(k_cg_direction_vec<float>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream,
                               p, rs, e, c->cg, Nv);
        else
            hipLaunchKernelGGL(k_cg_direction<float>, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, p,
                               rs, e, c->cg, Nv);
    }
    LS_LAUNCH_CHECK();
    return 0;
}

// Mixed precision with reliable updates, the FP32 inner iteration being CG
// on the red-black Schur complement.  Each replacement cycle starts from the
// FP32 copy r of the FP64 residual: the black right-hand side is
// r_b + N_br r_r, CG on S runs until the Schur residual has fallen by delta
// (or its norm reaches the tolerance), then the red correction is
// e_r = r_r + N_rb e_b, the full correction is added to the FP64 iterate and
// the FP64 residual recomputed.  Two-pass: k_rb_vec when nk is even (4 or 2
// floats per lane), else k_rb_scalar; fused (nk % 8 == 0): the inner vectors
// are compacted (k_schur_apply).
static int solve_mixed_rb(ls_ctx* c, ls_step_report* rep, bool fused) {
    const uint32_t N = (uint32_t)c->N;
    const size_t Nv = fused ? c->N / 2 : c->N;
    const int nbe = ew_blocks(N);
    InnerBatch ib(batch_size(c));
    const double delta = c->P.mixed_inner_rtol > 0 ? c->P.mixed_inner_rtol : 0.1;
    float* rs = (float*)c->pw[0];       // Schur residual (FP32, in the pressure scratch)
    float* w = (float*)c->pw[1];
    int st;
    if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
    if ((st = poll_cg(c))) return st;
    const double bnorm2 = c->h_cg->bnorm2;
    const double tol2 = c->h_cg->tol2;
    rep->press_bnorm = sqrt(bnorm2);
    rep->press_relres = c->h_cg->relres;
    if (!rep->press_outer) rep->press_relres0 = c->h_cg->relres;
    if (bnorm2 == 0.0) {
        LS_HIP(hipMemsetAsync(c->x64, 0, c->N * sizeof(double), c->stream));
        rep->press_relres = 0.0;
        return 0;
    }
    if (c->h_cg->converged) return 0;
    const double scale = std::ldexp(1.0, std::ilogb(sqrt(c->h_cg->rr)));
    double rr64 = c->h_cg->rr, wrr = c->h_cg->wrr;
    double last_wrr = wrr;
    int stalls = 0;
    bool first = true;
    const auto t_start = std::chrono::steady_clock::now();
    double rr_start32 = 0.0;
    c->race_tmp.clear();
    auto record = [&](double rr32) {
        if (!(rr32 > 0.0) || !(rr_start32 > 0.0)) return;
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t_start).count();
        c->race_tmp.emplace_back(ms, 0.5 * std::log10(rr_start32 / rr32));
    };
    const int maxit = std::min(c->P.press_maxit, std::max(2000, 3 * (c->P.press_maxit / 4)));
    while (true) {
        // r (FP32 copy of the FP64 residual), e = 0; Schur right-hand side
        hipLaunchKernelGGL(k_inner_start, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->r64, c->r32,
                           c->p32, c->x32, N, 1.0 / scale, 0, c->partials);
        const int nbs = fused ? launch_rb<5>(c, c->r32, c->r32, rs, nullptr)
                              : launch_rb<3>(c, c->r32, c->r32, rs, nullptr);
        hipLaunchKernelGGL(k_inner_init, dim3(1), dim3(1024), 0, c->stream, c->partials, nbs,
                           c->cg, 0.0, maxit, first ? 1 : 0);
        LS_LAUNCH_CHECK();
        if (first) {
            LS_HIP(hipMemcpyAsync(c->p32, rs, Nv * sizeof(float), hipMemcpyDeviceToDevice,
                                  c->stream));
        } else if (ew_vec<float>((uint32_t)Nv)) {   // xlive = 0 here: only the direction
            hipLaunchKernelGGL(k_cg_direction_vec<float>, dim3(vec_blocks<float>((uint32_t)Nv)),
                               dim3(EW_BLOCK), 0, c->stream, c->p32, rs, c->x32, c->cg,
                               (uint32_t)Nv);                  // carries over
            LS_LAUNCH_CHECK();
        } else {
            hipLaunchKernelGGL(k_cg_direction<float>, dim3(ew_blocks((uint32_t)Nv)),
                               dim3(EW_BLOCK), 0, c->stream, c->p32, rs, c->x32, c->cg,
                               (uint32_t)Nv);
            LS_LAUNCH_CHECK();
        }
        if ((st = poll_cg(c))) return st;
        // trigger: the Schur residual fell by delta, or its estimated
        // unscaled norm reaches the tolerance (the Schur residual is the
        // full residual once the red values are reconstructed)
        const double rrs = c->h_cg->rr;
        if (first) rr_start32 = rrs;
        const double w2 = wrr / rr64;
        double trig = delta * delta * rrs;
        const double t2 = 0.81 * tol2 / (w2 * scale * scale);
        if (t2 > trig) trig = t2;
        hipLaunchKernelGGL(k_set_tol2, dim3(1), dim3(1), 0, c->stream, c->cg, trig);
        LS_LAUNCH_CHECK();
        first = false;
        if (!(rrs <= trig)) {
            ib.start(rrs, c->h_cg->iter);
            int nb = ib.next(rrs, c->h_cg->iter, trig);
            while (!c->h_cg->done) {
                if ((st = cg_iterations_rb(c, fused, c->x32, rs, c->p32, c->ap32, w, nb)))
                    return st;
                if ((st = poll_cg(c))) return st;
                record(c->h_cg->rr);
                nb = ib.next(c->h_cg->rr, c->h_cg->iter, trig);
            }
        }
        const int iters = c->h_cg->iter;
        const bool broke = !c->h_cg->converged && iters < maxit && !(rrs <= trig);
        // red reconstruction e_r = r_r + N_rb e_b into w, then x64 += scale w
        // (compacted e: expanded into ap32 first, q being dead here)
        if (fused) {
            launch_unpack(c, c->x32, c->ap32);
            launch_reconstruct_add(c, c->ap32, c->r32, w, scale);
        } else {
            launch_reconstruct_add(c, c->x32, c->r32, w, scale);
        }
        LS_LAUNCH_CHECK();
        rep->press_outer += 1;
        if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
        if ((st = poll_cg(c))) return st;
        rr64 = c->h_cg->rr;
        wrr = c->h_cg->wrr;
        rep->press_relres = c->h_cg->relres;
        rep->press_iters = iters;
        record(rr64 / (scale * scale));
        if (c->h_cg->converged) {
            c->race.swap(c->race_tmp);
            c->have_rate = true;
            return 0;
        }
        stalls = (wrr > 0.25 * last_wrr) ? stalls + 1 : 0;
        last_wrr = wrr;
        if (broke || stalls >= 3 || iters >= maxit) {
            rep->press_fallback = 1;
            const int left = c->P.press_maxit - iters;
            if (left <= 0) return 1;
            return solve_plain64(c, rep, left);
        }
        hipLaunchKernelGGL(k_set_iter, dim3(1), dim3(1), 0, c->stream, c->cg, iters);
        LS_LAUNCH_CHECK();
    }
}

// The mixed solver of the context: red-black Schur inner CG when enabled and
// the vectorised layout applies, else inner CG on the operator.
static int solve_mixed_any(ls_ctx* c, ls_step_report* rep) {
    const int form = schur_form(c);
    if (form) return solve_mixed_rb(c, rep, form == 1);
    return solve_mixed(c, rep);
}

static double effective_top_pressure(double top, double t) {
    const double ramp_duration = 3600.0;
    double eff = top;
    if (t < ramp_duration) eff = top * (t / ramp_duration) + 1e-6;
    return eff;
}

static int solve_learned(ls_ctx* c, ls_step_report* rep);
static int warm_start(ls_ctx* c, double ptop, ls_step_report* rep);
static int warm_push(ls_ctx* c, double ptop, bool replace);

int solve_pressure(ls_ctx* c, double t, ls_step_report* rep) {
    const double ptop = effective_top_pressure(c->P.top_pressure, t);
    int st = assemble_pressure(c, t);
    if (st) return st;
    if ((st = warm_start(c, ptop, rep))) return st;
    if (c->P.solver == LS_SOLVER_PLAIN64) st = solve_plain64(c, rep, c->P.press_maxit);
    else if (c->P.solver == LS_SOLVER_MIXED) st = solve_mixed_any(c, rep);
    else st = solve_learned(c, rep);
    if (st < 0) return st;
    int st2 = finalize_pressure(c, ptop);
    if (st2) return st2;
    // a solve that ended on the projection itself adds no direction: it takes
    // the newest slot instead of a new one, so the history (and the cost of
    // the next projection) does not grow with copies of one span
    const bool same_span = rep->press_iters == 0 && rep->warm_basis > 0 && !rep->warm_net;
    if (st == 0 && (st2 = warm_push(c, ptop, same_span))) return st2;
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

int debug_schur_apply(ls_ctx* c, int form, const float* p, float* q, double* dot) {
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    SchurShape sh;
    if (form == 1 && !schur_shape(c, &sh)) {
        c->err = "the fused Schur apply needs nk % " + std::to_string(2 * SCHUR_VW) + " == 0 and nk <= " + std::to_string(SCHUR_MAX_NK);
        return -3;
    }
    if (form == 2 && !use_vec<float>(c)) {
        c->err = "the compacted two-pass form needs nk % 4 == 0 (form 4 takes any nk)";
        return -3;
    }
    const size_t N = c->N, Nh = N / 2;
    float* pc = (float*)c->pw[0];            // 2N floats per scratch vector
    float* qc = pc + Nh;
    int nbs;
    std::vector<float> full;
    if (form == 4) {
        // full-length vectors, any nk: the gather and apply passes the
        // solver runs (k_rb_vec with 4 or 2 floats per lane for nk % 4 == 0
        // or 2, else k_rb_scalar); the red entries of p are never read,
        // those of q come back zero
        float* pf = (float*)c->pw[1];
        float* wf = pf + N;
        LS_HIP(hipMemcpyAsync(pf, p, N * sizeof(float), hipMemcpyHostToDevice, c->stream));
        launch_rb<1>(c, pf, nullptr, wf, nullptr);
        nbs = launch_rb<2>(c, wf, pf, c->ap32, nullptr);
        LS_LAUNCH_CHECK();
        LS_HIP(hipMemcpyAsync(q, c->ap32, N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        std::vector<double> part(nbs);
        LS_HIP(hipMemcpyAsync(part.data(), c->partials, nbs * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        double a = 0.0;
        for (double x : part) a += x;
        *dot = a;
        return 0;
    }
    LS_HIP(hipMemcpyAsync(pc, p, Nh * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (form == 1) {
        nbs = launch_schur(c, pc, qc, nullptr);
        LS_LAUNCH_CHECK();
        LS_HIP(hipMemcpyAsync(q, qc, Nh * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    } else {
        float* pf = (float*)c->pw[1];
        float* wf = pf + N;
        launch_unpack(c, pc, pf);
        launch_rb<1>(c, pf, nullptr, wf, nullptr);
        nbs = launch_rb<2>(c, wf, pf, c->ap32, nullptr);
        LS_LAUNCH_CHECK();
        full.resize(N);
        LS_HIP(hipMemcpyAsync(full.data(), c->ap32, N * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    }
    std::vector<double> part(nbs);
    LS_HIP(hipMemcpyAsync(part.data(), c->partials, nbs * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    if (form == 2) {
        const size_t nk = (size_t)c->nk, nj = (size_t)c->nj;
        for (size_t v = 0; v < N; ++v) {
            const size_t row = v / nk;
            if (((row / nj + row % nj + v % nk) & 1) == 0) q[v >> 1] = full[v];
        }
    }
    double a = 0.0;
    for (double x : part) a += x;
    *dot = a;
    return 0;
}

int time_stencil(ls_ctx* c, int precision, int reps, double* ms, double* bytes) {
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    SchurShape sh;
    if (precision == LS_TIME_SCHUR && !schur_shape(c, &sh)) {
        c->err = "the fused Schur apply needs nk % " + std::to_string(2 * SCHUR_VW) + " == 0 and nk <= " + std::to_string(SCHUR_MAX_NK);
        return -3;
    }
    hipEvent_t e0 = c->ev[14], e1 = c->ev[15];
    for (int pass = 0; pass < 2; ++pass) {
        int n = pass == 0 ? 1 : reps;
        if (pass == 1) LS_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < n; ++r) {
            if (precision == LS_TIME_RB_GATHER) launch_rb<1>(c, c->p32, nullptr, c->ap32, nullptr);
            else if (precision == LS_TIME_RB_APPLY) launch_rb<2>(c, c->r32, c->p32, c->ap32, nullptr);
            else if (precision == LS_TIME_SCHUR) launch_schur(c, c->p32, c->ap32, nullptr);
            else if (precision == 32) launch_stencil<float>(c, c->p32, c->ap32, c->c32, nullptr);
            else launch_stencil<double>(c, c->p64, c->ap64, c->c64, nullptr);
        }
        LS_LAUNCH_CHECK();
    }
    LS_HIP(hipEventRecord(e1, c->stream));
    LS_HIP(hipEventSynchronize(e1));
    float f = 0.0f;
    LS_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = (double)f / reps;
    if (precision == LS_TIME_RB_GATHER) {
        *bytes = (double)c->N * 20.0;      // p, 3 faces, w
    } else if (precision == LS_TIME_RB_APPLY) {
        *bytes = (double)c->N * 24.0;      // w, 3 faces, p, q
    } else if (precision == LS_TIME_SCHUR) {
        *bytes = (double)c->N * 16.0;      // 3 faces; p, q compacted (half length)
    } else {
        const double sv = precision == 32 ? 4.0 : 8.0;
        *bytes = (double)c->N * (2.0 * sv + 3.0 * sv);
    }
    return 0;
}



// ===========================================================================
// Learned solver: flexible CG (Notay 2000, truncation 1) preconditioned by a
// two-level operator whose coarse level is the U-Net.
//
//   M(r) = 2w r - w^2 A^r + g (P e - w A^P e),   e = UNet(R r)
//
// i.e. one damped-Jacobi step (weight w, the scaled diagonal is 1), an
// additive coarse correction P e from the network applied to the residual
// restricted to the 2x coarser grid, with the Galerkin step length
// g = (Pe . (r - w A^r)) / (Pe . A^Pe), then one damped-Jacobi post-smoothing
// step (A^ z is known by linearity, no extra stencil).  R sums the 8 fine
// children of a coarse voxel; P copies the coarse value to the solved fine
// children.  Vectors and stencils are FP32 (dot products FP64); the FP64
// residual replaces the FP32 recurrence as in the mixed solver, and the
// solver falls back to mixed PCG on breakdown or stagnation.
// ===========================================================================

constexpr double kJacobiW = 0.6;

struct CGeo {
    int ni, nj, nk;       // fine grid
    int ci, cj, ck;       // network (coarse, padded) grid
    uint32_t si, sj;
    uint32_t cn;
};

static CGeo cgeo(const ls_ctx* c) {
    CGeo g;
    g.ni = c->ni; g.nj = c->nj; g.nk = c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    g.sj = (uint32_t)c->nk;
    g.ci = c->net_dims[0]; g.cj = c->net_dims[1]; g.ck = c->net_dims[2];
    g.cn = (uint32_t)g.ci * g.cj * g.ck;
    return g;
}

// Restriction and the static input channels.  One thread per coarse voxel.
// ch0 (written later, scaled): sum of the 8 children's residual; ch1: fraction
// of solved children; ch2-4: mean +face conductances; ch5-7: mean -face
// conductances; ch8-15: zero.
__global__ void k_lpack(const float* __restrict__ r, const int8_t* __restrict__ cls,
                        const float* __restrict__ cx, const float* __restrict__ cy,
                        const float* __restrict__ cz, float* __restrict__ rc,
                        __bf16* __restrict__ xin, CGeo g, double* __restrict__ partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double ss = 0.0;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < g.cn; q += gridDim.x * blockDim.x) {
        const uint32_t kq = q % g.ck, jq = (q / g.ck) % g.cj, iq = q / ((uint32_t)g.ck * g.cj);
        float rs = 0.f, m = 0.f, f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) {
                    const int i = 2 * iq + a, j = 2 * jq + b, k = 2 * kq + d;
                    if (i >= g.ni || j >= g.nj || k >= g.nk) continue;
                    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
                    rs += r[v];
                    m += cls[v] == 1 ? 1.f : 0.f;
                    f[0] += cx[v];
                    f[1] += cy[v];
                    f[2] += cz[v];
                    f[3] += i > 0 ? cx[v - g.si] : 0.f;
                    f[4] += j > 0 ? cy[v - g.sj] : 0.f;
                    f[5] += k > 0 ? cz[v - 1] : 0.f;
                }
        rc[q] = rs;
        ss += (double)rs * rs;
        __bf16* o = xin + (size_t)q * 16;
        o[1] = (__bf16)(m * 0.125f);
        for (int e = 0; e < 6; ++e) o[2 + e] = (__bf16)(f[e] * 0.125f);
        for (int e = 8; e < 16; ++e) o[e] = (__bf16)0.f;
    }
    double a[1] = {ss};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

// scale of ch0: 1 / rms of the restricted residual over the coarse grid
__global__ void k_lpack_scale(const double* __restrict__ partials, int n, uint32_t cn,
                              FcgScalars* fs) {
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) fs->scale_c = a[0] > 0.0 ? 1.0 / sqrt(a[0] / cn) : 0.0;
}

__global__ void k_lpack_ch0(const float* __restrict__ rc, __bf16* __restrict__ xin, uint32_t cn,
                            const FcgScalars* __restrict__ fs) {
    const float sc = (float)fs->scale_c;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < cn; q += gridDim.x * blockDim.x)
        xin[(size_t)q * 16] = (__bf16)(rc[q] * sc);
}

// P: w[v] = e[parent(v)] on solved voxels, 0 elsewhere
__global__ void k_lprolong(const float* __restrict__ out, const int8_t* __restrict__ cls,
                           float* __restrict__ w, CGeo g) {
    const uint32_t N = (uint32_t)g.ni * g.nj * g.nk;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const uint32_t k = v % g.nk, j = (v / g.nk) % g.nj, i = v / g.si;
        const uint32_t q = ((i >> 1) * g.cj + (j >> 1)) * g.ck + (k >> 1);
        w[v] = cls[v] == 1 ? out[(size_t)q * 16] : 0.f;
    }
}

// dots (w . (r - om sr)) and (w . sw)
__global__ void k_ldots(const float* __restrict__ w, const float* __restrict__ r,
                        const float* __restrict__ sr, const float* __restrict__ sw, float om,
                        uint32_t N, const FcgScalars* __restrict__ fs, double* __restrict__ partials) {
    if (fs->done) return;
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    double a0 = 0.0, a1 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const double wv = w[v];
        a0 += wv * (double)(r[v] - om * sr[v]);
        a1 += wv * (double)sw[v];
    }
    double q[2] = {a0, a1};
    block_reduce<2, RED_SUM, RED_SUM>(q, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = q[0];
        partials[gridDim.x + blockIdx.x] = q[1];
    }
}

__global__ void k_lgamma(const double* __restrict__ partials, int n, FcgScalars* fs) {
    if (fs->done) return;
    __shared__ double lds[2 * 1024 / WAVE];
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        a0 += partials[i];
        a1 += partials[n + i];
    }
    double a[2] = {a0, a1};
    block_reduce<2, RED_SUM, RED_SUM>(a, lds);
    if (threadIdx.x == 0) fs->gamma = (a[1] > 0.0 && isfinite(a[0] / a[1])) ? a[0] / a[1] : 0.0;
}

// z = 2 om r - om^2 sr + g (w - om sw); partial z . qprev
__global__ void k_lz(const float* __restrict__ r, const float* __restrict__ sr,
                     const float* __restrict__ w, const float* __restrict__ sw,
                     const float* __restrict__ qprev, float* __restrict__ z, float om, uint32_t N,
                     const FcgScalars* __restrict__ fs, double* __restrict__ partials) {
    if (fs->done) return;
    __shared__ double lds[EW_BLOCK / WAVE];
    const float g = (float)fs->gamma;
    double a0 = 0.0;
    const bool first = fs->first;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const float zv = 2.f * om * r[v] - om * om * sr[v] + g * (w[v] - om * sw[v]);
        z[v] = zv;
        if (!first) a0 += (double)zv * (double)qprev[v];
    }
    double a[1] = {a0};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

__global__ void k_lbeta(const double* __restrict__ partials, int n, FcgScalars* fs) {
    if (fs->done) return;
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) fs->beta = (fs->first || !(fs->pq_prev > 0.0)) ? 0.0 : a[0] / fs->pq_prev;
}

// p = z - beta pprev; partial p . r
__global__ void k_lp(const float* __restrict__ z, const float* __restrict__ pprev,
                     float* __restrict__ p, const float* __restrict__ r, uint32_t N,
                     const FcgScalars* __restrict__ fs, double* __restrict__ partials) {
    if (fs->done) return;
    __shared__ double lds[EW_BLOCK / WAVE];
    const float b = (float)fs->beta;
    double a0 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        // the first direction reads no pprev (its vector holds no earlier value)
        const float pv = b != 0.0f ? z[v] - b * pprev[v] : z[v];
        p[v] = pv;
        a0 += (double)pv * (double)r[v];
    }
    double a[1] = {a0};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

__global__ void k_lpr(const double* __restrict__ partials, int n, FcgScalars* fs) {
    if (fs->done) return;
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) fs->pr = a[0];
}

// alpha = p.r / p.q  (p.q from the stencil partials)
__global__ void k_lalpha(const double* __restrict__ partials, int n, FcgScalars* fs, CgScalars* cs) {
    if (fs->done) return;
    __shared__ double lds[1024 / WAVE];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        const double pq = a[0];
        fs->pq = pq;
        if (!(pq > 0.0) || !isfinite(pq) || !isfinite(fs->pr)) {
            fs->done = 1;
            fs->broke = 1;
            cs->done = 1;
        } else {
            fs->alpha = fs->pr / pq;
            cs->alpha = fs->alpha;   // consumed by the shared update kernel
            cs->done = 0;
        }
    }
}

// after the update: r.r, trigger for the FP64 replacement, bookkeeping
__global__ void k_lcheck(const double* __restrict__ partials, int n, FcgScalars* fs) {
    if (fs->done) return;
    __shared__ double lds[2 * 1024 / WAVE];
    double a0 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a0 += partials[i];
    double a[1] = {a0};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        const double rr = a[0];
        fs->iter += 1;
        fs->rr_prev = fs->rr;
        fs->rr = rr;
        fs->pq_prev = fs->pq;
        fs->first = 0;
        if (!isfinite(rr)) {
            fs->done = 1;
            fs->broke = 1;
        } else if (rr <= fs->tol2 || fs->iter >= fs->maxit) {
            fs->done = 1;
        }
    }
}

__global__ void k_lstart(FcgScalars* fs, double rr, double trig, int maxit, int first) {
    fs->rr = rr;
    fs->tol2 = trig;
    fs->maxit = maxit;
    fs->done = 0;
    fs->broke = 0;
    if (first) {
        fs->iter = 0;
        fs->first = 1;
        fs->pq_prev = 0.0;
    }
}

int learned_setup(ls_ctx* c) {
    if (c->net != nullptr) return 0;
    if (c->net_weights.empty()) {
        c->err = "learned solver: no network weights (ls_set_network)";
        return -20;
    }
    auto up16 = [](int n) { int h = (n + 1) / 2; return ((h + 15) / 16) * 16; };
    c->net_dims[0] = up16(c->ni);
    c->net_dims[1] = up16(c->nj);
    c->net_dims[2] = up16(c->nk);
    c->net = new UNet();
    int st = unet_alloc(c->net, c->net_dims[0], c->net_dims[1], c->net_dims[2], 1, false, c->stream);
    if (st) {
        c->err = "learned solver: " + c->net->err;
        return st;
    }
    const size_t cn = (size_t)c->net_dims[0] * c->net_dims[1] * c->net_dims[2];
    if (hipMalloc(&c->rc, cn * sizeof(float)) != hipSuccess) {
        c->err = "learned solver: hipMalloc";
        return -12;
    }
    c->grid_bytes += c->net->bytes + cn * sizeof(float);
    if (c->net_weights.size() != c->net->nw) {
        c->err = "learned solver: weight count mismatch";
        return -3;
    }
    LS_HIP(hipMemcpyAsync(c->net->w, c->net_weights.data(), c->net->nw * sizeof(float),
                          hipMemcpyHostToDevice, c->stream));
    st = unet_convert_weights(c->net);
    if (st) return st;
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

void learned_free(ls_ctx* c) {
    if (c->net) {
        if (c->rc) c->grid_bytes -= c->net->bytes + (size_t)c->net_dims[0] * c->net_dims[1] *
                                                        c->net_dims[2] * sizeof(float);
        unet_free(c->net);
        delete c->net;
        c->net = nullptr;
    }
    if (c->rc) (void)hipFree(c->rc);
    c->rc = nullptr;
}

// z = M(r): restriction, U-Net, prolongation, Galerkin step, smoothing
static int learned_precond(ls_ctx* c, const float* r, float* z, const float* qprev, float* sr,
                           float* w, float* sw) {
    const uint32_t N = (uint32_t)c->N;
    const int nbe = ew_blocks(N);
    CGeo g = cgeo(c);
    const float om = (float)kJacobiW;
    launch_stencil<float>(c, r, sr, c->c32, &c->fcg->done);
    const int nbc = ew_blocks(g.cn);
    hipLaunchKernelGGL(k_lpack, dim3(nbc), dim3(EW_BLOCK), 0, c->stream, r, c->cls, c->c32[0],
                       c->c32[1], c->c32[2], c->rc, c->net->x, g, c->partials);
    hipLaunchKernelGGL(k_lpack_scale, dim3(1), dim3(1024), 0, c->stream, c->partials, nbc, g.cn,
                       c->fcg);
    hipLaunchKernelGGL(k_lpack_ch0, dim3(nbc), dim3(EW_BLOCK), 0, c->stream, c->rc, c->net->x, g.cn,
                       c->fcg);
    int st = unet_forward(c->net, 1);
    if (st) {
        c->err = c->net->err;
        return st;
    }
    hipLaunchKernelGGL(k_lprolong, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->net->out, c->cls, w, g);
    launch_stencil<float>(c, w, sw, c->c32, &c->fcg->done);
    hipLaunchKernelGGL(k_ldots, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, w, r, sr, sw, om, N, c->fcg,
                       c->partials);
    hipLaunchKernelGGL(k_lgamma, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe, c->fcg);
    hipLaunchKernelGGL(k_lz, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, r, sr, w, sw, qprev, z, om, N,
                       c->fcg, c->partials);
    hipLaunchKernelGGL(k_lbeta, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe, c->fcg);
    LS_LAUNCH_CHECK();
    return 0;
}

__global__ void k_lk_reset(FcgScalars* fs) { fs->done = 0; }

__global__ void k_bf16_to_f32(const __bf16* __restrict__ a, float* __restrict__ b, uint32_t n) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        b[e] = (float)a[e];
}

int learned_pack_debug(ls_ctx* c, const float* r_dev, float* out_dev) {
    int st = learned_setup(c);
    if (st) return st;
    CGeo g = cgeo(c);
    const int nbc = ew_blocks(g.cn);
    hipLaunchKernelGGL(k_lk_reset, dim3(1), dim3(1), 0, c->stream, c->fcg);
    hipLaunchKernelGGL(k_lpack, dim3(nbc), dim3(EW_BLOCK), 0, c->stream, r_dev, c->cls, c->c32[0],
                       c->c32[1], c->c32[2], c->rc, c->net->x, g, c->partials);
    hipLaunchKernelGGL(k_lpack_scale, dim3(1), dim3(1024), 0, c->stream, c->partials, nbc, g.cn,
                       c->fcg);
    hipLaunchKernelGGL(k_lpack_ch0, dim3(nbc), dim3(EW_BLOCK), 0, c->stream, c->rc, c->net->x, g.cn,
                       c->fcg);
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(ew_blocks(g.cn * 16)), dim3(EW_BLOCK), 0, c->stream,
                       c->net->x, out_dev, g.cn * 16);
    LS_LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------
// Training data of the learned coarse correction, built on the device.
// ---------------------------------------------------------------------------
// A_c = P^T A^ P on the network grid, one thread per coarse voxel:
// ac[q] = solved children - 2 x the scaled faces inside the aggregate;
// ac[(1 + a) cn + q] = - the faces from the aggregate into its +a neighbour
// (a stored face is 0 unless both of its voxels are solved rows).
__global__ void k_lgalerkin(const int8_t* __restrict__ cls, const double* __restrict__ cx,
                            const double* __restrict__ cy, const double* __restrict__ cz, CGeo g,
                            float* __restrict__ ac) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < g.cn; q += gridDim.x * blockDim.x) {
        const uint32_t kq = q % g.ck, jq = (q / g.ck) % g.cj, iq = q / ((uint32_t)g.ck * g.cj);
        double d = 0.0, o[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e) {
                    const int i = 2 * iq + a, j = 2 * jq + b, k = 2 * kq + e;
                    if (i >= g.ni || j >= g.nj || k >= g.nk) continue;
                    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
                    d += cls[v] == 1 ? 1.0 : 0.0;
                    const int side[3] = {a, b, e};
                    const bool nxt[3] = {i + 1 < g.ni, j + 1 < g.nj, k + 1 < g.nk};
                    const double f[3] = {cx[v], cy[v], cz[v]};
                    for (int ax = 0; ax < 3; ++ax) {
                        if (!nxt[ax]) continue;
                        if (side[ax] == 0) d -= 2.0 * f[ax];
                        else o[ax] -= f[ax];
                    }
                }
        ac[q] = (float)d;
        ac[g.cn + q] = (float)o[0];
        ac[2 * (size_t)g.cn + q] = (float)o[1];
        ac[3 * (size_t)g.cn + q] = (float)o[2];
    }
}

// e = R (sd (p* - p)) in the scaled variables
__global__ void k_lsample_e(const double* __restrict__ pstar, const double* __restrict__ p,
                            const double* __restrict__ sd, const int8_t* __restrict__ cls,
                            double* __restrict__ e, uint32_t N) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        e[v] = cls[v] == 1 ? sd[v] * (pstar[v] - p[v]) : 0.0;
}

// r masked to the solved rows (FP32 copy for the restriction), partial e.r
__global__ void k_lsample_r(const double* __restrict__ e, const double* __restrict__ r,
                            const int8_t* __restrict__ cls, float* __restrict__ r32, uint32_t N,
                            double* __restrict__ partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double acc = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const double rv = cls[v] == 1 ? r[v] : 0.0;
        r32[v] = (float)rv;
        acc += e[v] * rv;
    }
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

int learned_galerkin_debug(ls_ctx* c, float* ac_dev) {
    int st = learned_setup(c);
    if (st) return st;
    CGeo g = cgeo(c);
    hipLaunchKernelGGL(k_lgalerkin, dim3(ew_blocks(g.cn)), dim3(EW_BLOCK), 0, c->stream, c->cls,
                       c->c64[0], c->c64[1], c->c64[2], g, ac_dev);
    LS_LAUNCH_CHECK();
    return 0;
}

// One sample of the energy loss from the current pressure field p (an
// iterate) and the reference solution p* (device, unscaled): the network
// input of r = A^ e, e = R sd (p* - p) (into x_dev, [cn][16]), rc = P^T r
// (rc_dev) and e.r = e.A^e.  Uses the solver's FP64 vectors as scratch.
int learned_sample_debug(ls_ctx* c, const double* pstar_dev, float* x_dev, float* rc_dev,
                         double* eae) {
    int st = learned_setup(c);
    if (st) return st;
    const uint32_t N = (uint32_t)c->N;
    const int nbe = ew_blocks(N);
    hipLaunchKernelGGL(k_lsample_e, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, pstar_dev, c->p,
                       c->sd, c->cls, c->x64, N);
    launch_stencil<double>(c, c->x64, c->r64, c->c64, nullptr);
    hipLaunchKernelGGL(k_lsample_r, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->x64, c->r64,
                       c->cls, c->r32, N, c->partials);
    LS_LAUNCH_CHECK();
    std::vector<double> part(nbe);
    LS_HIP(hipMemcpyAsync(part.data(), c->partials, nbe * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    double a = 0.0;
    for (double x : part) a += x;
    *eae = a;
    if ((st = learned_pack_debug(c, c->r32, x_dev))) return st;
    const CGeo g = cgeo(c);
    LS_HIP(hipMemcpyAsync(rc_dev, c->rc, g.cn * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

static int poll_fcg(ls_ctx* c) {
    LS_HIP(hipMemcpyAsync(c->h_fcg, c->fcg, sizeof(FcgScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Mixed-precision PCG inside the learned solver (a converged solve leaves
// its residual curve for the learned solver's race).
static int mixed_recorded(ls_ctx* c, ls_step_report* rep) {
    ls_step_report r2;
    std::memset(&r2, 0, sizeof(r2));
    const int st = solve_mixed_any(c, &r2);   // records the race curve
    if (!rep->press_outer && !rep->press_iters) {
        rep->press_relres0 = r2.press_relres0;
        rep->press_bnorm = r2.press_bnorm;
    }
    rep->press_iters += r2.press_iters;
    rep->press_outer += r2.press_outer;
    rep->press_relres = r2.press_relres;
    return st;
}

// Decades the recorded mixed solve had reached after `ms` (linear between
// the recorded points; its last value past the end).
static double race_decades(const std::vector<std::pair<double, double>>& r, double ms) {
    if (r.empty()) return 0.0;
    if (ms <= r.front().first) return r.front().second * (ms / std::max(r.front().first, 1e-9));
    for (size_t q = 1; q < r.size(); ++q)
        if (ms <= r[q].first) {
            const double f = (ms - r[q - 1].first) / std::max(r[q].first - r[q - 1].first, 1e-9);
            return r[q - 1].second + f * (r[q].second - r[q - 1].second);
        }
    return r.back().second;
}

static int solve_learned(ls_ctx* c, ls_step_report* rep) {
    int st = learned_setup(c);
    if (st) return st;
    // The plain method sets the bar: the network runs only against a known
    // convergence rate of mixed PCG on this context (the first solve sets
    // it), and after repeated losses it is rested for a few solves.
    const bool raced = c->P.learned_race > 0.0;
    if ((raced && !c->have_rate) || c->net_skip > 0) {
        if (c->net_skip > 0) --c->net_skip;
        rep->press_fallback = 1;
        return mixed_recorded(c, rep);
    }
    const uint32_t N = (uint32_t)c->N;
    const int nbe = ew_blocks(N);
    // FP32 work vectors in the pressure scratch, two per N-double vector
    float* z = (float*)c->pw[0];
    float* pa = z + N;
    float* pb = (float*)c->pw[1];
    float* qa = pb + N;
    float* qb = (float*)c->pw[2];
    float* sr = qb + N;
    float* w = (float*)c->pw[3];
    float* sw = w + N;
    double* xbackup = c->pw[4];
    if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
    if ((st = poll_cg(c))) return st;
    const double bnorm2 = c->h_cg->bnorm2;
    const double tol2 = c->h_cg->tol2;
    rep->press_bnorm = sqrt(bnorm2);
    rep->press_relres = c->h_cg->relres;
    if (!rep->press_outer) rep->press_relres0 = c->h_cg->relres;
    if (bnorm2 == 0.0) {
        LS_HIP(hipMemsetAsync(c->x64, 0, c->N * sizeof(double), c->stream));
        rep->press_relres = 0.0;
        return 0;
    }
    if (c->h_cg->converged) return 0;
    LS_HIP(hipMemcpyAsync(xbackup, c->x64, c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    const double scale = std::ldexp(1.0, std::ilogb(sqrt(c->h_cg->rr)));
    const double delta = c->P.mixed_inner_rtol > 0 ? c->P.mixed_inner_rtol : 0.1;
    double rr64 = c->h_cg->rr, wrr = c->h_cg->wrr, last_wrr = wrr;
    const double wrr_start = wrr;
    const int maxit = c->P.press_maxit < 2000 ? c->P.press_maxit : 2000;
    int stalls = 0, applies = 0;
    bool first = true;
    const double rr_start = rr64 / (scale * scale);
    bool lost = false;
    const auto t_fcg = std::chrono::steady_clock::now();
    while (true) {
        const double rr32 = rr64 / (scale * scale);
        const double w2 = wrr / rr64;
        double trig = delta * delta * rr32;
        const double t2 = 0.81 * tol2 / (w2 * scale * scale);
        if (t2 > trig) trig = t2;
        hipLaunchKernelGGL(k_inner_start, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->r64, c->r32,
                           c->p32, c->x32, N, 1.0 / scale, 0, c->partials);
        hipLaunchKernelGGL(k_lstart, dim3(1), dim3(1), 0, c->stream, c->fcg, rr32, trig, maxit,
                           first ? 1 : 0);
        LS_LAUNCH_CHECK();
        first = false;
        if ((st = poll_fcg(c))) return st;
        while (!c->h_fcg->done) {
            for (int it = 0; it < 2; ++it) {
                // z = M(r); beta = z.qprev / pq_prev
                if ((st = learned_precond(c, c->r32, z, qb, sr, w, sw))) return st;
                ++applies;
                hipLaunchKernelGGL(k_lp, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, z, pb, pa, c->r32,
                                   N, c->fcg, c->partials);
                hipLaunchKernelGGL(k_lpr, dim3(1), dim3(1024), 0, c->stream, c->partials, nbe, c->fcg);
                const int nbs = launch_stencil<float>(c, pa, qa, c->c32, &c->fcg->done);
                hipLaunchKernelGGL(k_lalpha, dim3(1), dim3(1024), 0, c->stream, c->partials, nbs,
                                   c->fcg, c->cg);
                if (ew_vec<float>(N))
                    hipLaunchKernelGGL((k_cg_update_vec<float, true>), dim3(vec_blocks<float>(N)),
                                       dim3(EW_BLOCK), 0, c->stream, c->x32, c->r32, pa, qa, nullptr,
                                       c->cg, N, c->partials);
                else
                    hipLaunchKernelGGL((k_cg_update<float, true>), dim3(nbe), dim3(EW_BLOCK), 0,
                                       c->stream, c->x32, c->r32, pa, qa, nullptr, c->cg, N,
                                       c->partials);
                const int nbu = ew_vec<float>(N) ? vec_blocks<float>(N) : nbe;
                hipLaunchKernelGGL(k_lcheck, dim3(1), dim3(1024), 0, c->stream, c->partials, nbu,
                                   c->fcg);
                LS_LAUNCH_CHECK();
                std::swap(pa, pb);
                std::swap(qa, qb);
            }
            if ((st = poll_fcg(c))) return st;
            // race against the last mixed PCG solve's residual
            // curve; after the same wall time the network must have reduced
            // the residual by at least learned_race x as many decades.  The
            // check repeats at every poll from the 6th iteration on (a
            // preconditioner that stalls after a fast start is caught
            // later; one that is behind is stopped after ~6 applies)
            if (raced && c->h_fcg->iter >= 6 && !c->h_fcg->done) {
                const double ms = std::chrono::duration<double, std::milli>(
                                      std::chrono::steady_clock::now() - t_fcg).count();
                const double dec = 0.5 * std::log10(rr_start / c->h_fcg->rr);
                if (!(dec >= c->P.learned_race * race_decades(c->race, ms))) {
                    lost = true;
                    break;
                }
            }
        }
        const int iters = c->h_fcg->iter;
        const bool broke = c->h_fcg->broke;
        rep->press_iters = iters;
        rep->press_unet_applies = applies;
        if (lost || broke) break;
        hipLaunchKernelGGL(k_refine_add, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->x64, c->x32, N,
                           scale);
        LS_LAUNCH_CHECK();
        rep->press_outer += 1;
        if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
        if ((st = poll_cg(c))) return st;
        rr64 = c->h_cg->rr;
        wrr = c->h_cg->wrr;
        rep->press_relres = c->h_cg->relres;
        if (c->h_cg->converged) {
            c->net_failures = 0;
            return 0;
        }
        stalls = (wrr > 0.25 * last_wrr) ? stalls + 1 : 0;
        last_wrr = wrr;
        if (stalls >= 2 || iters >= maxit) break;
    }
    // fallback: mixed-precision PCG from the saved start (an FCG iterate
    // with a smaller residual can still hold more error in the slowest
    // modes, so its progress is discarded)
    rep->press_fallback = 1;
    (void)wrr_start;
    LS_HIP(hipMemcpyAsync(c->x64, xbackup, c->N * sizeof(double), hipMemcpyDeviceToDevice,
                          c->stream));
    if (++c->net_failures >= 2) {
        c->net_skip = 8;
        c->net_failures = 0;
    }
    const int iters_fcg = rep->press_iters;
    rep->press_iters = 0;
    const int outer_fcg = rep->press_outer;
    rep->press_outer = 0;
    st = mixed_recorded(c, rep);
    rep->press_iters += iters_fcg;
    rep->press_outer += outer_fcg;
    return st;
}


// ---------------------------------------------------------------------------
// Warm start by Galerkin projection onto earlier solutions (Fischer 1998,
// "Projection techniques for iterative solution of Ax = b with successive
// right-hand sides", CMAME 163).  The last m pressure solutions, mapped to
// the scaled variables of the new system (v = sqrt(d) p on solved rows), span
// the space; x0 is the A^-orthogonal projection of the solution onto it,
// x0 = sum_a (q_a . b^) q_a over an A^-orthonormal basis q.  x0 minimises the
// A^-norm of the error over the span: never worse than the previous pressure,
// which is in it, and exact when the operator is unchanged and the right-hand
// side moves inside the span (the linear top-pressure ramp over a fixed
// evaporation term needs two vectors).
//
// The raw vectors are the newest solution and the differences of successive
// ones.  They are orthonormalised by classical Gram-Schmidt in the A^ inner
// product, run twice, with the energy of each orthogonalised vector taken
// from a fresh stencil application: successive solutions are close to
// collinear, and a Gram matrix V^T A^ V of them loses the small directions
// (the top-pressure part is nearly constant, close to the operator's null
// space) to cancellation.  With the network (LS_WARM_LEARNED) the prolonged
// U-Net correction of the projected residual joins the basis the same way, so
// the network can only lower the A^-norm error of x0.

struct WBasis {
    const double* v[LS_WARM_MAX];
};
struct WCoef {
    double c[LS_WARM_MAX];
};

// v = sqrt(d) (p - q)  (q == nullptr: v = sqrt(d) p)
__global__ void k_wscale(const double* __restrict__ p, const double* __restrict__ q,
                         const double* __restrict__ sd, double* __restrict__ v, uint32_t N) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        v[i] = sd[i] * (q ? p[i] - q[i] : p[i]);
}

__global__ void k_wfrom32(const float* __restrict__ a, double* __restrict__ b, uint32_t N) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

// block partials of x . y
__global__ void k_wdot(const double* __restrict__ x, const double* __restrict__ y, uint32_t N,
                       double* __restrict__ partials) {
    __shared__ double lds[EW_BLOCK / WAVE];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        acc += x[i] * y[i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

// out[q] = sum of the nb partials of value q (one block per value, fixed order)
__global__ void k_wsum(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    __shared__ double lds[1024 / WAVE];
    const int q = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[q * nb + i];
    double a[1] = {acc};
    block_reduce<1, RED_SUM>(a, lds);
    if (threadIdx.x == 0) out[q] = a[0];
}

// w -= sum_a C.c[a] B.v[a]
__global__ void k_wsub(WBasis B, WCoef C, int m, double* __restrict__ w, uint32_t N) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double s = w[i];
        for (int a = 0; a < m; ++a) s -= C.c[a] * B.v[a][i];
        w[i] = s;
    }
}

// x (+)= sum_a C.c[a] B.v[a]
__global__ void k_wcombine(WBasis B, WCoef C, int m, double* __restrict__ x, uint32_t N,
                           int accumulate) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double s = accumulate ? x[i] : 0.0;
        for (int a = 0; a < m; ++a) s += C.c[a] * B.v[a][i];
        x[i] = s;
    }
}

// The one-vector projected start and its residual in one pass: x = cc w and
// r = b - cc (A^ w), with A^ w from the projection's stencil pass (p = r as
// well when p is given), and the partials of the residual kernel
// (k_stencil_vec MODE 1): r.r, |sd r|^2, |sd b|^2.  w may be x itself.
__global__ __launch_bounds__(EW_BLOCK) void k_wfinish(const double* w,
                                                      const double* __restrict__ aw,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ sd, double cc,
                                                      double* x, double* __restrict__ r,
                                                      double* __restrict__ p, uint32_t N,
                                                      double* __restrict__ partials) {
    __shared__ double lds[3 * EW_BLOCK / WAVE];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const double wv = w[v], bv = b[v], sv = sd[v];
        const double rv = bv - cc * aw[v];
        __builtin_nontemporal_store(cc * wv, x + v);
        __builtin_nontemporal_store(rv, r + v);
        if (p != nullptr) __builtin_nontemporal_store(rv, p + v);
        const double sr = sv * rv, sb = sv * bv;
        a0 += rv * rv;
        a1 += sr * sr;
        a2 += sb * sb;
    }
    double a[3] = {a0, a1, a2};
    block_reduce<3, RED_SUM, RED_SUM, RED_SUM>(a, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = a[0];
        partials[gridDim.x + blockIdx.x] = a[1];
        partials[2 * gridDim.x + blockIdx.x] = a[2];
    }
}
__global__ void k_wscale_inplace(double* __restrict__ a, double* __restrict__ b, double s,
                                 uint32_t N) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        a[i] *= s;
        b[i] *= s;
    }
}

static int warm_alloc(ls_ctx* c) {
    if (c->wsum) return 0;
    if (hipMalloc((void**)&c->wsum, 64 * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_wsum, 64 * sizeof(double)) != hipSuccess) {
        c->err = "warm start: allocation failed";
        return -12;
    }
    c->allocations.push_back(c->wsum);
    c->other_bytes += 64 * sizeof(double);
    return 0;
}

// Dots x_a . y for a < n (n <= LS_WARM_MAX), returned on the host.
static int warm_dots(ls_ctx* c, const double* const* x, int n, const double* y, double* out) {
    const uint32_t N = (uint32_t)c->N;
    const int nb = std::min(ew_blocks(N), MAX_PARTIALS / LS_WARM_MAX);
    for (int a = 0; a < n; ++a)
        hipLaunchKernelGGL(k_wdot, dim3(nb), dim3(EW_BLOCK), 0, c->stream, x[a], y, N,
                           c->partials + (size_t)a * nb);
    hipLaunchKernelGGL(k_wsum, dim3(n), dim3(1024), 0, c->stream, c->partials, nb, c->wsum);
    LS_LAUNCH_CHECK();
    LS_HIP(hipMemcpyAsync(c->h_wsum, c->wsum, n * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < n; ++a) out[a] = c->h_wsum[a];
    return 0;
}

struct WarmSet {
    WBasis q{}, aq{};   // A^-orthonormal basis and A^ q
    int n = 0;
};

// Orthonormalise the raw vector in w (A^ w goes to aw) against the set and
// add it when it keeps a share of its energy above rounding.  Returns 0 and
// sets *added.
static int warm_add(ls_ctx* c, WarmSet& S, double* w, double* aw, bool* added) {
    const uint32_t N = (uint32_t)c->N;
    const int nbe = ew_blocks(N);
    *added = false;
    double e_raw = 0.0, h[LS_WARM_MAX];
    int st;
    launch_stencil<double>(c, w, aw, c->c64, nullptr);
    if ((st = warm_dots(c, (const double* const*)&w, 1, aw, &e_raw))) return st;
    if (!(e_raw > 0.0) || !std::isfinite(e_raw)) return 0;
    for (int pass = 0; pass < 2 && S.n > 0; ++pass) {
        // h_a = w . A^ q_a ; w -= sum h_a q_a
        if ((st = warm_dots(c, S.aq.v, S.n, w, h))) return st;
        WCoef C{};
        for (int a = 0; a < S.n; ++a) C.c[a] = h[a];
        hipLaunchKernelGGL(k_wsub, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, S.q, C, S.n, w, N);
        LS_LAUNCH_CHECK();
    }
    double e = e_raw;
    if (S.n > 0) {
        launch_stencil<double>(c, w, aw, c->c64, nullptr);
        if ((st = warm_dots(c, (const double* const*)&w, 1, aw, &e))) return st;
    }
    // keep directions whose energy is far above the rounding of the
    // orthogonalisation (relative energy ~ eps^2 times the stencil norm)
    if (!(e > 1e-20 * e_raw) || !std::isfinite(e)) return 0;
    hipLaunchKernelGGL(k_wscale_inplace, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, w, aw,
                       1.0 / std::sqrt(e), N);
    LS_LAUNCH_CHECK();
    S.q.v[S.n] = w;
    S.aq.v[S.n] = aw;
    ++S.n;
    *added = true;
    return 0;
}

// x64 += sum over q[from..n) of (q . b^) q   (from == 0: x64 = ...)
static int warm_accumulate(ls_ctx* c, const WarmSet& S, int from) {
    const uint32_t N = (uint32_t)c->N;
    double g[LS_WARM_MAX];
    int st;
    if ((st = warm_dots(c, S.q.v + from, S.n - from, c->bhat, g))) return st;
    WBasis B{};
    WCoef C{};
    for (int a = from; a < S.n; ++a) {
        B.v[a - from] = S.q.v[a];
        C.c[a - from] = g[a - from];
    }
    hipLaunchKernelGGL(k_wcombine, dim3(ew_blocks(N)), dim3(EW_BLOCK), 0, c->stream, B, C,
                       S.n - from, c->x64, N, from > 0 ? 1 : 0);
    LS_LAUNCH_CHECK();
    return 0;
}

// A solution of the history ring; a one-slot ring keeps none of its own:
// its solution is the pressure field, which only the pressure phase writes
// (an upload, a rollback or a new initial state empties the history).
static const double* hist_vec(const ls_ctx* c, int slot) {
    return c->hist_ring == 1 ? c->p : c->hist[slot];
}

// x64 <- projected warm start (the assembled system is in place and x64
// holds the previous-pressure start).  Work vectors: the pressure scratch
// (basis vector a in pw[2a], its image A^ w in pw[2a + 1]).
static int warm_start(ls_ctx* c, double ptop, ls_step_report* rep) {
    rep->warm_basis = 0;
    rep->warm_net = 0;
    c->resid_ready = 0;
    if (c->P.warm_start == LS_WARM_PREVIOUS || c->n_hist == 0) return 0;
    int st = warm_alloc(c);
    if (st) return st;
    const uint32_t N = (uint32_t)c->N;
    const int nbe = ew_blocks(N);
    // Only solutions driven by a comparable top pressure (within 100x) are
    // used: one solved at a near-zero drive (the ramp starts at 1e-6) is
    // dominated by terms that do not scale with it (the floored rows' lagged
    // values from the initial state, evaporation), and scaling it up puts
    // the start's error into the slowest modes of the operator: fewer
    // residual, many more CG iterations.
    int m = 0;
    while (m < std::min(c->n_hist, c->P.warm_history)) {
        const int slot = (c->hist_head + c->hist_ring - 1 - m) % c->hist_ring;
        const double ratio = c->hist_ptop[slot] / ptop;
        if (!(ratio >= 0.01 && ratio <= 100.0)) break;
        ++m;
    }
    // (with no usable solution the previous-pressure start stays: a network
    // correction of it, measured at 512^3, left an error that took mixed PCG
    // past 10000 iterations against 6007 without it)
    if (m == 0) return 0;
    // one vector and no network: x0 = (w.b^ / w.A^w) w with w = sqrt(d) p_newest,
    // its energy and projection from one stencil pass with both dots
    if (m == 1 && c->P.warm_start == LS_WARM_PROJECTED && use_vec<double>(c)) {
        const int slot = (c->hist_head + c->hist_ring - 1) % c->hist_ring;
        // with a one-solution history the solution is p itself, and w = sd p
        // is the previous-pressure start k_press_rhs_scale left in x64
        const bool wp = c->hist_ring == 1;
        double* w = wp ? c->x64 : c->pw[0];
        double* aw = c->pw[1];
        if (!wp)
            hipLaunchKernelGGL(k_wscale, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->hist[slot],
                               nullptr, c->sd, w, N);
        const dim3 grd = vec_grid<double>(c);
        const int nbv = (int)(grd.x * grd.y * grd.z);
        hipLaunchKernelGGL((k_stencil_vec<double, 2>), grd, dim3(WAVE, StencilShape<double>::TJV),
                           0, c->stream, w, aw, c->c64[0], c->c64[1], c->c64[2], sgeo(c),
                           c->partials, nullptr, c->bhat, nullptr, nullptr);
        hipLaunchKernelGGL(k_wsum, dim3(2), dim3(1024), 0, c->stream, c->partials, nbv, c->wsum);
        LS_LAUNCH_CHECK();
        LS_HIP(hipMemcpyAsync(c->h_wsum, c->wsum, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        const double e = c->h_wsum[0], gb = c->h_wsum[1];
        if (!(e > 0.0) || !std::isfinite(e) || !std::isfinite(gb)) return 0;
        if (c->P.solver != LS_SOLVER_LEARNED) {
            // x = (w.b^ / w.A^w) w and its residual b^ - x.A^w in one pass;
            // the solver's first residual call takes them
            hipLaunchKernelGGL(k_wfinish, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, w, aw,
                               c->bhat, c->sd, gb / e, c->x64, c->r64,
                               c->P.solver == LS_SOLVER_PLAIN64 ? c->p64 : nullptr, N,
                               c->partials);
            LS_LAUNCH_CHECK();
            c->resid_ready = nbe;
        } else {
            WBasis B{};
            WCoef C{};
            B.v[0] = w;
            C.c[0] = gb / e;
            hipLaunchKernelGGL(k_wcombine, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, B, C, 1,
                               c->x64, N, 0);
            LS_LAUNCH_CHECK();
        }
        rep->warm_basis = 1;
        return 0;
    }
    WarmSet S;
    int slot_used = 0;
    for (int a = 0; a < m; ++a) {   // newest first, then successive differences
        const int slot = (c->hist_head + c->hist_ring - 1 - a) % c->hist_ring;
        const int newer = (slot + 1) % c->hist_ring;
        double* w = c->pw[2 * slot_used];
        double* aw = c->pw[2 * slot_used + 1];
        hipLaunchKernelGGL(k_wscale, dim3(nbe), dim3(EW_BLOCK), 0, c->stream,
                           a == 0 ? hist_vec(c, slot) : hist_vec(c, newer),
                           a == 0 ? nullptr : hist_vec(c, slot), c->sd, w, N);
        LS_LAUNCH_CHECK();
        bool added = false;
        if ((st = warm_add(c, S, w, aw, &added))) return st;
        if (added) ++slot_used;
    }
    if (S.n == 0) return 0;
    if ((st = warm_accumulate(c, S, 0))) return st;
    rep->warm_basis = S.n;
    if (c->P.warm_start != LS_WARM_LEARNED) return 0;
    // network correction of the projected residual, in FP32 like the
    // preconditioner: r = b^ - A^ x0, e = P U(R r), then e joins the basis
    if ((st = learned_setup(c))) return st;
    if ((st = cg64_residual(c, c->P.press_rtol, 1, false))) return st;
    if ((st = poll_cg(c))) return st;
    if (c->h_cg->converged || !(c->h_cg->rr > 0.0)) return 0;
    const double scale = std::ldexp(1.0, std::ilogb(std::sqrt(c->h_cg->rr)));
    hipLaunchKernelGGL(k_inner_start, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->r64, c->r32,
                       c->p32, c->x32, N, 1.0 / scale, 0, c->partials);
    hipLaunchKernelGGL(k_lk_reset, dim3(1), dim3(1), 0, c->stream, c->fcg);
    CGeo g = cgeo(c);
    const int nbc = ew_blocks(g.cn);
    hipLaunchKernelGGL(k_lpack, dim3(nbc), dim3(EW_BLOCK), 0, c->stream, c->r32, c->cls,
                       c->c32[0], c->c32[1], c->c32[2], c->rc, c->net->x, g, c->partials);
    hipLaunchKernelGGL(k_lpack_scale, dim3(1), dim3(1024), 0, c->stream, c->partials, nbc, g.cn,
                       c->fcg);
    hipLaunchKernelGGL(k_lpack_ch0, dim3(nbc), dim3(EW_BLOCK), 0, c->stream, c->rc, c->net->x,
                       g.cn, c->fcg);
    LS_LAUNCH_CHECK();
    if ((st = unet_forward(c->net, 1))) {
        c->err = c->net->err;
        return st;
    }
    hipLaunchKernelGGL(k_lprolong, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->net->out, c->cls,
                       c->x32, g);
    double* w = c->pw[2 * slot_used];
    double* aw = c->pw[2 * slot_used + 1];
    hipLaunchKernelGGL(k_wfrom32, dim3(nbe), dim3(EW_BLOCK), 0, c->stream, c->x32, w, N);
    LS_LAUNCH_CHECK();
    bool added = false;
    if ((st = warm_add(c, S, w, aw, &added))) return st;
    if (!added) return 0;
    if ((st = warm_accumulate(c, S, S.n - 1))) return st;
    rep->warm_basis = S.n;
    rep->warm_net = 1;
    return 0;
}

// After a solve: the new pressure joins the history (replace: it overwrites
// the newest entry).
static int warm_push(ls_ctx* c, double ptop, bool replace) {
    if (c->P.warm_start == LS_WARM_PREVIOUS) {
        c->n_hist = 0;
        return 0;
    }
    // the ring holds warm_history solutions: a new length starts it afresh
    // and frees the slots beyond it
    const int ring = std::max(1, std::min(c->P.warm_history, LS_WARM_MAX));
    if (ring != c->hist_ring) {
        for (int q = ring == 1 ? 0 : ring; q < LS_WARM_MAX; ++q)
            if (c->hist[q]) {
                LS_HIP(hipStreamSynchronize(c->stream));
                for (size_t a = 0; a < c->allocations.size(); ++a)
                    if (c->allocations[a] == c->hist[q]) {
                        c->allocations.erase(c->allocations.begin() + a);
                        break;
                    }
                (void)hipFree(c->hist[q]);
                c->hist[q] = nullptr;
                c->grid_bytes -= c->N * sizeof(double);
            }
        c->hist_ring = ring;
        c->n_hist = 0;
        c->hist_head = 0;
        replace = false;
    }
    if (ring == 1) {            // the solution is p itself: nothing to copy
        c->hist_ptop[0] = ptop;
        c->hist_head = 0;
        c->n_hist = 1;
        return 0;
    }
    if (replace && c->n_hist > 0) {
        const int newest = (c->hist_head + c->hist_ring - 1) % c->hist_ring;
        LS_HIP(hipMemcpyAsync(c->hist[newest], c->p, c->N * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
        c->hist_ptop[newest] = ptop;
        return 0;
    }
    const int slot = c->hist_head;
    if (!c->hist[slot]) {
        if (hipMalloc((void**)&c->hist[slot], c->N * sizeof(double)) != hipSuccess) {
            c->err = "out of device memory (pressure history)";
            return -12;
        }
        c->allocations.push_back(c->hist[slot]);
        c->grid_bytes += c->N * sizeof(double);
    }
    LS_HIP(hipMemcpyAsync(c->hist[slot], c->p, c->N * sizeof(double), hipMemcpyDeviceToDevice,
                          c->stream));
    c->hist_ptop[slot] = ptop;
    c->hist_head = (slot + 1) % c->hist_ring;
    c->n_hist = std::min(c->n_hist + 1, c->hist_ring);
    return 0;
}

}  // namespace ls
