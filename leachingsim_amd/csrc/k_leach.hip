// k_leach.hip — reactive-fraction update (replaces updateFrac,
// src/Simulation.cpp:315-617).
//
// The reference sweeps in place in loop order k -> j -> i (i innermost):
// a voxel's update reads the labels and acid of its six neighbours and writes
// its own fraction/label and the acid of its six neighbours, so two voxels
// interact only when their Manhattan distance is <= 2.  Every such pair has a
// strictly increasing value of  w = i + 2 j + 3 k  from the earlier voxel (in
// loop order) to the later one, and no two voxels within distance 2 share a
// value of w.  Processing the hyperplanes w = 0, 1, 2, ... in order, each in
// parallel, therefore reproduces the sequential sweep operation for
// operation: same operands, same order of the acid decrements at every voxel.
//
// Fast path: if no reactive voxel (label >= Sulphide, frac > 0) has a
// non-zero surface flux and none is at or below the leaching threshold, the
// sweep cannot change anything; only the sub-step count is reported.
#include <cmath>

#include "ls_context.h"

namespace ls {

struct LGeo {
    int ni, nj, nk;
    uint32_t si, sj;
};

__device__ __forceinline__ bool lnbr(const LGeo& g, int i, int j, int k, uint32_t v, int f,
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

// calc_surface_flux (src/Simulation.cpp:383-405)
__device__ __forceinline__ double surface_flux(const uint8_t* img, const double* acid,
                                               const LGeo& g, int i, int j, int k, uint32_t v,
                                               int f, double k_eff) {
    uint32_t n;
    if (lnbr(g, i, j, k, v, f, n)) {
        uint8_t n_type = img[n];
        if (n_type == PORE || n_type == ROCK) return k_eff * acid[n];
    }
    return 0.0;
}

__global__ void k_leach_scan(const uint8_t* __restrict__ img, const float* __restrict__ frac,
                             const double* __restrict__ acid, LGeo g, uint32_t N, double k_eff,
                             double dx, double* __restrict__ partials) {
    __shared__ double lds[2 * EW_BLOCK / WAVE];
    double max_rate = 0.0, needed = 0.0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const uint8_t t = img[v];
        if (t < SUL) continue;
        const uint32_t ij = v / (uint32_t)g.nk;
        const int k = (int)(v - ij * (uint32_t)g.nk);
        const int i = (int)(ij / (uint32_t)g.nj);
        const int j = (int)(ij - (uint32_t)i * (uint32_t)g.nj);
        double fs[6];
        for (int f = 0; f < 6; ++f) fs[f] = surface_flux(img, acid, g, i, j, k, v, f, k_eff);
        if (t == SUL) {
            double flux_sum = 0.0;
            for (int f = 0; f < 6; ++f) flux_sum += fs[f];
            double area_per_face = dx * dx;
            double voxel_volume = dx * dx * dx;
            double vol_rate = flux_sum * area_per_face;
            double fractional_rate = vol_rate / voxel_volume;
            if (fractional_rate > max_rate) max_rate = fractional_rate;
        }
        const float fr = frac[v];
        if (fr > 0.0) {
            double flux_sum = fs[0] + fs[1] + fs[2] + fs[3] + fs[4] + fs[5];
            if (flux_sum != 0.0 || !(flux_sum == flux_sum) || fr <= 1e-4) needed = 1.0;
        }
    }
    double a[2] = {max_rate, needed};
    block_reduce<2, RED_MAX, RED_MAX>(a, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = a[0];
        partials[gridDim.x + blockIdx.x] = a[1];
    }
}

__global__ void k_leach_plan(const double* __restrict__ partials, int n, double dt,
                             double time_scalar, StepScalars* ss) {
    __shared__ double lds[2 * 1024 / WAVE];
    double m = 0.0, need = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        m = fmax(m, partials[i]);
        need = fmax(need, partials[n + i]);
    }
    double a[2] = {m, need};
    block_reduce<2, RED_MAX, RED_MAX>(a, lds);
    if (threadIdx.x == 0) {
        double global_max_rate = a[0];
        double effective_fractional_rate = global_max_rate * time_scalar;
        double dt_safe = dt;
        if (effective_fractional_rate > 1e-20) dt_safe = 0.1 / effective_fractional_rate;
        double calculated_steps = ceil(dt / dt_safe);
        int num_steps = 1;
        if (calculated_steps > 1000) num_steps = 1000;
        else if (calculated_steps > 1.0) num_steps = (int)calculated_steps;
        ss->leach_max_rate = global_max_rate;
        ss->leach_nsub = num_steps;
        ss->leach_dt_sub = dt / num_steps;
        ss->leach_needed = a[1] > 0.0 ? 1 : 0;
        ss->leached = 0;
    }
}

// One hyperplane w = i + 2j + 3k of the in-place sweep (src/Simulation.cpp:
// 469-607).
__global__ void k_leach_wave(uint8_t* __restrict__ img, float* __restrict__ frac,
                             double* __restrict__ acid, double* __restrict__ inv,
                             double* __restrict__ conc, LGeo g, int w, double k_eff, double dx,
                             double time_scalar, double c_sat, StepScalars* __restrict__ ss) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= g.nk || j >= g.nj) return;
    const int i = w - 2 * j - 3 * k;
    if (i < 0 || i >= g.ni) return;
    const uint32_t v = (uint32_t)i * g.si + (uint32_t)j * g.sj + (uint32_t)k;
    if (!(img[v] >= SUL && frac[v] > 0.0)) return;
    const double dt_sub = ss->leach_dt_sub;
    double f_xp = surface_flux(img, acid, g, i, j, k, v, 0, k_eff);
    double f_xm = surface_flux(img, acid, g, i, j, k, v, 1, k_eff);
    double f_yp = surface_flux(img, acid, g, i, j, k, v, 2, k_eff);
    double f_ym = surface_flux(img, acid, g, i, j, k, v, 3, k_eff);
    double f_zp = surface_flux(img, acid, g, i, j, k, v, 4, k_eff);
    double f_zm = surface_flux(img, acid, g, i, j, k, v, 5, k_eff);
    double flux_sum = f_xp + f_xm + f_yp + f_ym + f_zp + f_zm;
    double area_per_face = dx * dx;
    double voxel_volume = dx * dx * dx;
    double mass_loss = flux_sum * area_per_face * dt_sub;
    double fractional_loss = mass_loss / voxel_volume;
    float fr = (float)((double)frac[v] - (fractional_loss * time_scalar));
    frac[v] = fr;
    if (flux_sum > 0.0) {
        const double fs[6] = {f_xp, f_xm, f_yp, f_ym, f_zp, f_zm};
        for (int f = 0; f < 6; ++f) {
            uint32_t n;
            if (lnbr(g, i, j, k, v, f, n)) acid[n] -= fractional_loss * (fs[f] / flux_sum);
        }
    }
    if (fr <= 1e-4) {
        frac[v] = 0.0f;
        if (img[v] == SUL) {
            img[v] = PORE;
            inv[v] = 0.0;
            conc[v] = c_sat;
            atomicAdd(&ss->leached, 1);
        }
    }
}

int phase_leaching(ls_ctx* c, double dt, ls_step_report* rep) {
    if (c->acid_zero && c->frac_clean) {
        // every surface flux is k_eff * 0 and no fraction sits at or below
        // the threshold: the scan would find max rate 0 (one sub-step) and
        // nothing to do, and the sweep would change nothing
        if (rep) {
            rep->leach_substeps = 1;
            rep->leach_max_rate = 0.0;
            rep->leached = 0;
            rep->leach_general_path = 0;
        }
        return 0;
    }
    // the sweep writes frac: the step's snapshot of it first
    if (c->bfrac_stale) {
        if (int st = copy_d2d(c, c->b_frac, c->frac, c->N * 4)) return st;
        c->bfrac_stale = false;
    }
    LGeo g;
    g.ni = c->ni; g.nj = c->nj; g.nk = c->nk;
    g.sj = (uint32_t)c->nk;
    g.si = (uint32_t)c->nk * (uint32_t)c->nj;
    const double dx = c->P.dx;
    const double k_trans = c->P.D / dx;
    const double k_eff = 1.0 / ((1.0 / k_trans) + (1.0 / c->P.kreac));
    const uint32_t N = (uint32_t)c->N;
    const int nb = ew_blocks(N);
    hipLaunchKernelGGL(k_leach_scan, dim3(nb), dim3(EW_BLOCK), 0, c->stream, c->img, c->frac,
                       c->acid, g, N, k_eff, dx, c->partials);
    hipLaunchKernelGGL(k_leach_plan, dim3(1), dim3(1024), 0, c->stream, c->partials, nb, dt,
                       c->P.time_scalar, c->ss);
    LS_LAUNCH_CHECK();
    LS_HIP(hipMemcpyAsync(c->h_ss, c->ss, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    const int nsub = c->h_ss->leach_nsub;
    int general = 0;
    if (c->h_ss->leach_needed) {
        general = 1;
        c->frac_clean = false;    // the sweep may leave fractions anywhere in (0, 1]
        const int W = (c->ni - 1) + 2 * (c->nj - 1) + 3 * (c->nk - 1) + 1;
        dim3 blk(64, 4), grd((c->nk + 63) / 64, (c->nj + 3) / 4);
        for (int step = 0; step < nsub; ++step) {
            for (int w = 0; w < W; ++w)
                hipLaunchKernelGGL(k_leach_wave, grd, blk, 0, c->stream, c->img, c->frac, c->acid,
                                   c->inv, c->conc, g, w, k_eff, dx, c->P.time_scalar,
                                   c->P.c_sat, c->ss);
            LS_LAUNCH_CHECK();
        }
        LS_HIP(hipMemcpyAsync(c->h_ss, c->ss, sizeof(StepScalars), hipMemcpyDeviceToHost,
                              c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
    }
    if (rep) {
        rep->leach_substeps = nsub;
        rep->leach_max_rate = c->h_ss->leach_max_rate;
        rep->leached = general ? c->h_ss->leached : 0;
        rep->leach_general_path = general;
    }
    return 0;
}

}  // namespace ls
