// ls_common.h — shared device helpers for the gfx950 hot path kernels.
// Wave64 reductions, label constants, launch geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lsgpu.h"

namespace ls {

constexpr uint8_t AIR = LS_AIR, PORE = LS_PORE, ROCK = LS_ROCK, SUL = LS_SULPHIDE,
                  PREC = LS_PRECIPITATE;

constexpr int WAVE = 64;          // CDNA wavefront width
constexpr int EW_BLOCK = 256;     // elementwise / reduction block size
// block partials of one reduction: the vectorised kernels of a
// 1024 x 1024 x 768 grid launch up to 786432 blocks
constexpr int MAX_PARTIALS = 1 << 20;

// Tile of the i-marching stencil kernels: 64 threads along k (one wave, one
// coalesced 512 B / 256 B row), TJ rows along j, each thread marches ICHUNK
// planes along i keeping x[i-1], x[i], x[i+1] and the -i face in registers.
constexpr int TK = 64;
constexpr int TJ = 4;
constexpr int ICHUNK = 32;

__device__ __forceinline__ int flat_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ int flat_bid() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, WAVE));
    return v;
}

enum RedOp { RED_SUM = 0, RED_MAX = 1, RED_MIN = 2 };

template <int OP>
__device__ __forceinline__ double wave_red(double v) {
    if (OP == RED_SUM) return wave_sum(v);
    if (OP == RED_MAX) return wave_max(v);
    return wave_min(v);
}
template <int OP>
__device__ __forceinline__ double red_identity() {
    if (OP == RED_SUM) return 0.0;
    if (OP == RED_MAX) return -1.0 / 0.0;
    return 1.0 / 0.0;
}
template <int OP>
__device__ __forceinline__ double red_combine(double a, double b) {
    if (OP == RED_SUM) return a + b;
    if (OP == RED_MAX) return fmax(a, b);
    return fmin(a, b);
}

// Deterministic block reduction of NV values; result valid in thread 0.
// lds must hold NV * (blockDim / 64) doubles.
template <int NV, int OP0, int OP1 = RED_SUM, int OP2 = RED_SUM, int OP3 = RED_SUM>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* lds) {
    const int tid = flat_tid();
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int nw = (nthreads + WAVE - 1) / WAVE;
    const int lane = tid % WAVE, w = tid / WAVE;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const int op = q == 0 ? OP0 : q == 1 ? OP1 : q == 2 ? OP2 : OP3;
        if (op == RED_SUM) v[q] = wave_sum(v[q]);
        else if (op == RED_MAX) v[q] = wave_max(v[q]);
        else v[q] = wave_min(v[q]);
        if (lane == 0) lds[q * nw + w] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int op = q == 0 ? OP0 : q == 1 ? OP1 : q == 2 ? OP2 : OP3;
            double a = lds[q * nw];
            for (int i = 1; i < nw; ++i) {
                double b = lds[q * nw + i];
                a = op == RED_SUM ? a + b : op == RED_MAX ? fmax(a, b) : fmin(a, b);
            }
            v[q] = a;
        }
    }
}

// Row-block kernels (a block walks rows (i, j) of the grid) fold their row
// blocks over grid y and z, so a grid of more than 65535 row blocks (1024 x
// 1024 rows and more) still launches: the block's row block, its linear
// index and the block count.
__device__ __forceinline__ uint32_t row_block() { return blockIdx.y + gridDim.y * blockIdx.z; }
__device__ __forceinline__ int block_linear() {
    return (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}
__device__ __forceinline__ int block_count() { return (int)(gridDim.x * gridDim.y * gridDim.z); }
// The grid of gx x rowblocks blocks, the row blocks folded over y and z.
inline dim3 row_grid(uint32_t gx, uint32_t rowblocks) {
    uint32_t gz = 1;
    while ((rowblocks + gz - 1) / gz > 65535) gz *= 2;
    return dim3(gx, (rowblocks + gz - 1) / gz, gz);
}

// Number of blocks for grid-stride elementwise/reduction kernels: enough to
// fill 256 CUs several times over, bounded so partial arrays stay small.
inline int ew_blocks(size_t n) {
    size_t b = (n + EW_BLOCK - 1) / EW_BLOCK;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

// Remap a linear block id so that consecutive tiles land on the same XCD.
// The dispatcher hands block b to XCD (b % 8); tile t = remap(b) makes each
// XCD own a contiguous range of tiles, so j-neighbour tiles share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    const int nx = 8;
    int per = nblocks / nx;
    if (per == 0 || b >= per * nx) return b;
    return (b % nx) * per + b / nx;
}

}  // namespace ls
