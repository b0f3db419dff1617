// k_unet.hip — the 3-level 3D U-Net of the learned pressure preconditioner:
// BF16 activations (channels-last), BF16 weights, FP32 accumulation on MFMA
// (gfx950's 32-deep v_mfma_f32_16x16x32_bf16), hand-written forward,
// backward and Adam.
//
// Convolution as implicit GEMM, D = W . X^T per 16 output channels x 16
// voxels, 32 deep: the A operand is a 16 (cout) x 32 weight tile (8 channels
// of one tap per lane, one 16-byte load), the B operand 32 x 16 voxels (the
// same 8 channels of one voxel per lane), and each lane ends with 4
// consecutive output channels of one voxel (one 8-byte store).  A wave
// computes 64 voxels x all output channels; pairs of (tap, 16-channel chunk)
// units form the K loop.  Concatenated inputs ([2x upsampled deep, skip]) are
// read in place from the two tensors.
//
// Backward-data is the same convolution on the ReLU-masked output gradient
// with the mirrored, transposed weights W'[26 - t][ci][co] = W[t][co][ci].
// Backward-weight is a GEMM over voxels per tap: per (chunk, wave) FP32
// partials reduced in a fixed order (deterministic).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ls_unet.h"

namespace ls {

typedef __bf16 bf16;
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf4 load4(const bf16* p) { return *(const bf4*)p; }
__device__ __forceinline__ bf4 load4(const float* p) {
    const f4 v = *(const f4*)p;
    bf4 r;
    r[0] = (bf16)v[0]; r[1] = (bf16)v[1]; r[2] = (bf16)v[2]; r[3] = (bf16)v[3];
    return r;
}
__device__ __forceinline__ void store4(bf16* p, f4 v) {
    bf4 r;
    r[0] = (bf16)v[0]; r[1] = (bf16)v[1]; r[2] = (bf16)v[2]; r[3] = (bf16)v[3];
    *(bf4*)p = r;
}
__device__ __forceinline__ void store4(float* p, f4 v) { *(f4*)p = v; }

__device__ __forceinline__ void vdecode(uint32_t v, const NetLvl& L, int& i, int& j, int& k) {
    const uint32_t ij = v / (uint32_t)L.nk;
    k = (int)(v - ij * (uint32_t)L.nk);
    i = (int)(ij / (uint32_t)L.nj);
    j = (int)(ij - (uint32_t)i * (uint32_t)L.nj);
}

// ---------------------------------------------------------------------------
// Convolution forward (also backward-data with mirrored weights)
//   Y[b][v][co] = act(bias[co] + sum_t sum_ci W[t][co][ci] X[b][v + off(t)][ci])
//   X = [UPA ? up2(A) : A  (CA channels), B (CB channels)]
//
// The reduction runs over units u = t NCK + ck of 16 input channels (tap t,
// channel chunk ck); unit pairs (2p, 2p + 1) feed one 32-deep MFMA
// (v_mfma_f32_16x16x32_bf16): lane group h = q >> 1 of the A and B operands
// takes unit 2p + h, its 8 channels 16 ck + 8 (q & 1) .. + 7 (an odd unit
// count pads the last pair with zeros).  Every kernel below issues the pairs
// in this order per output tile, so they agree bit for bit.
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf8 load8(const bf16* p) { return *(const bf8*)p; }
__device__ __forceinline__ bf8 load8(const float* p) {
    const f4 a = *(const f4*)p, b = *(const f4*)(p + 4);
    bf8 r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        r[e] = (bf16)a[e];
        r[e + 4] = (bf16)b[e];
    }
    return r;
}

__device__ __forceinline__ bf8 zero8() {
    bf8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (bf16)0.0f;
    return r;
}

__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int CA, int CB, int COUT, bool UPA, bool RELU, typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void k_conv(const TIN* __restrict__ A, const TIN* __restrict__ B,
                                              const bf16* __restrict__ W,
                                              const float* __restrict__ bias,
                                              TOUT* __restrict__ Y, NetLvl L, NetLvl LA) {
    constexpr int CIN = CA + CB, NCK = CIN / 16, U = 27 * NCK, NP = (U + 1) / 2;
    constexpr int MT = COUT / 16;
    constexpr int NT = 4;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int b = blockIdx.y;
    const uint32_t v0 = (blockIdx.x * 4u + wave) * 64u;
    if (v0 >= L.n) return;  // uniform per wave
    const int col = lane & 15;
    const int q = lane >> 4, h = q >> 1, cq = 8 * (q & 1);
    const int kq = 4 * q;
    int vi[NT], vj[NT], vk[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) vdecode(v0 + n * 16 + col, L, vi[n], vj[n], vk[n]);
    f4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    const size_t bA = (size_t)b * (UPA ? LA.n : L.n);
    const size_t bB = (size_t)b * L.n;
    for (int pr = 0; pr < NP; ++pr) {
        const int u = 2 * pr + h;
        const bool uv = u < U;
        const int t = uv ? u / NCK : 0;
        const int c0 = 16 * (uv ? u - t * NCK : 0) + cq;     // this lane's first channel
        const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
        bf8 wa[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
            wa[m] = uv ? load8(W + ((size_t)t * COUT + m * 16 + col) * CIN + c0) : zero8();
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int ii = vi[n] + di, jj = vj[n] + dj, kk = vk[n] + dk;
            const bool inb = uv && ii >= 0 && ii < L.ni && jj >= 0 && jj < L.nj && kk >= 0 &&
                             kk < L.nk;
            bf8 xb = zero8();
            if (inb) {
                if (c0 < CA) {
                    const uint32_t na = UPA ? ((uint32_t)(ii >> 1) * LA.nj + (uint32_t)(jj >> 1)) *
                                                      LA.nk + (uint32_t)(kk >> 1)
                                            : ((uint32_t)ii * L.nj + (uint32_t)jj) * L.nk +
                                                  (uint32_t)kk;
                    xb = load8(A + (bA + na) * CA + c0);
                } else {
                    const uint32_t nb = ((uint32_t)ii * L.nj + (uint32_t)jj) * L.nk + (uint32_t)kk;
                    xb = load8(B + (bB + nb) * CB + (c0 - CA));
                }
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m][n] = mfma32(wa[m], xb, acc[m][n]);
        }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int co0 = m * 16 + kq;
        f4 bv = f4{0.f, 0.f, 0.f, 0.f};
        if (bias != nullptr) bv = f4{bias[co0], bias[co0 + 1], bias[co0 + 2], bias[co0 + 3]};
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            f4 o = acc[m][n] + bv;
            if (RELU) {
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = o[e] > 0.f ? o[e] : 0.f;
            }
            const uint32_t v = v0 + n * 16 + col;
            store4(Y + (bB + v) * COUT + co0, o);
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled convolution (forward, and backward-data with the mirrored weights):
// the same result as k_conv, bit for bit.  A block of TJ x MS waves owns TJ
// rows (j) x TK voxels (k) of an output plane and marches CT_ICH planes along
// i; the MS waves of a row split its output channels.  The input planes i - 1 .. i + 1 of the tile and its
// one-voxel halo sit in LDS as BF16 ([up2(A), B] channels for a decoder
// layer): four rotating slots, the plane after next fetched into registers
// while the current one is computed, one barrier per plane.  A slot is laid
// out [row][channel quad][voxel][4 channels], so the 16 lanes that read one
// quad of a B operand read 128 contiguous bytes (no bank conflict).  With
// one 16-channel output block per wave and at most 48 input channels the
// weight fragments (27 per 16 input channels) stay in registers, at most 256
// per lane (two waves per SIMD); otherwise each tap's fragments are loaded
// (L1/L2) one tap ahead.  The MFMA sequence of an output tile (taps, then
// 16-channel chunks, in order) is that of k_conv.  Tile shapes keep a
// block's LDS at 26-41 KB.  The grid of the first level has extents that are
// multiples of 16, so a level's extents are multiples of 4.
// ---------------------------------------------------------------------------
constexpr int CT_ICH = 16;
// Weight fragments of a wave's output block stay in registers up to this
// many 16-channel input chunks per tap (else they are loaded one MFMA pair
// ahead).
#ifndef LS_CONV_WREG_NCK
#define LS_CONV_WREG_NCK 3
#endif

// Tile of a layer: TJ rows x TK voxels per block, and MS waves per row that
// split the output channels (each wave computes COUT / MS of them from the
// shared LDS tile, so that its weight fragments fit in registers).
template <int CIN, int COUT> struct ConvTile;
template <> struct ConvTile<16, 32> { static constexpr int TJ = 2, TK = 64, MS = 2; };
template <> struct ConvTile<32, 32> { static constexpr int TJ = 2, TK = 32, MS = 2; };
template <> struct ConvTile<32, 64> { static constexpr int TJ = 1, TK = 32, MS = 4; };
#ifndef LS_CT48_TJ
#define LS_CT48_TJ 4
#endif
#ifndef LS_CT48_TK
#define LS_CT48_TK 16
#endif
template <> struct ConvTile<48, 16> { static constexpr int TJ = LS_CT48_TJ, TK = LS_CT48_TK, MS = 1; };

template <int CA, int CB, int COUT, bool UPA, bool RELU, typename TIN, typename TOUT>
__global__ __launch_bounds__((64 * ConvTile<CA + CB, COUT>::TJ * ConvTile<CA + CB, COUT>::MS), 2) void k_conv_tiled(
    const TIN* __restrict__ A, const TIN* __restrict__ B, const bf16* __restrict__ W,
    const float* __restrict__ bias, TOUT* __restrict__ Y, NetLvl L, NetLvl LA) {
    constexpr int CIN = CA + CB, NQ = CIN / 4, NCK = CIN / 16, MT = COUT / 16;
    constexpr int TJ = ConvTile<CIN, COUT>::TJ, TK = ConvTile<CIN, COUT>::TK, NT = TK / 16;
    constexpr int MS = ConvTile<CIN, COUT>::MS, MW = MT / MS;   // output blocks per wave
    constexpr int NTH = 64 * TJ * MS, VOX = TK + 2, G8 = CIN / 8;
    constexpr int PLANE = (TJ + 2) * NQ * VOX * 4;           // BF16 elements per slot
    constexpr int CHUNKS = (TJ + 2) * VOX * G8;              // 8-channel chunks per plane
    constexpr int PER = (CHUNKS + NTH - 1) / NTH;
    constexpr int U = 27 * NCK, NP = (U + 1) / 2;           // channel units, MFMA pairs
    constexpr bool WREG = MW == 1 && NCK <= LS_CONV_WREG_NCK;   // weights in registers
    __shared__ bf16 tile[4 * PLANE];
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) / MS, mg = (threadIdx.x >> 6) % MS;
    const int col = lane & 15, q = lane >> 4, h = q >> 1, cq = 8 * (q & 1);
    const int nkt = L.nk / TK + (L.nk % TK ? 1 : 0);
    const int k0 = (int)(blockIdx.x % (unsigned)nkt) * TK;
    const int j0 = (int)(blockIdx.x / (unsigned)nkt) * TJ;
    const int i0 = blockIdx.y * CT_ICH, i1 = min(i0 + CT_ICH, L.ni);
    const int b = blockIdx.z;
    const size_t offA = (size_t)b * (UPA ? LA.n : L.n), offB = (size_t)b * L.n;
    // the lane's unit of MFMA pair pr: 2 pr + h (tap, channel chunk)
    auto unit = [&](int pr, int& t, int& c0) {
        const int u = 2 * pr + h;
        t = u < U ? u / NCK : -1;
        c0 = u < U ? 16 * (u - (u / NCK) * NCK) + cq : 0;
    };
    bf8 wr[WREG ? NP : 1];
    if constexpr (WREG) {
#pragma unroll
        for (int pr = 0; pr < NP; ++pr) {
            int t, c0;
            unit(pr, t, c0);
            wr[pr] = t >= 0 ? load8(W + ((size_t)t * COUT + mg * 16 + col) * CIN + c0) : zero8();
        }
    }
    // chunk c of a plane: row c / (VOX G8), voxel (c / G8) % VOX, channels
    // 8 g .. 8 g + 7 with g = c % G8 (quads 2 g, 2 g + 1)
    bf8 pre[PER];
    auto fetch = [&](int ip) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int c = threadIdx.x + NTH * r;
            const int row = c / (VOX * G8), rem = c - row * (VOX * G8);
            const int vx = rem / G8, g = rem - vx * G8;
            const int jj = j0 - 1 + row, kk = k0 - 1 + vx;
            bf8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (bf16)0.0f;
            if (c < CHUNKS && ip >= 0 && ip < L.ni && jj >= 0 && jj < L.nj && kk >= 0 &&
                kk < L.nk) {
                if (8 * g < CA) {
                    const size_t n = UPA ? ((size_t)(ip >> 1) * LA.nj + (jj >> 1)) * LA.nk + (kk >> 1)
                                         : ((size_t)ip * L.nj + jj) * L.nk + kk;
                    v = load8(A + (offA + n) * CA + 8 * g);
                } else {
                    const size_t n = ((size_t)ip * L.nj + jj) * L.nk + kk;
                    v = load8(B + (offB + n) * CB + (8 * g - CA));
                }
            }
            pre[r] = v;
        }
    };
    auto publish = [&](int slot) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int c = threadIdx.x + NTH * r;
            if (c < CHUNKS) {
                const int row = c / (VOX * G8), rem = c - row * (VOX * G8);
                const int vx = rem / G8, g = rem - vx * G8;
                bf16* d = tile + slot * PLANE + ((row * NQ + 2 * g) * VOX + vx) * 4;
                *(bf4*)d = bf4{pre[r][0], pre[r][1], pre[r][2], pre[r][3]};
                *(bf4*)(d + 4 * VOX) = bf4{pre[r][4], pre[r][5], pre[r][6], pre[r][7]};
            }
        }
    };
    // this wave's output blocks are mg MW .. mg MW + MW - 1
    auto wload = [&](int pr, bf8 (&wa)[MW]) {
        int t, c0;
        unit(pr, t, c0);
#pragma unroll
        for (int m = 0; m < MW; ++m)
            wa[m] = t >= 0 ? load8(W + ((size_t)t * COUT + (mg * MW + m) * 16 + col) * CIN + c0)
                           : zero8();
    };
    fetch(i0 - 1);
    publish((i0 - 1) & 3);
    fetch(i0);
    publish(i0 & 3);
    fetch(i0 + 1);
    const int j = j0 + wave;
    for (int i = i0; i < i1; ++i) {
        // slot (i + 1) & 3 held plane i - 3, last read for output plane
        // i - 2, which every wave finished before the barrier of plane i - 1
        publish((i + 1) & 3);
        if (i + 1 < i1) fetch(i + 2);
        __syncthreads();
        f4 acc[MW][NT];
#pragma unroll
        for (int m = 0; m < MW; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = f4{0.f, 0.f, 0.f, 0.f};
        // one MFMA pair: every column, also those past nk in a partial k-tile
        // (their LDS voxels are zero and their results are not stored; a
        // branch here would serialise the LDS reads).  The lane reads the 8
        // channels of its unit at its tap: two quads, VOX x 4 elements apart.
        auto pair = [&](int pr, const bf8 (&wa)[MW]) {
            int t, c0;
            unit(pr, t, c0);
            const int tt = t >= 0 ? t : 13;            // a padded unit reads the centre tap
            const int di = tt / 9 - 1, dj = (tt / 3) % 3 - 1, dk = tt % 3 - 1;
            const bf16* base = tile + ((i + di) & 3) * PLANE +
                               (((wave + 1 + dj) * NQ + c0 / 4) * VOX + 1 + dk + col) * 4;
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const bf4 lo = *(const bf4*)(base + n * 64);
                const bf4 hi = *(const bf4*)(base + 4 * VOX + n * 64);
                bf8 xb = bf8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                if (t < 0) xb = zero8();
#pragma unroll
                for (int m = 0; m < MW; ++m) acc[m][n] = mfma32(wa[m], xb, acc[m][n]);
            }
        };
        if constexpr (WREG) {
#pragma unroll
            for (int pr = 0; pr < NP; ++pr) {
                bf8 w1[1] = {wr[pr]};
                pair(pr, w1);
            }
        } else {   // weights one pair ahead; the pair loop stays rolled
            bf8 wa[MW], wn[MW];
            wload(0, wa);
#pragma unroll 1
            for (int pr = 0; pr < NP; ++pr) {
                if (pr + 1 < NP) wload(pr + 1, wn);
                pair(pr, wa);
#pragma unroll
                for (int m = 0; m < MW; ++m) wa[m] = wn[m];
            }
        }
        if (j < L.nj) {
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                const int co0 = (mg * MW + m) * 16 + 4 * q;
                f4 bv = f4{0.f, 0.f, 0.f, 0.f};
                if (bias != nullptr) bv = f4{bias[co0], bias[co0 + 1], bias[co0 + 2], bias[co0 + 3]};
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    const int k = k0 + 16 * n + col;   // nk is a multiple of 4 only
                    if (k >= L.nk) continue;           // at the coarser levels
                    f4 o = acc[m][n] + bv;
                    if (RELU) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) o[e] = o[e] > 0.f ? o[e] : 0.f;
                    }
                    store4(Y + (offB + ((size_t)i * L.nj + j) * L.nk + k) * COUT + co0, o);
                }
            }
        }
    }
}

// The one-block case (16 -> 16 channels: e0a, e0b, d0b, out and their
// backward-data) as a kernel of its own: weights in registers, 4 rows x 64
// voxels, the taps unrolled; the general kernel's form of it is ~60 % slower.
constexpr int CT_TJ = 4, CT_TK = 64;
constexpr int CT_ROWS = CT_TJ + 2, CT_VOX = CT_TK + 2;
constexpr int CT_PLANE = CT_ROWS * CT_VOX * 16;            // BF16 elements per slot
constexpr int CT_CHUNKS = CT_ROWS * CT_VOX * 2;            // 8-channel chunks per plane
constexpr int CT_PER_THREAD = (CT_CHUNKS + 255) / 256;

template <bool RELU, typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void k_conv16_tiled(const TIN* __restrict__ X,
                                                      const bf16* __restrict__ W,
                                                      const float* __restrict__ bias,
                                                      TOUT* __restrict__ Y, NetLvl L) {
    __shared__ bf16 tile[4 * CT_PLANE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, q = lane >> 4, h = q >> 1, cq = 8 * (q & 1);
    const int nkt = L.nk / CT_TK + (L.nk % CT_TK ? 1 : 0);
    const int k0 = (int)(blockIdx.x % (unsigned)nkt) * CT_TK;
    const int j0 = (int)(blockIdx.x / (unsigned)nkt) * CT_TJ;
    const int i0 = blockIdx.y * CT_ICH, i1 = min(i0 + CT_ICH, L.ni);
    const size_t xoff = (size_t)blockIdx.z * L.n;
    // MFMA pair pr takes taps 2 pr (lane groups 0, 1) and 2 pr + 1 (2, 3);
    // tap 27 is the zero padding of the last pair
    bf8 wr[14];
#pragma unroll
    for (int pr = 0; pr < 14; ++pr) {
        const int t = 2 * pr + h;
        wr[pr] = t < 27 ? load8(W + ((size_t)t * 16 + col) * 16 + cq) : zero8();
    }
    // chunk c of a plane: row c / (2 CT_VOX), voxel (c / 2) % CT_VOX, channels
    // 8 h .. 8 h + 7 with h = c % 2, i.e. quads 2 h and 2 h + 1
    bf8 pre[CT_PER_THREAD];
    auto fetch = [&](int ip) {
#pragma unroll
        for (int r = 0; r < CT_PER_THREAD; ++r) {
            const int c = threadIdx.x + 256 * r;
            const int row = c / (2 * CT_VOX), rem = c - row * (2 * CT_VOX);
            const int jj = j0 - 1 + row, kk = k0 - 1 + (rem >> 1);
            bf8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (bf16)0.0f;
            if (c < CT_CHUNKS && ip >= 0 && ip < L.ni && jj >= 0 && jj < L.nj && kk >= 0 &&
                kk < L.nk)
                v = load8(X + (xoff + ((size_t)ip * L.nj + jj) * L.nk + kk) * 16 + (rem & 1) * 8);
            pre[r] = v;
        }
    };
    auto publish = [&](int slot) {
#pragma unroll
        for (int r = 0; r < CT_PER_THREAD; ++r) {
            const int c = threadIdx.x + 256 * r;
            if (c < CT_CHUNKS) {
                const int row = c / (2 * CT_VOX), rem = c - row * (2 * CT_VOX);
                bf16* d = tile + slot * CT_PLANE +
                          ((row * 4 + 2 * (rem & 1)) * CT_VOX + (rem >> 1)) * 4;
                *(bf4*)d = bf4{pre[r][0], pre[r][1], pre[r][2], pre[r][3]};
                *(bf4*)(d + 4 * CT_VOX) = bf4{pre[r][4], pre[r][5], pre[r][6], pre[r][7]};
            }
        }
    };
    fetch(i0 - 1);
    publish((i0 - 1) & 3);
    fetch(i0);
    publish(i0 & 3);
    fetch(i0 + 1);
    const int j = j0 + wave;
    f4 bv = f4{0.f, 0.f, 0.f, 0.f};
    if (bias != nullptr) bv = f4{bias[4 * q], bias[4 * q + 1], bias[4 * q + 2], bias[4 * q + 3]};
    for (int i = i0; i < i1; ++i) {
        // slot (i + 1) & 3 held plane i - 3, last read for output plane
        // i - 2, which every wave finished before the barrier of plane i - 1
        publish((i + 1) & 3);
        if (i + 1 < i1) fetch(i + 2);
        __syncthreads();
        f4 acc[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int pr = 0; pr < 14; ++pr) {
            const int t = 2 * pr + h < 27 ? 2 * pr + h : 13;   // the padding reads the centre
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            const bf16* base = tile + ((i + di) & 3) * CT_PLANE +
                               (((wave + 1 + dj) * 4 + 2 * (q & 1)) * CT_VOX + 1 + dk + col) * 4;
            // all four columns, also those past nk in a partial k-tile (their
            // LDS voxels are zero and their results are not stored): a branch
            // here would serialise the LDS reads
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const bf4 lo = *(const bf4*)(base + n * 64);
                const bf4 hi = *(const bf4*)(base + 4 * CT_VOX + n * 64);
                bf8 xb = bf8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                if (pr == 13 && h == 1) xb = zero8();
                acc[n] = mfma32(wr[pr], xb, acc[n]);
            }
        }
        if (j < L.nj) {
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const int k = k0 + 16 * n + col;
                if (k0 + 16 * n >= L.nk) continue;
                f4 o = acc[n] + bv;
                if (RELU) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) o[e] = o[e] > 0.f ? o[e] : 0.f;
                }
                store4(Y + (xoff + ((size_t)i * L.nj + j) * L.nk + k) * 16 + 4 * q, o);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: part[c][t][co][ci] = sum over the voxels of chunk c of
// g[v][co] X[v + off(t)][ci]; one wave per (chunk, tap), blocks of 32 voxels
// along K (v_mfma_f32_16x16x32_bf16: lane group q takes voxels 8 q .. 8 q + 7
// of the block).  X is the layer input ([up(A), B]), g the FP32 gradient of the
// pre-activation.
// ---------------------------------------------------------------------------
template <int CA, int CB, int COUT, bool UPA>
__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ A, const bf16* __restrict__ B,
                                               const float* __restrict__ G, float* __restrict__ part,
                                               NetLvl L, NetLvl LA, uint32_t chunk, int nchunk) {
    constexpr int CIN = CA + CB;
    constexpr int MT = COUT / 16, CT = CIN / 16;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    const int b = blockIdx.z;
    const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
    const uint32_t c = blockIdx.x * 4u + wave;     // chunk id within the sample
    const int row = lane & 15;
    const int kq = (lane >> 4) * 4;      // output rows
    const int kv = (lane >> 4) * 8;      // voxels of the block
    f4 acc[MT][CT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int q = 0; q < CT; ++q) acc[m][q] = f4{0.f, 0.f, 0.f, 0.f};
    const size_t bA = (size_t)b * (UPA ? LA.n : L.n);
    const size_t bB = (size_t)b * L.n;
    const uint32_t vbeg = c * chunk;
    const uint32_t vend = min(vbeg + chunk, L.n);
    if (c < (uint32_t)nchunk) {
        for (uint32_t vb = vbeg; vb < vend; vb += 32) {
            // this lane's 8 voxels along K
            bf8 ga[MT];
            bf8 xb[CT];
#pragma unroll
            for (int m = 0; m < MT; ++m) ga[m] = zero8();
#pragma unroll
            for (int q = 0; q < CT; ++q) xb[q] = zero8();
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t v = vb + kv + e;
                if (v >= vend) continue;
                int i, j, k;
                vdecode(v, L, i, j, k);
#pragma unroll
                for (int m = 0; m < MT; ++m) ga[m][e] = (bf16)G[(bB + v) * COUT + m * 16 + row];
                const int ii = i + di, jj = j + dj, kk = k + dk;
                if (ii < 0 || ii >= L.ni || jj < 0 || jj >= L.nj || kk < 0 || kk >= L.nk) continue;
                const uint32_t nb = ((uint32_t)ii * L.nj + (uint32_t)jj) * L.nk + (uint32_t)kk;
                const uint32_t na = UPA ? ((uint32_t)(ii >> 1) * LA.nj + (uint32_t)(jj >> 1)) * LA.nk +
                                              (uint32_t)(kk >> 1)
                                        : nb;
#pragma unroll
                for (int q = 0; q < CT; ++q) {
                    const int c0 = q * 16;
                    xb[q][e] = (c0 < CA) ? A[(bA + na) * CA + c0 + row]
                                         : B[(bB + nb) * CB + (c0 - CA) + row];
                }
            }
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int q = 0; q < CT; ++q)
                    acc[m][q] = mfma32(ga[m], xb[q], acc[m][q]);
        }
    }
    // lane holds D[co = m*16 + kq + r][ci = q*16 + row]
    const size_t slot = (size_t)b * gridDim.x * 4 + c;
    float* dst = part + (slot * 27 + t) * (size_t)(COUT * CIN);
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int q = 0; q < CT; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                dst[(size_t)(m * 16 + kq + r) * CIN + q * 16 + row] = acc[m][q][r];
}

// sum the partial slots in a fixed order into the FP32 gradient
__global__ void k_wgrad_reduce(const float* __restrict__ part, float* __restrict__ gw,
                               uint32_t elems, int nslots) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < elems; e += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int q = 0; q < nslots; ++q) s += part[(size_t)q * elems + e];
        gw[e] = s;
    }
}

// bias gradient: gb[co] = sum_b sum_v G[b][v][co]
__global__ void k_bias_grad(const float* __restrict__ G, float* __restrict__ gb, uint32_t nvox,
                            int cout) {
    __shared__ float red[256];
    const int co = blockIdx.x;
    float s = 0.f;
    for (uint32_t v = threadIdx.x; v < nvox; v += blockDim.x) s += G[(size_t)v * cout + co];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) gb[co] = red[0];
}

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
__global__ void k_relu_mask(float* __restrict__ g, const bf16* __restrict__ y, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        if (!((float)y[e] > 0.f)) g[e] = 0.f;
}

// 2x2x2 average pool, channels-last, per sample
__global__ void k_pool(const bf16* __restrict__ X, bf16* __restrict__ Y, NetLvl Lf, NetLvl Lc,
                       int C, int batch) {
    const size_t total = (size_t)batch * Lc.n * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int ch = (int)(e % C);
        const size_t bv = e / C;
        const uint32_t v = (uint32_t)(bv % Lc.n);
        const size_t b = bv / Lc.n;
        int I, J, K;
        vdecode(v, Lc, I, J, K);
        float s = 0.f;
        for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb)
                for (int cc = 0; cc < 2; ++cc) {
                    const uint32_t f = ((uint32_t)(2 * I + a) * Lf.nj + (uint32_t)(2 * J + bb)) * Lf.nk +
                                       (uint32_t)(2 * K + cc);
                    s += (float)X[(b * Lf.n + f) * C + ch];
                }
        Y[e] = (bf16)(s * 0.125f);
    }
}

// G0 (fine, C channels) (+)= G1(coarse parent) / 8    (pool backward)
__global__ void k_pool_bwd(const float* __restrict__ G1, float* __restrict__ G0, NetLvl Lf,
                           NetLvl Lc, int C, int batch, int accumulate) {
    const size_t total = (size_t)batch * Lf.n * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int ch = (int)(e % C);
        const size_t bv = e / C;
        const uint32_t v = (uint32_t)(bv % Lf.n);
        const size_t b = bv / Lf.n;
        int i, j, k;
        vdecode(v, Lf, i, j, k);
        const uint32_t p = ((uint32_t)(i >> 1) * Lc.nj + (uint32_t)(j >> 1)) * Lc.nk + (uint32_t)(k >> 1);
        const float val = G1[(b * Lc.n + p) * C + ch] * 0.125f;
        G0[e] = accumulate ? G0[e] + val : val;
    }
}

// From the full input gradient of a concat layer (CIN = CD + CS channels):
// deep part summed over the 8 children into Gdeep (coarse), skip part into
// Gskip (fine).
__global__ void k_cat_bwd(const float* __restrict__ Gin, float* __restrict__ Gdeep,
                          float* __restrict__ Gskip, NetLvl Lf, NetLvl Lc, int CD, int CS,
                          int batch, int acc_deep, int acc_skip) {
    const int CIN = CD + CS;
    const size_t tdeep = (size_t)batch * Lc.n * CD;
    const size_t tskip = (size_t)batch * Lf.n * CS;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tdeep + tskip;
         e += (size_t)gridDim.x * blockDim.x) {
        if (e < tdeep) {
            const int ch = (int)(e % CD);
            const size_t bv = e / CD;
            const uint32_t v = (uint32_t)(bv % Lc.n);
            const size_t b = bv / Lc.n;
            int I, J, K;
            vdecode(v, Lc, I, J, K);
            float s = 0.f;
            for (int a = 0; a < 2; ++a)
                for (int bb = 0; bb < 2; ++bb)
                    for (int cc = 0; cc < 2; ++cc) {
                        const uint32_t f = ((uint32_t)(2 * I + a) * Lf.nj + (uint32_t)(2 * J + bb)) *
                                               Lf.nk + (uint32_t)(2 * K + cc);
                        s += Gin[(b * Lf.n + f) * CIN + ch];
                    }
            Gdeep[e] = acc_deep ? Gdeep[e] + s : s;
        } else {
            const size_t e2 = e - tdeep;
            const int ch = (int)(e2 % CS);
            const size_t bv = e2 / CS;
            const float val = Gin[bv * CIN + CD + ch];
            Gskip[e2] = acc_skip ? Gskip[e2] + val : val;
        }
    }
}

__global__ void k_add_into(float* __restrict__ dst, const float* __restrict__ src, size_t n, int acc) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        dst[e] = acc ? dst[e] + src[e] : src[e];
}

// masked MSE on output channel 0: loss partials and the output gradient
__global__ void k_mse(const float* __restrict__ out, const float* __restrict__ target,
                      const float* __restrict__ mask, float* __restrict__ gout, size_t nvox,
                      double inv_count, double* __restrict__ partials) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox;
         v += (size_t)gridDim.x * blockDim.x) {
        const float m = mask[v];
        const float d = (out[v * 16] - target[v]) * m;
        s += (double)d * d;
        gout[v * 16] = (float)(2.0 * inv_count) * d;
#pragma unroll
        for (int c = 1; c < 16; ++c) gout[v * 16 + c] = 0.f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0] * inv_count;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, const uint8_t* __restrict__ frozen, size_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        if (frozen[e]) continue;
        const float ge = g[e];
        const float me = b1 * m[e] + (1.f - b1) * ge;
        const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        w[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
    }
}

// forward BF16 weights and the mirrored-transposed copy for backward-data
__global__ void k_w_convert(const float* __restrict__ w, bf16* __restrict__ wf,
                            bf16* __restrict__ wm, int cout, int cin) {
    const int n = 27 * cout * cin;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int t = e / (cout * cin);
        const int r = e - t * cout * cin;
        const int co = r / cin, ci = r - co * cin;
        const bf16 x = (bf16)w[e];
        wf[e] = x;
        wm[((size_t)(26 - t) * cin + ci) * cout + co] = x;
    }
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        b[e] = (bf16)a[e];
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
static int eb(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

#define NET_LAUNCH_CHECK()                                                       \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess) {                                                  \
            net->err = std::string("kernel launch: ") + hipGetErrorString(e_);   \
            return -11;                                                          \
        }                                                                        \
    } while (0)

template <int CA, int CB, int COUT, bool UPA, bool RELU, typename TIN, typename TOUT>
static void conv(UNet* net, const TIN* A, const TIN* B, const bf16* W, const float* bias, TOUT* Y,
                 const NetLvl& L, const NetLvl& LA, int batch) {
    if (net->conv_tiled) {
        if constexpr (CA == 16 && CB == 0 && COUT == 16 && !UPA) {
            const int nkt = (L.nk + CT_TK - 1) / CT_TK, njt = (L.nj + CT_TJ - 1) / CT_TJ;
            dim3 grd(nkt * njt, (L.ni + CT_ICH - 1) / CT_ICH, batch);
            hipLaunchKernelGGL((k_conv16_tiled<RELU, TIN, TOUT>), grd, dim3(256), 0, net->stream, A,
                               W, bias, Y, L);
            return;
        } else if constexpr ((CA == 16 && CB == 0 && COUT == 32 && !UPA) ||
                             (CA == 32 && CB == 0 && COUT == 32 && !UPA) ||
                             (CA == 32 && CB == 0 && COUT == 64 && !UPA) ||
                             (CA == 32 && CB == 16 && COUT == 16 && UPA)) {
            // the layers where the tiled form was measured faster: e1a, e1b,
            // d1b (and their backward-data), e2a, d0a; with 64 or more input
            // channels the weights no longer fit in registers and the general
            // kernel's 64-voxel weight reuse wins
            using T = ConvTile<CA + CB, COUT>;
            const int nkt = (L.nk + T::TK - 1) / T::TK, njt = (L.nj + T::TJ - 1) / T::TJ;
            dim3 grd(nkt * njt, (L.ni + CT_ICH - 1) / CT_ICH, batch);
            hipLaunchKernelGGL((k_conv_tiled<CA, CB, COUT, UPA, RELU, TIN, TOUT>), grd,
                               dim3(64 * T::TJ * T::MS), 0, net->stream, A, B, W, bias, Y, L, LA);
            return;
        }
    }
    const uint32_t waves = (L.n + 63) / 64;
    dim3 grd((waves + 3) / 4, batch);
    hipLaunchKernelGGL((k_conv<CA, CB, COUT, UPA, RELU, TIN, TOUT>), grd, dim3(256), 0, net->stream,
                       A, B, W, bias, Y, L, LA);
}

template <int CA, int CB, int COUT, bool UPA>
static void wgrad(UNet* net, const bf16* A, const bf16* B, const float* G, float* gw,
                  const NetLvl& L, const NetLvl& LA, int batch) {
    constexpr int CIN = CA + CB;
    uint32_t chunk = 4096;
    int nchunk = (int)((L.n + chunk - 1) / chunk);
    const int bx = (nchunk + 3) / 4;
    const int nslots = batch * bx * 4;
    hipLaunchKernelGGL((k_wgrad<CA, CB, COUT, UPA>), dim3(bx, 27, batch), dim3(256), 0, net->stream,
                       A, B, G, net->part, L, LA, chunk, nchunk);
    const uint32_t elems = 27u * COUT * CIN;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((elems + 255) / 256), dim3(256), 0, net->stream,
                       net->part, gw, elems, nslots);
}

size_t unet_part_floats(const NetLvl& L0, int batch) {
    // largest slots * 27 * cout * cin over the layers
    size_t best = 0;
    for (int l = 0; l < NET_LAYERS; ++l) {
        const NetLvl& L = L0;
        (void)L;
        const uint32_t n = L0.n >> (3 * kLayerLevel[l]);
        const size_t nchunk = (n + 4095) / 4096;
        const size_t slots = (size_t)batch * ((nchunk + 3) / 4) * 4;
        const size_t s = slots * 27 * kLayerCout[l] * kLayerCin[l];
        if (s > best) best = s;
    }
    return best;
}

int unet_forward(UNet* net, int batch) {
    const NetLvl* L = net->L;
    bf16** a = net->act;
    const bf16* const* W = net->wf;
    float* const* bias = net->bias;
    conv<16, 0, 16, false, true>(net, net->x, (const bf16*)nullptr, W[0], bias[0], a[0], L[0], L[0], batch);
    conv<16, 0, 16, false, true>(net, (const bf16*)a[0], (const bf16*)nullptr, W[1], bias[1], a[1], L[0], L[0], batch);
    hipLaunchKernelGGL(k_pool, dim3(eb((size_t)batch * L[1].n * 16)), dim3(256), 0, net->stream, a[1], net->pool0, L[0], L[1], 16, batch);
    conv<16, 0, 32, false, true>(net, (const bf16*)net->pool0, (const bf16*)nullptr, W[2], bias[2], a[2], L[1], L[1], batch);
    conv<32, 0, 32, false, true>(net, (const bf16*)a[2], (const bf16*)nullptr, W[3], bias[3], a[3], L[1], L[1], batch);
    hipLaunchKernelGGL(k_pool, dim3(eb((size_t)batch * L[2].n * 32)), dim3(256), 0, net->stream, a[3], net->pool1, L[1], L[2], 32, batch);
    conv<32, 0, 64, false, true>(net, (const bf16*)net->pool1, (const bf16*)nullptr, W[4], bias[4], a[4], L[2], L[2], batch);
    conv<64, 0, 64, false, true>(net, (const bf16*)a[4], (const bf16*)nullptr, W[5], bias[5], a[5], L[2], L[2], batch);
    conv<64, 32, 32, true, true>(net, (const bf16*)a[5], (const bf16*)a[3], W[6], bias[6], a[6], L[1], L[2], batch);
    conv<32, 0, 32, false, true>(net, (const bf16*)a[6], (const bf16*)nullptr, W[7], bias[7], a[7], L[1], L[1], batch);
    conv<32, 16, 16, true, true>(net, (const bf16*)a[7], (const bf16*)a[1], W[8], bias[8], a[8], L[0], L[1], batch);
    conv<16, 0, 16, false, true>(net, (const bf16*)a[8], (const bf16*)nullptr, W[9], bias[9], a[9], L[0], L[0], batch);
    conv<16, 0, 16, false, false>(net, (const bf16*)a[9], (const bf16*)nullptr, W[10], bias[10], net->out, L[0], L[0], batch);
    NET_LAUNCH_CHECK();
    return 0;
}

// Backward from net->gout (gradient of the FP32 output, 16 channels) to the
// FP32 weight gradient net->g (flat layout of the weights).
int unet_backward(UNet* net, int batch) {
    const NetLvl* L = net->L;
    bf16** a = net->act;
    float** g = net->gact;
    const bf16* const* Wm = net->wm;
    float* gw[NET_LAYERS];
    float* gb[NET_LAYERS];
    for (int l = 0; l < NET_LAYERS; ++l) {
        gw[l] = net->g + net->woff[l];
        gb[l] = net->g + net->boff[l];
    }
    auto relu_mask = [&](int l) {
        const size_t n = (size_t)batch * L[kLayerLevel[l]].n * kLayerCout[l];
        hipLaunchKernelGGL(k_relu_mask, dim3(eb(n)), dim3(256), 0, net->stream, g[l], a[l], n);
    };
    auto bias_grad = [&](int l) {
        hipLaunchKernelGGL(k_bias_grad, dim3(kLayerCout[l]), dim3(256), 0, net->stream, g[l], gb[l],
                           (uint32_t)(batch * L[kLayerLevel[l]].n), kLayerCout[l]);
    };
    // out (10): input a9
    bias_grad(10);
    wgrad<16, 0, 16, false>(net, a[9], nullptr, net->gout, gw[10], L[0], L[0], batch);
    conv<16, 0, 16, false, false>(net, (const float*)net->gout, (const float*)nullptr, Wm[10], nullptr, g[9], L[0], L[0], batch);
    // d0b (9): input a8
    relu_mask(9);
    bias_grad(9);
    wgrad<16, 0, 16, false>(net, a[8], nullptr, g[9], gw[9], L[0], L[0], batch);
    conv<16, 0, 16, false, false>(net, (const float*)g[9], (const float*)nullptr, Wm[9], nullptr, g[8], L[0], L[0], batch);
    // d0a (8): input [up(a7) 32, a1 16]
    relu_mask(8);
    bias_grad(8);
    wgrad<32, 16, 16, true>(net, a[7], a[1], g[8], gw[8], L[0], L[1], batch);
    conv<16, 0, 48, false, false>(net, (const float*)g[8], (const float*)nullptr, Wm[8], nullptr, net->gtmp, L[0], L[0], batch);
    hipLaunchKernelGGL(k_cat_bwd, dim3(eb((size_t)batch * (L[1].n * 32 + L[0].n * 16))), dim3(256), 0, net->stream,
                       net->gtmp, g[7], g[1], L[0], L[1], 32, 16, batch, 0, 0);
    // d1b (7): input a6
    relu_mask(7);
    bias_grad(7);
    wgrad<32, 0, 32, false>(net, a[6], nullptr, g[7], gw[7], L[1], L[1], batch);
    conv<32, 0, 32, false, false>(net, (const float*)g[7], (const float*)nullptr, Wm[7], nullptr, g[6], L[1], L[1], batch);
    // d1a (6): input [up(a5) 64, a3 32]
    relu_mask(6);
    bias_grad(6);
    wgrad<64, 32, 32, true>(net, a[5], a[3], g[6], gw[6], L[1], L[2], batch);
    conv<32, 0, 96, false, false>(net, (const float*)g[6], (const float*)nullptr, Wm[6], nullptr, net->gtmp, L[1], L[1], batch);
    hipLaunchKernelGGL(k_cat_bwd, dim3(eb((size_t)batch * (L[2].n * 64 + L[1].n * 32))), dim3(256), 0, net->stream,
                       net->gtmp, g[5], g[3], L[1], L[2], 64, 32, batch, 0, 0);
    // e2b (5): input a4
    relu_mask(5);
    bias_grad(5);
    wgrad<64, 0, 64, false>(net, a[4], nullptr, g[5], gw[5], L[2], L[2], batch);
    conv<64, 0, 64, false, false>(net, (const float*)g[5], (const float*)nullptr, Wm[5], nullptr, g[4], L[2], L[2], batch);
    // e2a (4): input pool1 = pool(a3)
    relu_mask(4);
    bias_grad(4);
    wgrad<32, 0, 64, false>(net, net->pool1, nullptr, g[4], gw[4], L[2], L[2], batch);
    conv<64, 0, 32, false, false>(net, (const float*)g[4], (const float*)nullptr, Wm[4], nullptr, net->gtmp, L[2], L[2], batch);
    hipLaunchKernelGGL(k_pool_bwd, dim3(eb((size_t)batch * L[1].n * 32)), dim3(256), 0, net->stream,
                       net->gtmp, g[3], L[1], L[2], 32, batch, 1);
    // e1b (3): input a2
    relu_mask(3);
    bias_grad(3);
    wgrad<32, 0, 32, false>(net, a[2], nullptr, g[3], gw[3], L[1], L[1], batch);
    conv<32, 0, 32, false, false>(net, (const float*)g[3], (const float*)nullptr, Wm[3], nullptr, g[2], L[1], L[1], batch);
    // e1a (2): input pool0 = pool(a1)
    relu_mask(2);
    bias_grad(2);
    wgrad<16, 0, 32, false>(net, net->pool0, nullptr, g[2], gw[2], L[1], L[1], batch);
    conv<32, 0, 16, false, false>(net, (const float*)g[2], (const float*)nullptr, Wm[2], nullptr, net->gtmp, L[1], L[1], batch);
    hipLaunchKernelGGL(k_pool_bwd, dim3(eb((size_t)batch * L[0].n * 16)), dim3(256), 0, net->stream,
                       net->gtmp, g[1], L[0], L[1], 16, batch, 1);
    // e0b (1): input a0
    relu_mask(1);
    bias_grad(1);
    wgrad<16, 0, 16, false>(net, a[0], nullptr, g[1], gw[1], L[0], L[0], batch);
    conv<16, 0, 16, false, false>(net, (const float*)g[1], (const float*)nullptr, Wm[1], nullptr, g[0], L[0], L[0], batch);
    // e0a (0): input x
    relu_mask(0);
    bias_grad(0);
    wgrad<16, 0, 16, false>(net, net->x, nullptr, g[0], gw[0], L[0], L[0], batch);
    NET_LAUNCH_CHECK();
    return 0;
}

int unet_convert_weights(UNet* net) {
    for (int l = 0; l < NET_LAYERS; ++l) {
        const int n = 27 * kLayerCout[l] * kLayerCin[l];
        hipLaunchKernelGGL(k_w_convert, dim3((n + 255) / 256), dim3(256), 0, net->stream,
                           net->w + net->woff[l], net->wf[l], net->wm[l], kLayerCout[l], kLayerCin[l]);
        net->bias[l] = net->w + net->boff[l];
    }
    NET_LAUNCH_CHECK();
    return 0;
}

int unet_mse(UNet* net, const float* target, const float* mask, int batch, double* loss) {
    const size_t nvox = (size_t)batch * net->L[0].n;
    double cnt = net->mask_count > 0 ? net->mask_count : 1.0;
    const int nb = eb(nvox) > 4096 ? 4096 : eb(nvox);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, net->stream, net->out, target, mask, net->gout,
                       nvox, 1.0 / cnt, net->dpart);
    NET_LAUNCH_CHECK();
    std::vector<double> h(nb);
    if (hipMemcpyAsync(h.data(), net->dpart, nb * sizeof(double), hipMemcpyDeviceToHost, net->stream) != hipSuccess ||
        hipStreamSynchronize(net->stream) != hipSuccess) {
        net->err = "loss readback failed";
        return -10;
    }
    double s = 0.0;
    for (double x : h) s += x;
    *loss = s;
    return 0;
}

int unet_adam(UNet* net, float lr) {
    net->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)net->adam_t), c2 = 1.f - powf(b2, (float)net->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(eb(net->nw)), dim3(256), 0, net->stream, net->w, net->am, net->av,
                       net->g, net->frozen, net->nw, lr, b1, b2, eps, c1, c2);
    NET_LAUNCH_CHECK();
    return unet_convert_weights(net);
}

int unet_load_input(UNet* net, const float* in_dev, int batch) {
    const size_t n = (size_t)batch * net->L[0].n * 16;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(eb(n)), dim3(256), 0, net->stream, in_dev, net->x, n);
    NET_LAUNCH_CHECK();
    return 0;
}

}  // namespace ls
