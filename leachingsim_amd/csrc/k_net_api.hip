// k_net_api.hip — C ABI of the U-Net (training and standalone inference on
// host buffers) and its device allocation.  See include/lsgpu.h.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ls_unet.h"

using namespace ls;

namespace ls {

template <typename T>
static int nalloc(UNet* net, T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) + 256);
    if (e != hipSuccess) {
        net->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return -12;
    }
    net->allocations.push_back(q);
    net->bytes += n * sizeof(T) + 256;
    *p = (T*)q;
    return 0;
}

int unet_alloc(UNet* net, int ni, int nj, int nk, int batch, bool trainable, hipStream_t stream) {
    if (ni % 16 || nj % 16 || nk % 16 || ni < 16 || nj < 16 || nk < 16 || batch < 1) {
        net->err = "network grid must be multiples of 16 in every direction";
        return -2;
    }
    net->batch = batch;
    net->trainable = trainable;
    for (int l = 0; l < 3; ++l) {
        net->L[l].ni = ni >> l;
        net->L[l].nj = nj >> l;
        net->L[l].nk = nk >> l;
        net->L[l].n = (uint32_t)net->L[l].ni * net->L[l].nj * net->L[l].nk;
    }
    if (stream) {
        net->stream = stream;
    } else {
        if (hipStreamCreate(&net->stream) != hipSuccess) {
            net->err = "no HIP device";
            return -4;
        }
        net->own_stream = true;
    }
    size_t o = 0;
    for (int l = 0; l < NET_LAYERS; ++l) {
        net->woff[l] = o;
        o += 27 * (size_t)kLayerCout[l] * kLayerCin[l];
        net->boff[l] = o;
        o += kLayerCout[l];
    }
    net->nw = o;
    int st = 0;
    st |= nalloc(net, &net->w, o);
    st |= nalloc(net, &net->frozen, o);
    for (int l = 0; l < NET_LAYERS; ++l) {
        st |= nalloc(net, &net->wf[l], 27 * (size_t)kLayerCout[l] * kLayerCin[l]);
        st |= nalloc(net, &net->wm[l], 27 * (size_t)kLayerCout[l] * kLayerCin[l]);
    }
    const size_t B = (size_t)batch;
    st |= nalloc(net, &net->x, B * net->L[0].n * NET_CIN);
    for (int l = 0; l < NET_LAYERS - 1; ++l)
        st |= nalloc(net, &net->act[l], B * net->L[kLayerLevel[l]].n * kLayerCout[l]);
    st |= nalloc(net, &net->pool0, B * net->L[1].n * 16);
    st |= nalloc(net, &net->pool1, B * net->L[2].n * 32);
    st |= nalloc(net, &net->out, B * net->L[0].n * 16);
    st |= nalloc(net, &net->dpart, 4096);
    if (trainable) {
        st |= nalloc(net, &net->g, o);
        st |= nalloc(net, &net->am, o);
        st |= nalloc(net, &net->av, o);
        for (int l = 0; l < NET_LAYERS - 1; ++l)
            st |= nalloc(net, &net->gact[l], B * net->L[kLayerLevel[l]].n * kLayerCout[l]);
        st |= nalloc(net, &net->gout, B * net->L[0].n * 16);
        st |= nalloc(net, &net->gtmp, B * net->L[0].n * 48);
        st |= nalloc(net, &net->part, unet_part_floats(net->L[0], batch));
    }
    if (st) return -12;
    // frozen: output-layer channels 1..15 (weights and biases)
    std::vector<uint8_t> fz(o, 0);
    const int lo = NET_LAYERS - 1;
    for (int t = 0; t < 27; ++t)
        for (int co = 1; co < 16; ++co)
            for (int ci = 0; ci < kLayerCin[lo]; ++ci)
                fz[net->woff[lo] + ((size_t)t * 16 + co) * kLayerCin[lo] + ci] = 1;
    for (int co = 1; co < 16; ++co) fz[net->boff[lo] + co] = 1;
    if (hipMemcpy(net->frozen, fz.data(), o, hipMemcpyHostToDevice) != hipSuccess) return -10;
    if (hipMemset(net->w, 0, o * sizeof(float)) != hipSuccess) return -10;
    if (trainable) {
        if (hipMemset(net->am, 0, o * sizeof(float)) != hipSuccess ||
            hipMemset(net->av, 0, o * sizeof(float)) != hipSuccess)
            return -10;
    }
    return unet_convert_weights(net);
}

void unet_free(UNet* net) {
    if (!net) return;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    for (void* q : net->allocations) (void)hipFree(q);
    net->allocations.clear();
    if (net->own_stream && net->stream) (void)hipStreamDestroy(net->stream);
    net->stream = nullptr;
}

// ---------------------------------------------------------------------------
// Galerkin energy loss of the coarse correction, on the network's grid.
//
// With u = P out (P copies a coarse value to the solved fine children) the
// share of the A-norm error e that the best multiple of u leaves is
//   loss = 1 - (u.r)^2 / (u.A u  e.A e),   r = A e,
// and since u.r = out.(P^T r) and u.A u = out.(P^T A P) out, it is evaluated
// exactly on the coarse grid from the restricted residual rc = P^T r, the
// scalar e.A e and the Galerkin operator A_c = P^T A P, which for the 2x2x2
// aggregation of a 7-point operator is again 7-point and symmetric (stored as
// its diagonal and the couplings to the +i, +j, +k neighbours).  The gradient
//   d loss / d out = -2 (u.r) / (u.Au e.Ae) rc + 2 (u.r)^2 / ((u.Au)^2 e.Ae) A_c out
// is written as the output gradient of the backward pass, averaged over the
// batch.
// ---------------------------------------------------------------------------
constexpr int EN_MAXB = 8;        // samples per energy step
constexpr int EN_BLOCKS = 1024;   // dot partials per sample

struct EnBatch {
    const float* ac[EN_MAXB];
    const float* rc[EN_MAXB];
    double eae[EN_MAXB];
};

__global__ __launch_bounds__(256) void k_en_apply(const float* __restrict__ out, EnBatch eb_,
                                                  NetLvl L, float* __restrict__ au,
                                                  double* __restrict__ part) {
    __shared__ double lds[2 * 256 / WAVE];
    const int b = blockIdx.y;
    const float* __restrict__ ac = eb_.ac[b];
    const float* __restrict__ rc = eb_.rc[b];
    const uint32_t n = L.n, sj = (uint32_t)L.nk, si = (uint32_t)L.nk * L.nj;
    const float* __restrict__ u = out + (size_t)b * n * 16;
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const uint32_t ij = q / sj;
        const int k = (int)(q - ij * sj), i = (int)(ij / (uint32_t)L.nj),
                  j = (int)(ij - (uint32_t)i * L.nj);
        const float uq = u[(size_t)q * 16];
        float a = ac[q] * uq;
        if (i + 1 < L.ni) a += ac[n + q] * u[(size_t)(q + si) * 16];
        if (i > 0) a += ac[n + q - si] * u[(size_t)(q - si) * 16];
        if (j + 1 < L.nj) a += ac[2 * (size_t)n + q] * u[(size_t)(q + sj) * 16];
        if (j > 0) a += ac[2 * (size_t)n + q - sj] * u[(size_t)(q - sj) * 16];
        if (k + 1 < L.nk) a += ac[3 * (size_t)n + q] * u[(size_t)(q + 1) * 16];
        if (k > 0) a += ac[3 * (size_t)n + q - 1] * u[(size_t)(q - 1) * 16];
        au[(size_t)b * n + q] = a;
        s0 += (double)uq * (double)rc[q];
        s1 += (double)uq * (double)a;
    }
    double v[2] = {s0, s1};
    block_reduce<2, RED_SUM, RED_SUM>(v, lds);
    if (threadIdx.x == 0) {
        part[((size_t)b * 2 + 0) * gridDim.x + blockIdx.x] = v[0];
        part[((size_t)b * 2 + 1) * gridDim.x + blockIdx.x] = v[1];
    }
}

// per sample: loss and the two gradient coefficients (divided by the batch)
__global__ void k_en_coef(const double* __restrict__ part, int nb, EnBatch eb_, int batch,
                          double* __restrict__ coef) {
    __shared__ double lds[2 * 1024 / WAVE];
    const int b = blockIdx.x;
    double a0 = 0.0, a1 = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        a0 += part[((size_t)b * 2 + 0) * nb + t];
        a1 += part[((size_t)b * 2 + 1) * nb + t];
    }
    double v[2] = {a0, a1};
    block_reduce<2, RED_SUM, RED_SUM>(v, lds);
    if (threadIdx.x == 0) {
        const double ur = v[0], uau = v[1], eae = eb_.eae[b];
        double loss = 1.0, al = 0.0, be = 0.0;
        if (uau > 0.0 && eae > 0.0 && isfinite(ur) && isfinite(uau)) {
            loss = 1.0 - ur * ur / (uau * eae);
            al = -2.0 * ur / (uau * eae) / batch;
            be = 2.0 * ur * ur / (uau * uau * eae) / batch;
        }
        coef[4 * b + 0] = loss;
        coef[4 * b + 1] = al;
        coef[4 * b + 2] = be;
        coef[4 * b + 3] = ur;
    }
}

// output gradient: channel 0 = alpha rc + beta A_c out, channels 1..15 = 0
__global__ __launch_bounds__(256) void k_en_grad(EnBatch eb_, const float* __restrict__ au,
                                                 const double* __restrict__ coef, uint32_t n,
                                                 float* __restrict__ gout) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const int b = blockIdx.y;
    const float al = (float)coef[4 * b + 1], be = (float)coef[4 * b + 2];
    const float* __restrict__ rc = eb_.rc[b];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const size_t o = ((size_t)b * n + q) * 16;
        const f4v z = f4v{0.f, 0.f, 0.f, 0.f};
        f4v g0 = z;
        g0[0] = al * rc[q] + be * au[(size_t)b * n + q];
        *(f4v*)(gout + o) = g0;
        *(f4v*)(gout + o + 4) = z;
        *(f4v*)(gout + o + 8) = z;
        *(f4v*)(gout + o + 12) = z;
    }
}

__global__ void k_f32_to_bf16_store(const float* __restrict__ a, __bf16* __restrict__ b, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        b[e] = (__bf16)a[e];
}

}  // namespace ls

struct ls_net : UNet {};

static int sync(ls_net* n) {
    hipError_t e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) {
        n->err = std::string("stream: ") + hipGetErrorString(e);
        return -10;
    }
    return 0;
}

extern "C" {

size_t ls_net_num_weights(void) {
    size_t o = 0;
    for (int l = 0; l < NET_LAYERS; ++l) o += 27 * (size_t)kLayerCout[l] * kLayerCin[l] + kLayerCout[l];
    return o;
}

int ls_net_create(ls_net** out, int ni, int nj, int nk, int batch, int trainable) {
    if (!out) return -1;
    ls_net* n = new ls_net();
    int st = unet_alloc(n, ni, nj, nk, batch, trainable != 0, nullptr);
    *out = n;
    if (st == 0) st = sync(n);
    return st;
}

void ls_net_destroy(ls_net* n) {
    if (!n) return;
    unet_free(n);
    delete n;
}

const char* ls_net_last_error(const ls_net* n) { return n ? n->err.c_str() : "null network"; }

int ls_net_set_weights(ls_net* n, const float* w, size_t count) {
    if (!n || !w || count != n->nw) {
        if (n) n->err = "weight count mismatch";
        return -3;
    }
    if (hipMemcpyAsync(n->w, w, count * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess)
        return -10;
    n->adam_t = 0;
    if (n->trainable) {
        (void)hipMemsetAsync(n->am, 0, count * sizeof(float), n->stream);
        (void)hipMemsetAsync(n->av, 0, count * sizeof(float), n->stream);
    }
    int st = unet_convert_weights(n);
    return st ? st : sync(n);
}

int ls_net_get_weights(ls_net* n, float* w, size_t count) {
    if (!n || !w || count != n->nw) return -3;
    if (hipMemcpyAsync(w, n->w, count * sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess)
        return -10;
    return sync(n);
}

// in: [batch][ni][nj][nk][16] FP32 (rounded to BF16 on the device);
// out: [batch][ni][nj][nk] FP32, output channel 0.
int ls_net_forward(ls_net* n, const float* in, float* out) {
    if (!n || !in || !out) return -1;
    const size_t nv = (size_t)n->batch * n->L[0].n;
    float* tmp = nullptr;
    if (hipMalloc(&tmp, nv * 16 * sizeof(float)) != hipSuccess) return -12;
    int st = 0;
    if (hipMemcpyAsync(tmp, in, nv * 16 * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess)
        st = -10;
    if (!st) st = unet_load_input(n, tmp, n->batch);
    if (!st) st = unet_forward(n, n->batch);
    std::vector<float> full;
    if (!st) {
        full.resize(nv * 16);
        if (hipMemcpyAsync(full.data(), n->out, nv * 16 * sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess)
            st = -10;
    }
    if (!st) st = sync(n);
    (void)hipFree(tmp);
    if (st) return st;
    for (size_t v = 0; v < nv; ++v) out[v] = full[v * 16];
    return 0;
}

// Gradient of sum(gout * out[..., 0]) with respect to the weights, for the
// input of the last ls_net_forward call.
int ls_net_backward(ls_net* n, const float* gout, float* gw) {
    if (!n || !gout || !gw) return -1;
    if (!n->trainable) {
        n->err = "network was created without training buffers";
        return -3;
    }
    const size_t nv = (size_t)n->batch * n->L[0].n;
    std::vector<float> full(nv * 16, 0.0f);
    for (size_t v = 0; v < nv; ++v) full[v * 16] = gout[v];
    if (hipMemcpyAsync(n->gout, full.data(), nv * 16 * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess)
        return -10;
    int st = unet_backward(n, n->batch);
    if (st) return st;
    if (hipMemcpyAsync(gw, n->g, n->nw * sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess)
        return -10;
    return sync(n);
}

// One Adam step on the masked MSE between output channel 0 and target.
// in: [batch][N][16]; target, mask: [batch][N].  Returns the loss.
int ls_net_train_step(ls_net* n, const float* in, const float* target, const float* mask, float lr,
                      double* loss) {
    if (!n || !in || !target || !mask || !loss) return -1;
    if (!n->trainable) return -3;
    const size_t nv = (size_t)n->batch * n->L[0].n;
    float* tmp = nullptr;
    if (hipMalloc(&tmp, nv * 18 * sizeof(float)) != hipSuccess) return -12;
    int st = 0;
    if (hipMemcpyAsync(tmp, in, nv * 16 * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess ||
        hipMemcpyAsync(tmp + nv * 16, target, nv * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess ||
        hipMemcpyAsync(tmp + nv * 17, mask, nv * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess)
        st = -10;
    double cnt = 0.0;
    for (size_t v = 0; v < nv; ++v) cnt += mask[v];
    n->mask_count = cnt;
    if (!st) st = unet_load_input(n, tmp, n->batch);
    if (!st) st = unet_forward(n, n->batch);
    if (!st) st = unet_mse(n, tmp + nv * 16, tmp + nv * 17, n->batch, loss);
    if (!st) st = unet_backward(n, n->batch);
    if (!st) st = unet_adam(n, lr);
    if (!st) st = sync(n);
    (void)hipFree(tmp);
    return st;
}

// One Adam step from an output gradient given by the caller (any loss): the
// backward of the last forward with gout as d(loss)/d(output channel 0).
int ls_net_step(ls_net* n, const float* gout, float lr) {
    if (!n || !gout) return -1;
    if (!n->trainable) return -3;
    const size_t nv = (size_t)n->batch * n->L[0].n;
    std::vector<float> full(nv * 16, 0.0f);
    for (size_t v = 0; v < nv; ++v) full[v * 16] = gout[v];
    if (hipMemcpyAsync(n->gout, full.data(), nv * 16 * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess)
        return -10;
    int st = unet_backward(n, n->batch);
    if (!st) st = unet_adam(n, lr);
    if (!st) st = sync(n);
    return st;
}

// Device-resident training store for ls_net_energy_step: `samples` slots
// (input [n][16] as BF16, restricted residual [n], e.A e) and `operators`
// Galerkin operators ([4][n]).  Needs a trainable network.
int ls_net_store_alloc(ls_net* n, int samples, int operators) {
    if (!n) return -1;
    if (!n->trainable || samples < 1 || operators < 1 || n->batch > EN_MAXB) {
        n->err = "store: needs a trainable network, batch <= 8 and positive capacities";
        return -3;
    }
    if (n->st_cap) {
        n->err = "store: already allocated";
        return -3;
    }
    const size_t cn = n->L[0].n;
    int st = 0;
    st |= nalloc(n, &n->st_x, (size_t)samples * cn * 16);
    st |= nalloc(n, &n->st_rc, (size_t)samples * cn);
    st |= nalloc(n, &n->st_ac, (size_t)operators * 4 * cn);
    st |= nalloc(n, &n->st_tmp, cn * 16);
    st |= nalloc(n, &n->st_au, (size_t)n->batch * cn);
    st |= nalloc(n, &n->st_part, (size_t)n->batch * 2 * EN_BLOCKS);
    st |= nalloc(n, &n->st_coef, (size_t)n->batch * 4);
    if (st) return -12;
    n->st_cap = samples;
    n->st_nops = operators;
    n->st_eae.assign(samples, 0.0);
    n->st_op.assign(samples, -1);
    return 0;
}

// ac: [4][n] FP32 (diagonal; couplings of voxel q to q + i, q + j, q + k,
// zero on the last plane of each direction).
int ls_net_store_operator(ls_net* n, int slot, const float* ac) {
    if (!n || !ac) return -1;
    if (slot < 0 || slot >= n->st_nops) {
        n->err = "store: operator slot out of range";
        return -3;
    }
    const size_t cn = n->L[0].n;
    if (hipMemcpyAsync(n->st_ac + (size_t)slot * 4 * cn, ac, 4 * cn * sizeof(float),
                       hipMemcpyHostToDevice, n->stream) != hipSuccess)
        return -10;
    return sync(n);
}

// x: [n][16] FP32 network input (rounded to BF16), rc: [n] FP32 = P^T r,
// eae = e.A e of the sample's error; op_slot: its Galerkin operator.
int ls_net_store_sample(ls_net* n, int slot, int op_slot, const float* x, const float* rc,
                        double eae) {
    if (!n || !x || !rc) return -1;
    if (slot < 0 || slot >= n->st_cap || op_slot < 0 || op_slot >= n->st_nops) {
        n->err = "store: sample or operator slot out of range";
        return -3;
    }
    const size_t cn = n->L[0].n;
    if (hipMemcpyAsync(n->st_tmp, x, cn * 16 * sizeof(float), hipMemcpyHostToDevice,
                       n->stream) != hipSuccess ||
        hipMemcpyAsync(n->st_rc + (size_t)slot * cn, rc, cn * sizeof(float),
                       hipMemcpyHostToDevice, n->stream) != hipSuccess)
        return -10;
    const size_t ne = cn * 16;
    const int nb = (int)std::min<size_t>((ne + 255) / 256, 65536);
    hipLaunchKernelGGL(k_f32_to_bf16_store, dim3(nb), dim3(256), 0, n->stream, n->st_tmp,
                       n->st_x + (size_t)slot * ne, ne);
    if (hipGetLastError() != hipSuccess) return -10;
    n->st_eae[slot] = eae;
    n->st_op[slot] = op_slot;
    return sync(n);
}

// One step of the Galerkin energy loss on the stored samples `slots`
// (batch of them): forward, loss and output gradient on the device; when
// lr > 0 backward and Adam, else when grad_w is given backward only, with
// the weight gradient copied to grad_w.  losses[b] receives each sample's
// loss.
int ls_net_energy_step(ls_net* n, const int* slots, float lr, double* losses, float* grad_w) {
    if (!n || !slots || !losses) return -1;
    if (!n->st_cap) {
        n->err = "store: not allocated";
        return -3;
    }
    const int B = n->batch;
    const size_t cn = n->L[0].n;
    EnBatch eb_;
    for (int b = 0; b < EN_MAXB; ++b) {
        eb_.ac[b] = nullptr;
        eb_.rc[b] = nullptr;
        eb_.eae[b] = 0.0;
    }
    for (int b = 0; b < B; ++b) {
        const int s = slots[b];
        if (s < 0 || s >= n->st_cap || n->st_op[s] < 0) {
            n->err = "store: empty or out-of-range sample slot";
            return -3;
        }
        if (hipMemcpyAsync(n->x + (size_t)b * cn * 16, n->st_x + (size_t)s * cn * 16,
                           cn * 16 * sizeof(__bf16), hipMemcpyDeviceToDevice,
                           n->stream) != hipSuccess)
            return -10;
        eb_.ac[b] = n->st_ac + (size_t)n->st_op[s] * 4 * cn;
        eb_.rc[b] = n->st_rc + (size_t)s * cn;
        eb_.eae[b] = n->st_eae[s];
    }
    int st = unet_forward(n, B);
    if (st) return st;
    hipLaunchKernelGGL(k_en_apply, dim3(EN_BLOCKS, B), dim3(256), 0, n->stream, n->out, eb_,
                       n->L[0], n->st_au, n->st_part);
    hipLaunchKernelGGL(k_en_coef, dim3(B), dim3(1024), 0, n->stream, n->st_part, EN_BLOCKS, eb_, B,
                       n->st_coef);
    if (lr > 0.0f || grad_w != nullptr) {
        hipLaunchKernelGGL(k_en_grad, dim3(EN_BLOCKS, B), dim3(256), 0, n->stream, eb_, n->st_au,
                           n->st_coef, (uint32_t)cn, n->gout);
        if (hipGetLastError() != hipSuccess) return -10;
        st = unet_backward(n, B);
        if (!st && lr > 0.0f) st = unet_adam(n, lr);
        if (st) return st;
        if (lr <= 0.0f &&
            hipMemcpyAsync(grad_w, n->g, n->nw * sizeof(float), hipMemcpyDeviceToHost,
                           n->stream) != hipSuccess)
            return -10;
    }
    if (hipGetLastError() != hipSuccess) return -10;
    std::vector<double> h(4 * (size_t)B);
    if (hipMemcpyAsync(h.data(), n->st_coef, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                       n->stream) != hipSuccess)
        return -10;
    st = sync(n);
    if (st) return st;
    for (int b = 0; b < B; ++b) losses[b] = h[4 * b];
    return 0;
}

int ls_net_set_conv(ls_net* n, int tiled) {
    if (!n) return -1;
    n->conv_tiled = tiled != 0;
    return 0;
}

// Time `reps` forward passes on the device-resident input (HIP events).
int ls_net_time_forward(ls_net* n, int reps, double* ms) {
    if (!n || !ms || reps < 1) return -1;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -10;
    int st = unet_forward(n, n->batch);
    if (!st && hipEventRecord(a, n->stream) != hipSuccess) st = -10;
    for (int r = 0; r < reps && !st; ++r) st = unet_forward(n, n->batch);
    if (!st && hipEventRecord(b, n->stream) != hipSuccess) st = -10;
    float f = 0.0f;
    if (!st && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&f, a, b) != hipSuccess))
        st = -10;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)f / reps;
    return st;
}

}  // extern "C"
