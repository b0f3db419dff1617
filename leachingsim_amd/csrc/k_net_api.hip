// k_net_api.hip — C ABI of the U-Net (training and standalone inference on
// host buffers) and its device allocation.  See include/lsgpu.h.
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
