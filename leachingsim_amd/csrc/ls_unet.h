// ls_unet.h — device state of the 3-level U-Net (see k_unet.hip) and its
// internal entry points.  Layer order and weight layout match
// leachingsim_amd/unet_spec.py: e0a e0b e1a e1b e2a e2b d1a d1b d0a d0b out,
// each W[27][Cout][Cin] (FP32 master copy) then b[Cout].
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ls_common.h"

constexpr int NET_LAYERS = 11;
constexpr int NET_CIN = 16;
constexpr int kLayerCin[NET_LAYERS] = {16, 16, 16, 32, 32, 64, 96, 32, 48, 16, 16};
constexpr int kLayerCout[NET_LAYERS] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16};
constexpr int kLayerLevel[NET_LAYERS] = {0, 0, 1, 1, 2, 2, 1, 1, 0, 0, 0};

struct NetLvl {
    int ni, nj, nk;
    uint32_t n;  // voxels per sample
};

struct UNet {
    int batch = 1;
    NetLvl L[3];
    bool conv_tiled = true;      // LDS-tiled convolution kernels (else the general one)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    // weights
    size_t nw = 0;
    size_t woff[NET_LAYERS], boff[NET_LAYERS];
    float* w = nullptr;          // FP32 master (flat)
    float* g = nullptr;          // FP32 gradient (flat)
    float* am = nullptr;         // Adam moments
    float* av = nullptr;
    uint8_t* frozen = nullptr;   // 1 = never updated (unused output channels)
    int adam_t = 0;
    __bf16* wf[NET_LAYERS];      // forward BF16 weights
    __bf16* wm[NET_LAYERS];      // mirrored-transposed BF16 weights (backward-data)
    float* bias[NET_LAYERS];
    // activations
    __bf16* x = nullptr;         // input, 16 channels
    __bf16* act[NET_LAYERS - 1];
    __bf16* pool0 = nullptr;     // pool(e0b), level 1, 16 ch
    __bf16* pool1 = nullptr;     // pool(e1b), level 2, 32 ch
    float* out = nullptr;        // output, 16 channels FP32 (channel 0 used)
    // backward buffers (allocated only for training)
    float* gact[NET_LAYERS - 1];
    float* gout = nullptr;
    float* gtmp = nullptr;
    float* part = nullptr;
    double* dpart = nullptr;
    double mask_count = 0.0;
    bool trainable = false;
    // device-resident training store of the Galerkin energy loss
    // (ls_net_store_alloc): inputs, restricted residuals P^T r, e.A e per
    // sample; coarse Galerkin operators P^T A P per operator slot
    int st_cap = 0, st_nops = 0;
    __bf16* st_x = nullptr;      // [cap][n][16]
    float* st_rc = nullptr;      // [cap][n]
    std::vector<double> st_eae;  // [cap]
    std::vector<int> st_op;      // operator slot of each sample
    float* st_ac = nullptr;      // [nops][4][n]: diagonal, couplings to +i, +j, +k
    float* st_tmp = nullptr;     // [n][16] FP32 upload staging
    float* st_au = nullptr;      // [batch][n] A_c u
    double* st_part = nullptr;   // [batch][2][EN_BLOCKS] dot partials
    double* st_coef = nullptr;   // [batch][4] loss and gradient coefficients
    std::vector<void*> allocations;
    size_t bytes = 0;      // device bytes of the allocations
};

namespace ls {
size_t unet_part_floats(const NetLvl& L0, int batch);
int unet_forward(UNet* net, int batch);
int unet_backward(UNet* net, int batch);
int unet_convert_weights(UNet* net);
int unet_mse(UNet* net, const float* target, const float* mask, int batch, double* loss);
int unet_adam(UNet* net, float lr);
int unet_load_input(UNet* net, const float* in_dev, int batch);
// k_net_api.hip
int unet_alloc(UNet* net, int ni, int nj, int nk, int batch, bool trainable, hipStream_t stream);
void unet_free(UNet* net);
}  // namespace ls
