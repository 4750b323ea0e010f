// The learned coarse-grid correction's network driver: the U-Net of 3x3 convolutions run natively on
// one grid of a solve plan, from that grid's binary32 forcing to its binary32 correction, every launch
// only enqueued on a stream (so a refinement step's graph captures it). The layers and their order are
// those of cg_unet_layer (include/cgocean.h); the convolutions, pools and upsamplings are the library's
// bf16 matrix-core kernels (unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace cg {

constexpr int kNetLayers = 15;

struct NetLayer {
    int c1, c2, cout, relu, level;  // input channels (first, skip), outputs, ReLU, U-Net level 0..3
};

// The layer table for `base` channels at the finest U-Net level (base, 2 base, 4 base, 8 base).
void net_layers(int base, NetLayer out[kNetLayers]);
// bf16 weights of all layers, [9][cout][c1 + c2] each, in table order.
size_t net_weight_count(int base);

struct CoarseNet {
    int level = -1;  // the plan grid it serves; -1: none
    int base = 0;
    int n = 0, down = 0, across = 0;  // instances and the grid's shape
    long vstride = 0;                 // floats between two instances' grids
    float out_scale = 1.0f;           // the correction is out_scale x rms(f) x the network's channel 0
    NetLayer layers[kNetLayers];
    uint16_t *w = nullptr;  // every layer's weights, table order
    size_t w_off[kNetLayers] = {0};
    uint16_t *fixed = nullptr;  // per cell: mask, depth / 6000 m (bf16)
    const uint8_t *water = nullptr;  // the grid's mask (the stencil's, 0 or 1)
    uint16_t *x = nullptr;           // the packed input, 16 channels
    uint16_t *act[kNetLayers] = {nullptr};
    uint16_t *pooled[3] = {nullptr};
    uint16_t *upsampled[3] = {nullptr};  // decoders the staged kernel cannot read through upsampling
    float *y32 = nullptr;                // the output layer, 16 channels binary32
    float *scale = nullptr;              // per instance: rms of the forcing over the grid
    std::vector<void *> allocs;
};

// Shapes the network serves: down % 8 == 0, across % 128 == 0 (four U-Net levels of 16-column rows).
bool coarse_net_shape_ok(int down, int across);
// Allocates the network's buffers and uploads its weights (host, bf16) and depth (host, metres per
// cell, or NULL for zeros). Returns 0 or a negative status (CG_ERR_*); frees what it allocated on
// failure.
int coarse_net_setup(CoarseNet &net, int level, int base, int n, int down, int across, long vstride,
                     const uint16_t *weights, size_t count, const float *depth, const uint8_t *water,
                     float out_scale);
void coarse_net_free(CoarseNet &net);
// x (n instances, vstride apart) = the network's correction for forcing f (likewise), zero on land.
int coarse_net_forward(const CoarseNet &net, const float *f, float *x, hipStream_t s);

}  // namespace cg
