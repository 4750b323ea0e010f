// The learned coarse-grid correction's network driver (see coarse_net.hpp): packing a plan grid's
// forcing into the U-Net's input, the U-Net's forward on the library's convolution kernels, and the
// correction unpacked back onto the grid. Not a reference component: the reference has no learned part.
#include "coarse_net.hpp"

#include <cmath>
#include <cstring>

#include "cgocean.h"

namespace cg {

namespace {

__device__ inline uint16_t bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Per instance (blockIdx.x): the root mean square of its forcing over the grid's cells, summed in
// binary64 in a fixed order (each of 1024 threads four cells apart a step, then a tree), 1 where the
// forcing is zero.
__global__ __launch_bounds__(1024) void k_grid_rms(const float *__restrict__ f, long vstride, long cells,
                                                   float *__restrict__ scale) {
    __shared__ double sh[1024];
    const float *g = f + blockIdx.x * vstride;
    double s = 0.0;
    const long quads = cells / 4;
    for (long i = threadIdx.x; i < quads; i += 1024) {
        const float4 v = reinterpret_cast<const float4 *>(g)[i];
        s += (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w;
    }
    for (long i = 4 * quads + threadIdx.x; i < cells; i += 1024) s += (double)g[i] * g[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = sqrt(sh[0] / (double)cells);
        scale[blockIdx.x] = r > 0.0 ? (float)r : 1.0f;
    }
}

// The input of pixel (instance, cell): mask and scaled depth, the forcing over its rms, zeros.
__global__ void k_pack_grid(const uint16_t *__restrict__ fixed, const float *__restrict__ f, long vstride, long cells,
                            int n, const float *__restrict__ scale, uint16_t *__restrict__ x) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cells * n) return;
    const long inst = p / cells, cell = p - inst * cells;
    const uint32_t fx = *(const uint32_t *)(fixed + cell * 2);
    const float v = f[inst * vstride + cell] / scale[inst];
    *(uint4 *)(x + p * 16) = (uint4){fx, (uint32_t)bf16_round(v), 0u, 0u};
    *(uint4 *)(x + p * 16 + 8) = (uint4){0u, 0u, 0u, 0u};
}

// The correction of (instance, cell): channel 0 of the output times the forcing's rms and the
// network's output scale on water, zero on land.
__global__ void k_unpack_grid(const float *__restrict__ y32, const uint8_t *__restrict__ water, long cells, int n,
                              const float *__restrict__ scale, float out_scale, float *__restrict__ x, long vstride) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cells * n) return;
    const long inst = p / cells, cell = p - inst * cells;
    x[inst * vstride + cell] = water[cell] ? y32[p * 16] * (scale[inst] * out_scale) : 0.0f;
}

inline unsigned nblocks(long count, int per) { return (unsigned)((count + per - 1) / per); }

// Decoders of fewer pixels than this read an upsampled copy: the staged kernel's 4 x 64-pixel
// workgroups leave a small image's launch a handful of them.
constexpr long kUpStagedPixels = 65536;

template <typename T>
int alloc(CoarseNet &net, T **p, size_t count) {
    *p = nullptr;
    if (hipMalloc((void **)p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
        *p = nullptr;
        return CG_ERR_NOMEM;
    }
    net.allocs.push_back(*p);
    return CG_OK;
}

}  // namespace

void net_layers(int base, NetLayer out[kNetLayers]) {
    const int b1 = base, b2 = 2 * base, b4 = 4 * base, b8 = 8 * base;
    const NetLayer t[kNetLayers] = {
        {16, 0, b1, 1, 0}, {b1, 0, b1, 1, 0},   // enc0a, enc0b
        {b1, 0, b2, 1, 1}, {b2, 0, b2, 1, 1},   // enc1a, enc1b
        {b2, 0, b4, 1, 2}, {b4, 0, b4, 1, 2},   // enc2a, enc2b
        {b4, 0, b8, 1, 3}, {b8, 0, b8, 1, 3},   // mid_a, mid_b
        {b8, b4, b4, 1, 2}, {b4, 0, b4, 1, 2},  // dec2a, dec2b
        {b4, b2, b2, 1, 1}, {b2, 0, b2, 1, 1},  // dec1a, dec1b
        {b2, b1, b1, 1, 0}, {b1, 0, b1, 1, 0},  // dec0a, dec0b
        {b1, 0, 16, 0, 0},                      // out
    };
    for (int i = 0; i < kNetLayers; ++i) out[i] = t[i];
}

size_t net_weight_count(int base) {
    NetLayer L[kNetLayers];
    net_layers(base, L);
    size_t n = 0;
    for (int i = 0; i < kNetLayers; ++i) n += (size_t)9 * L[i].cout * (L[i].c1 + L[i].c2);
    return n;
}

bool coarse_net_shape_ok(int down, int across) { return down > 0 && across > 0 && down % 8 == 0 && across % 128 == 0; }

void coarse_net_free(CoarseNet &net) {
    for (void *p : net.allocs) (void)hipFree(p);
    net = CoarseNet{};
}

int coarse_net_setup(CoarseNet &net, int level, int base, int n, int down, int across, long vstride,
                     const uint16_t *weights, size_t count, const float *depth, const uint8_t *water,
                     float out_scale) {
    coarse_net_free(net);
    if (base < 16 || base % 16 || n < 1 || !coarse_net_shape_ok(down, across) || !weights || !water ||
        count != net_weight_count(base))
        return CG_ERR_ARG;
    net.level = level;
    net.base = base;
    net.n = n;
    net.down = down;
    net.across = across;
    net.vstride = vstride;
    net.out_scale = out_scale;
    net.water = water;
    net_layers(base, net.layers);
    int rc;
    auto bail = [&](int code) {
        coarse_net_free(net);
        return code;
    };
    size_t off = 0;
    for (int i = 0; i < kNetLayers; ++i) {
        net.w_off[i] = off;
        off += (size_t)9 * net.layers[i].cout * (net.layers[i].c1 + net.layers[i].c2);
    }
    if ((rc = alloc(net, &net.w, off))) return bail(rc);
    if (hipMemcpy(net.w, weights, sizeof(uint16_t) * off, hipMemcpyHostToDevice) != hipSuccess)
        return bail(CG_ERR_DEVICE);
    const long cells = (long)down * across;
    // The constant channels: the mask is the grid's, read back from the device, the depth the caller's.
    std::vector<uint8_t> w8(cells);
    if (hipMemcpy(w8.data(), water, cells, hipMemcpyDeviceToHost) != hipSuccess) return bail(CG_ERR_DEVICE);
    std::vector<uint16_t> fx((size_t)cells * 2);
    for (long c = 0; c < cells; ++c) {
        fx[2 * c] = host_bf16(w8[c] ? 1.0f : 0.0f);
        fx[2 * c + 1] = host_bf16(depth ? depth[c] / 6000.0f : 0.0f);
    }
    if ((rc = alloc(net, &net.fixed, fx.size()))) return bail(rc);
    if (hipMemcpy(net.fixed, fx.data(), sizeof(uint16_t) * fx.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(CG_ERR_DEVICE);
    const long px = cells * n;
    if ((rc = alloc(net, &net.x, (size_t)px * 16))) return bail(rc);
    for (int i = 0; i < kNetLayers; ++i) {
        const NetLayer &L = net.layers[i];
        const long lpx = px >> (2 * L.level);
        if (i == kNetLayers - 1) {
            if ((rc = alloc(net, &net.y32, (size_t)lpx * L.cout))) return bail(rc);
        } else if ((rc = alloc(net, &net.act[i], (size_t)lpx * L.cout))) {
            return bail(rc);
        }
    }
    for (int k = 0; k < 3; ++k) {
        const long lpx = px >> (2 * (k + 1));
        const int ch = base << k;  // the pooled encoder output of level k
        if ((rc = alloc(net, &net.pooled[k], (size_t)lpx * ch))) return bail(rc);
        // a decoder of level k reads the level k + 1 output upsampled: (8, 4, 2) x base channels
        if ((rc = alloc(net, &net.upsampled[k], (size_t)(px >> (2 * k)) * (base << (k + 1))))) return bail(rc);
    }
    if ((rc = alloc(net, &net.scale, n))) return bail(rc);
    return CG_OK;
}

int coarse_net_forward(const CoarseNet &net, const float *f, float *x, hipStream_t s) {
    const long cells = (long)net.down * net.across, px = cells * net.n;
    const int n = net.n;
    hipLaunchKernelGGL(k_grid_rms, dim3(n), dim3(1024), 0, s, f, net.vstride, cells, net.scale);
    hipLaunchKernelGGL(k_pack_grid, dim3(nblocks(px, 256)), dim3(256), 0, s, net.fixed, f, net.vstride, cells, n,
                       net.scale, net.x);
    if (hipGetLastError() != hipSuccess) return CG_ERR_DEVICE;
    auto dims = [&](int lvl, int *d, int *a) {
        *d = net.down >> lvl;
        *a = net.across >> lvl;
    };
    auto conv = [&](int i, const uint16_t *a1, const uint16_t *a2) -> int {
        const NetLayer &L = net.layers[i];
        int d, a;
        dims(L.level, &d, &a);
        const uint16_t *w = net.w + net.w_off[i];
        if (i == kNetLayers - 1)
            return cg_conv3x3(a1, L.c1, a2, L.c2, n, d, a, w, L.cout, L.relu, nullptr, 0, nullptr, net.y32, s);
        return cg_conv3x3(a1, L.c1, a2, L.c2, n, d, a, w, L.cout, L.relu, net.act[i], L.cout, nullptr, nullptr, s);
    };
    auto pool = [&](int lvl, const uint16_t *in, int ch, uint16_t *out) -> int {
        int d, a;
        dims(lvl, &d, &a);
        return cg_pool2(in, ch, n, d, a, 1.0f, out, s);
    };
    // A decoder of level lvl on (its coarse input upsampled, the skip): read through the upsampling by
    // the staged kernel where it serves the shape and the image is large (UNet.forward's rule), else
    // upsampled into a tensor first.
    auto decoder = [&](int i, const uint16_t *coarse, const uint16_t *skip, int lvl) -> int {
        const NetLayer &L = net.layers[i];
        int d, a;
        dims(lvl, &d, &a);
        const uint16_t *w = net.w + net.w_off[i];
        if (L.c1 % 32 == 0 && L.c2 % 32 == 0 && a % 64 == 0 && d % 4 == 0 && (long)n * d * a >= kUpStagedPixels)
            return cg_conv3x3_up(coarse, L.c1, skip, L.c2, n, d, a, w, L.cout, L.relu, net.act[i], L.cout, nullptr,
                                 nullptr, s);
        int rc = cg_upsample2(coarse, L.c1, n, d / 2, a / 2, 1.0f, net.upsampled[lvl], s);
        if (rc) return rc;
        return cg_conv3x3(net.upsampled[lvl], L.c1, skip, L.c2, n, d, a, w, L.cout, L.relu, net.act[i], L.cout,
                          nullptr, nullptr, s);
    };
    int rc;
    const NetLayer &e0 = net.layers[0];
    int d0, a0;
    dims(0, &d0, &a0);
    // The input's channels 3..15 are zero by construction: the first layer reads its three data channels.
    if (e0.cout <= 32 && e0.cout % 8 == 0)
        rc = cg_conv3x3_in3(net.x, n, d0, a0, net.w, e0.cout, e0.relu, net.act[0], s);
    else
        rc = conv(0, net.x, nullptr);
    if (rc) return rc;
    const int b = net.base;
    if ((rc = conv(1, net.act[0], nullptr))) return rc;                       // enc0b
    if ((rc = pool(0, net.act[1], b, net.pooled[0]))) return rc;             // p1
    if ((rc = conv(2, net.pooled[0], nullptr))) return rc;                    // enc1a
    if ((rc = conv(3, net.act[2], nullptr))) return rc;                       // enc1b
    if ((rc = pool(1, net.act[3], 2 * b, net.pooled[1]))) return rc;         // p2
    if ((rc = conv(4, net.pooled[1], nullptr))) return rc;                    // enc2a
    if ((rc = conv(5, net.act[4], nullptr))) return rc;                       // enc2b
    if ((rc = pool(2, net.act[5], 4 * b, net.pooled[2]))) return rc;         // p3
    if ((rc = conv(6, net.pooled[2], nullptr))) return rc;                    // mid_a
    if ((rc = conv(7, net.act[6], nullptr))) return rc;                       // mid_b
    if ((rc = decoder(8, net.act[7], net.act[5], 2))) return rc;              // dec2a
    if ((rc = conv(9, net.act[8], nullptr))) return rc;                       // dec2b
    if ((rc = decoder(10, net.act[9], net.act[3], 1))) return rc;             // dec1a
    if ((rc = conv(11, net.act[10], nullptr))) return rc;                     // dec1b
    if ((rc = decoder(12, net.act[11], net.act[1], 0))) return rc;            // dec0a
    if ((rc = conv(13, net.act[12], nullptr))) return rc;                     // dec0b
    if ((rc = conv(14, net.act[13], nullptr))) return rc;                     // out (binary32)
    hipLaunchKernelGGL(k_unpack_grid, dim3(nblocks(px, 256)), dim3(256), 0, s, net.y32, net.water, cells, n, net.scale,
                       net.out_scale, x, net.vstride);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

}  // namespace cg
