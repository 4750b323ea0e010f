// unet.hip: the learned preconditioner's kernels. A 3x3 convolution on bfloat16 activations with
// binary32 accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16), its adjoint (the data
// gradient) and its weight gradient, plus the few elementwise steps a U-Net needs (average pool,
// nearest upsampling, ReLU gradient mask, sum) and the packing between a solver vector and the
// network's input and output. Not a reference component: the reference has no learned part.
//
// Layout: activations are NHWC bfloat16, act[((b * down + row) * across + col) * channels + c];
// weights are [9][cout][cin] bfloat16 with tap = 3 * (dr + 1) + (dc + 1). The convolution pads rows
// with zeros and wraps columns (the solve grid is periodic east-west). A convolution may read its
// input from two tensors (channels [0, c1) from x1, [c1, c1 + c2) from x2): the U-Net's skip
// concatenation without a copy. Channel counts are multiples of 16 and `across` a multiple of 16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>

#include "cgocean.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t bf16_of(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                    // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ inline float float_of(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// Two bfloat16 values in one 32-bit word (the lower channel in the low half).
__device__ inline uint32_t pack2(float lo, float hi) { return (uint32_t)bf16_of(lo) | ((uint32_t)bf16_of(hi) << 16); }
__device__ inline float lo_of(uint32_t u) { return __uint_as_float(u << 16); }
__device__ inline float hi_of(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

__device__ inline int wrap(int c, int across) { return c < 0 ? c + across : (c >= across ? c - across : c); }

// y[p][o] = act(sum over taps t and input channels i of x[p + off(t)][i] * W(t, o, i)).
// Forward: W(t, o, i) = w[t][o][i]. ADJ (the data gradient of a forward convolution whose weights
// are w[9][cin][cout]): W(t, o, i) = w[8 - t][i][o], so y is the gradient with respect to that
// convolution's input when x is the gradient with respect to its output.
// One wavefront computes 16 consecutive pixels of one row (the MFMA's M) times 16 * NB output
// channels (N), summing over K = 9 taps x cin channels in steps of 16. Lane l feeds A[l % 16][4 (l / 16) + j]
// (pixel l % 16, channels 4 (l / 16) .. + 3: 8 contiguous bytes of NHWC) and B[4 (l / 16) + j][l % 16],
// and holds D[4 (l / 16) + j][l % 16] afterwards.
// Output channels [0, o1) go to y1 (o1 channels per pixel), [o1, cout) to y2; with y32 set, all
// cout channels go there as binary32 instead.
template <int NB, bool ADJ>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ x1, int c1, const uint16_t *__restrict__ x2,
                                                 int c2, int n, int down, int across, const uint16_t *__restrict__ w,
                                                 int cout, int relu, uint16_t *__restrict__ y1, int o1,
                                                 uint16_t *__restrict__ y2, float *__restrict__ y32) {
    const int lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int per_row = across >> 4;
    if (tile >= (long)n * down * per_row) return;  // the whole wavefront leaves together
    const long row_id = tile / per_row;             // b * down + row
    const int row = (int)(row_id % down);
    const long img = row_id / down;
    const int c0 = (int)(tile % per_row) * 16;
    const int co0 = blockIdx.y * 16 * NB;
    const int cin = c1 + c2;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    f4 acc[NB];
    for (int b = 0; b < NB; ++b) acc[b] = (f4){0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int rr = row + tap / 3 - 1;
        if (rr < 0 || rr >= down) continue;  // zero rows contribute nothing; uniform per wavefront
        const long pix = (img * down + rr) * across + wrap(c0 + m + tap % 3 - 1, across);
        for (int ci0 = 0; ci0 < cin; ci0 += 16) {
            const s4 a = ci0 < c1 ? *(const s4 *)(x1 + pix * c1 + ci0 + kq)
                                  : *(const s4 *)(x2 + pix * c2 + (ci0 - c1) + kq);
            for (int b = 0; b < NB; ++b) {
                const int co = co0 + b * 16 + m;
                s4 bv;
                if (!ADJ) {
                    bv = *(const s4 *)(w + ((long)tap * cout + co) * cin + ci0 + kq);
                } else {
                    const uint16_t *src = w + ((long)(8 - tap) * cin + ci0 + kq) * cout + co;
                    bv = (s4){(short)src[0], (short)src[cout], (short)src[2 * cout], (short)src[3 * cout]};
                }
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[b], 0, 0, 0);
            }
        }
    }
    const long p0 = row_id * across + c0 + kq;
    for (int b = 0; b < NB; ++b) {
        const int co = co0 + b * 16 + m;
        for (int j = 0; j < 4; ++j) {
            float v = acc[b][j];
            if (relu) v = v > 0.f ? v : 0.f;
            const long p = p0 + j;
            if (y32)
                y32[p * cout + co] = v;
            else if (co < o1)
                y1[p * o1 + co] = bf16_of(v);
            else
                y2[p * (cout - o1) + (co - o1)] = bf16_of(v);
        }
    }
}

// The forward convolution when every input channel count is a multiple of 32: gfx950's
// v_mfma_f32_16x16x32_bf16 (K = 32 per instruction, 16-byte operand loads) and MT tiles of 16
// pixels per wavefront (16 MT consecutive pixels of one row), so each B operand (weights) serves MT
// instructions and each A operand (pixels) NB. Lane l feeds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16], j = 0..7, and holds D[4 (l / 16) + j][l % 16] as in k_conv3x3.
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

// The weights of one 32-channel chunk (9 taps x 16 NB output channels) are staged in LDS by the
// whole workgroup and read by its four wavefronts; rows padded to 40 elements against bank conflicts.
constexpr int kSbStride = 40;

template <int MT, int NB>
__global__ __launch_bounds__(256) void k_conv3x3_k32(const uint16_t *__restrict__ x1, int c1,
                                                     const uint16_t *__restrict__ x2, int c2, int n, int down,
                                                     int across, const uint16_t *__restrict__ w, int cout, int relu,
                                                     uint16_t *__restrict__ y1, int o1, uint16_t *__restrict__ y2,
                                                     float *__restrict__ y32) {
    __shared__ __attribute__((aligned(16))) uint16_t sb[9 * 16 * NB * kSbStride];
    const int lane = threadIdx.x & 63;
    const long tile0 = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int per_row = across / (16 * MT);
    const long tiles = (long)n * down * per_row;
    // A wavefront past the last tile still stages weights and meets every barrier; it computes the
    // last tile again and stores nothing.
    const bool valid = tile0 < tiles;
    const long tile = valid ? tile0 : tiles - 1;
    const long row_id = tile / per_row;
    const int row = (int)(row_id % down);
    const long img = row_id / down;
    const int c0 = (int)(tile % per_row) * 16 * MT;
    const int co0 = blockIdx.y * 16 * NB;
    const int cin = c1 + c2;
    const int m = lane & 15, k8 = (lane >> 4) * 8, kq = (lane >> 4) * 4;
    f4 acc[MT][NB];
    for (int t = 0; t < MT; ++t)
        for (int b = 0; b < NB; ++b) acc[t][b] = (f4){0.f, 0.f, 0.f, 0.f};
    for (int ci0 = 0; ci0 < cin; ci0 += 32) {
        __syncthreads();  // every wavefront is done with the previous chunk's weights
        for (int e = threadIdx.x; e < 9 * 16 * NB * 4; e += 256) {
            const int tap = e / (16 * NB * 4), r = e % (16 * NB * 4), co_l = r >> 2, q = r & 3;
            *(uint4 *)&sb[(tap * 16 * NB + co_l) * kSbStride + q * 8] =
                *(const uint4 *)(w + ((long)tap * cout + co0 + co_l) * cin + ci0 + q * 8);
        }
        __syncthreads();
        const bool first = ci0 < c1;
        const uint16_t *xs = first ? x1 + ci0 + k8 : x2 + (ci0 - c1) + k8;
        const int cs = first ? c1 : c2;
        for (int tap = 0; tap < 9; ++tap) {
            const int rr = row + tap / 3 - 1;
            if (rr < 0 || rr >= down) continue;  // no barrier inside this loop
            const long rbase = (img * down + rr) * across;
            b8 a[MT], bv[NB];
            for (int t = 0; t < MT; ++t)
                a[t] = *(const b8 *)(xs + (rbase + wrap(c0 + 16 * t + m + tap % 3 - 1, across)) * cs);
            for (int b = 0; b < NB; ++b)
                bv[b] = *(const b8 *)&sb[(tap * 16 * NB + b * 16 + m) * kSbStride + k8];
            for (int t = 0; t < MT; ++t)
                for (int b = 0; b < NB; ++b)
                    acc[t][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], bv[b], acc[t][b], 0, 0, 0);
        }
    }
    if (!valid) return;
    for (int t = 0; t < MT; ++t) {
        const long p0 = row_id * across + c0 + 16 * t + kq;
        for (int b = 0; b < NB; ++b) {
            const int co = co0 + b * 16 + m;
            for (int j = 0; j < 4; ++j) {
                float v = acc[t][b][j];
                if (relu) v = v > 0.f ? v : 0.f;
                const long p = p0 + j;
                if (y32)
                    y32[p * cout + co] = v;
                else if (co < o1)
                    y1[p * o1 + co] = bf16_of(v);
                else
                    y2[p * (cout - o1) + (co - o1)] = bf16_of(v);
            }
        }
    }
}

// The forward convolution with both operands staged in LDS: a workgroup computes tiles of 4 rows x
// 64 pixels (one row per wavefront, four 16-pixel MFMA tiles each) times 16 NB output channels. Per
// 32-channel chunk it stages the 6 input rows a tile reads (66 pixels each: the tile and one halo
// column either side, wrapped; rows beyond the grid as zeros) and the chunk's weights for all nine
// taps, so each input value is fetched from L2 once per chunk rather than once per tap and output
// row, and the nine taps read shifted windows of LDS. Needs across % 64 == 0 and down % 4 == 0.
// With up1 the first input is read through nearest upsampling from a tensor of half the height and
// width (the U-Net's decoder input), so the upsampled tensor is never written.
//
// The grid holds as many workgroups as stay resident; workgroup k takes tiles k, k + grid, ... and
// walks their (tile, chunk) items in order, loading the next item's operands into registers while
// the current item's MFMAs and output stores run, so every workgroup keeps one item's loads in flight
// from its first item to its last. With one chunk (32 input channels) the weights are staged once.
//
// LDS holds 16-byte slots of 8 channels, four a pixel (row-major over the 6 x 66 staged pixels)
// or output channel (tap-major), the channel group q of item p in slot 4 p + (q ^ ((p >> 1) & 3)).
// Both a fragment read (lane m + 16 g takes item base + m's group g) and a staging store (lane l
// writes item l / 4's group l % 4) then touch every bank once for each pair of 8-lane blocks that
// gfx950's LDS serves together (measured: with 80-byte padded pixel rows both conflicted two-way).
constexpr int kApx = 66;

__device__ inline int swz(int p, int q) { return p * 4 + (q ^ ((p >> 1) & 3)); }

// 16 bytes as a plain vector (HIP's uint4 is a struct, whose copies can keep an array in scratch).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int NB>
__global__ __launch_bounds__(256, 2) void k_conv3x3_tile(const uint16_t *__restrict__ x1, int c1,
                                                      const uint16_t *__restrict__ x2, int c2, int n, int down,
                                                      int across, const uint16_t *__restrict__ w, int cout, int relu,
                                                      uint16_t *__restrict__ y1, int o1, uint16_t *__restrict__ y2,
                                                      float *__restrict__ y32, int up1) {
    __shared__ __attribute__((aligned(16))) uint16_t sa[6 * kApx * 4 * 8];
    __shared__ __attribute__((aligned(16))) uint16_t sb[9 * 16 * NB * 4 * 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col_tiles = across / 64, row_blocks = down / 4;
    const long tiles = (long)n * row_blocks * col_tiles;
    const int co0 = blockIdx.y * 16 * NB;
    const int cin = c1 + c2;
    const int m = lane & 15, g = lane >> 4, kq = (lane >> 4) * 4;
    long tile = blockIdx.x;
    if (tile >= tiles) return;  // the whole workgroup leaves together
    f4 acc[4][NB];
    #pragma unroll
    for (int t = 0; t < 4; ++t)
        #pragma unroll
        for (int b = 0; b < NB; ++b) acc[t][b] = (f4){0.f, 0.f, 0.f, 0.f};
    // Staging units of 16 bytes: pixel unit u is channel group u & 3 of staged pixel u >> 2 (row-major
    // over the 6 x 66 staged pixels), weight unit e channel group e & 3 of output channel (e >> 2) %
    // (16 NB) of tap e / (64 NB). Units are loaded into registers (kAU and kBU a thread; the weights
    // only with at most 32 output channels a workgroup: with 64 their registers cost more occupancy
    // than the overlap gains, so they are loaded as they are stored) and go to LDS between two barriers.
    constexpr int kAUnits = 6 * kApx * 4, kBUnits = 9 * 16 * NB * 4;
    constexpr int kAU = (kAUnits + 255) / 256, kBU = (kBUnits + 255) / 256;
    constexpr bool kPrefetchB = NB <= 2;
    u32x4 ra[kAU], rw[kPrefetchB ? kBU : 1];
    auto load_a = [&](long tl, int ci0) {
        const int c0 = (int)(tl % col_tiles) * 64;
        const long rb = tl / col_tiles;
        const int r0 = (int)(rb % row_blocks) * 4;
        const long img = rb / row_blocks;
        const bool first = ci0 < c1;
        const uint16_t *xs = first ? x1 + ci0 : x2 + (ci0 - c1);
        const int cs = first ? c1 : c2;
        #pragma unroll
        for (int k = 0; k < kAU; ++k) {
            const int u = threadIdx.x + 256 * k;
            ra[k] = (u32x4){0u, 0u, 0u, 0u};
            if (u < kAUnits) {
                const int i = u / (kApx * 4), px = (u >> 2) % kApx, q = u & 3;
                const int rr = r0 - 1 + i;
                if (rr >= 0 && rr < down) {
                    const int cc = wrap(c0 - 1 + px, across);
                    const long pix = first && up1 ? (img * (down >> 1) + (rr >> 1)) * (across >> 1) + (cc >> 1)
                                                  : (img * down + rr) * across + cc;
                    ra[k] = *(const u32x4 *)(xs + pix * cs + q * 8);
                }
            }
        }
    };
    auto weight_unit = [&](int ci0, int e) -> const u32x4 * {
        const int tap = e / (16 * NB * 4), r = e % (16 * NB * 4), co_l = r >> 2, q = r & 3;
        return (const u32x4 *)(w + ((long)tap * cout + co0 + co_l) * cin + ci0 + q * 8);
    };
    auto weight_slot = [&](int e) -> u32x4 * {
        const int tap = e / (16 * NB * 4), r = e % (16 * NB * 4), co_l = r >> 2, q = r & 3;
        return (u32x4 *)&sb[swz(tap * 16 * NB + co_l, q) * 8];
    };
    auto load_b = [&](int ci0) {
        if constexpr (kPrefetchB)
            #pragma unroll
            for (int k = 0; k < kBU; ++k) {
                const int e = threadIdx.x + 256 * k;
                if (e < kBUnits) rw[k] = *weight_unit(ci0, e);
            }
    };
    const bool one_chunk = cin == 32;
    bool weights_staged = false;
    int ci0 = 0;
    load_a(tile, 0);
    load_b(0);
    for (;;) {
        __syncthreads();  // every wavefront is done with the previous item
        #pragma unroll
        for (int k = 0; k < kAU; ++k) {
            const int u = threadIdx.x + 256 * k;
            if (u < kAUnits) {
                *(u32x4 *)&sa[swz(u >> 2, u & 3) * 8] = ra[k];
            }
        }
        if (!weights_staged) {
            if constexpr (kPrefetchB) {
                #pragma unroll
                for (int k = 0; k < kBU; ++k) {
                    const int e = threadIdx.x + 256 * k;
                    if (e < kBUnits) *weight_slot(e) = rw[k];
                }
            } else {
                for (int e = threadIdx.x; e < kBUnits; e += 256) *weight_slot(e) = *weight_unit(ci0, e);
            }
            weights_staged = one_chunk;
        }
        __syncthreads();
        // The next item: the next chunk of this tile, else this workgroup's next tile from chunk 0.
        const bool last_chunk = ci0 + 32 == cin;
        const long next_tile = last_chunk ? tile + gridDim.x : tile;
        const int next_ci = last_chunk ? 0 : ci0 + 32;
        const bool more = next_tile < tiles;
        if (more) {
            load_a(next_tile, next_ci);
            if (!one_chunk) load_b(next_ci);
        }
        #pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int i = wave + tap / 3;  // staged row of input row r0 + wave + dr, dr = tap / 3 - 1
            const int dc = tap % 3;        // staged pixel offset: px = 16 t + m + 1 + (tap % 3 - 1)
            b8 a[4], bv[NB];
            #pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = *(const b8 *)&sa[swz(i * kApx + 16 * t + m + dc, g) * 8];
            #pragma unroll
            for (int b = 0; b < NB; ++b) bv[b] = *(const b8 *)&sb[swz(tap * 16 * NB + b * 16 + m, g) * 8];
            #pragma unroll
            for (int t = 0; t < 4; ++t)
                #pragma unroll
                for (int b = 0; b < NB; ++b)
                    acc[t][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], bv[b], acc[t][b], 0, 0, 0);
        }
        if (last_chunk) {
            const int c0 = (int)(tile % col_tiles) * 64;
            const long rb = tile / col_tiles;
            const long prow = ((rb / row_blocks) * down + (rb % row_blocks) * 4 + wave) * across + c0;
            #pragma unroll
            for (int t = 0; t < 4; ++t) {
                #pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int co = co0 + b * 16 + m;
                    #pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float v = acc[t][b][j];
                        if (relu) v = v > 0.f ? v : 0.f;
                        const long p = prow + 16 * t + kq + j;
                        if (y32)
                            y32[p * cout + co] = v;
                        else if (co < o1)
                            y1[p * o1 + co] = bf16_of(v);
                        else
                            y2[p * (cout - o1) + (co - o1)] = bf16_of(v);
                    }
                    acc[t][b] = (f4){0.f, 0.f, 0.f, 0.f};
                }
            }
        }
        if (!more) break;
        tile = next_tile;
        ci0 = next_ci;
    }
}

// Weight gradient partials: part[s][t][o][i] = sum over the pixels p of slice s of
// dy[p][o] * x[p + off(t)][i]. One wavefront per (16 output channels, 16 input channels, slice),
// all nine taps at once (A = dy^T is loaded once per 16 pixels and used for every tap). M = output
// channels, N = input channels, K = 16 pixels of one row (across is a multiple of 16).
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const uint16_t *__restrict__ x1, int c1,
                                                       const uint16_t *__restrict__ x2, int c2,
                                                       const uint16_t *__restrict__ dy, int cout, int n, int down,
                                                       int across, int slices, float *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int slice = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (slice >= slices) return;
    const int cin = c1 + c2;
    const int cit = cin >> 4;
    const int co0 = (blockIdx.x / cit) * 16, ci0 = (blockIdx.x % cit) * 16;
    const bool first = ci0 < c1;
    const uint16_t *xs = first ? x1 : x2;
    const int cs = first ? c1 : c2, cofs = first ? ci0 : ci0 - c1;
    const long groups = (long)n * down * across / 16;
    const long g_lo = groups * slice / slices, g_hi = groups * (slice + 1) / slices;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    f4 acc[9];
    for (int t = 0; t < 9; ++t) acc[t] = (f4){0.f, 0.f, 0.f, 0.f};
    for (long g = g_lo; g < g_hi; ++g) {
        const long p0 = g * 16;
        const long row_id = p0 / across;
        const int c0 = (int)(p0 - row_id * across);
        const int row = (int)(row_id % down);
        const long img = row_id / down;
        const uint16_t *d = dy + (p0 + kq) * cout + co0 + m;
        const s4 a = (s4){(short)d[0], (short)d[cout], (short)d[2 * cout], (short)d[3 * cout]};
        for (int t = 0; t < 9; ++t) {
            const int rr = row + t / 3 - 1;
            if (rr < 0 || rr >= down) continue;
            const long base = (img * down + rr) * across;
            short v[4];
            for (int j = 0; j < 4; ++j)
                v[j] = (short)xs[(base + wrap(c0 + kq + j + t % 3 - 1, across)) * cs + cofs + m];
            const s4 bv = (s4){v[0], v[1], v[2], v[3]};
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[t], 0, 0, 0);
        }
    }
    float *dst = part + (long)slice * 9 * cout * cin;
    for (int t = 0; t < 9; ++t)
        for (int j = 0; j < 4; ++j) dst[((long)t * cout + co0 + kq + j) * cin + ci0 + m] = acc[t][j];
}

// The weight gradient partials as k_conv3x3_wgrad computes them, one wavefront (one workgroup) per
// (16 output channels, 16 input channels, slice), on v_mfma_f32_16x16x32_bf16 over groups of 32
// pixels of one row (across % 32 == 0). Per group the wavefront stages, with 8-byte loads, the
// output gradient's 32 x 16 block and the input's 3 rows x 34 pixels (a halo column either side,
// wrapped; rows beyond the grid as zeros) x 16 channels in LDS, and forms the pixel-major operands
// there (lane l: channel l % 16, pixels 8 (l / 16) .. + 7). Pixel rows are 20 elements apart so the
// four lane groups fall in different banks.
constexpr int kWgStride = 20;
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(64) void k_conv3x3_wgrad32(const uint16_t *__restrict__ x1, int c1,
                                                        const uint16_t *__restrict__ x2, int c2,
                                                        const uint16_t *__restrict__ dy, int cout, int n, int down,
                                                        int across, int slices, float *__restrict__ part) {
    __shared__ __attribute__((aligned(16))) uint16_t sdy[32 * kWgStride];
    __shared__ __attribute__((aligned(16))) uint16_t sx[3 * 34 * kWgStride];
    const int lane = threadIdx.x;
    const int slice = blockIdx.y;
    const int cin = c1 + c2;
    const int cit = cin >> 4;
    const int co0 = (blockIdx.x / cit) * 16, ci0 = (blockIdx.x % cit) * 16;
    const bool first = ci0 < c1;
    const uint16_t *xs = first ? x1 : x2;
    const int cs = first ? c1 : c2, cofs = first ? ci0 : ci0 - c1;
    const long groups = (long)n * down * across / 32;
    const long g_lo = groups * slice / slices, g_hi = groups * (slice + 1) / slices;
    const int m = lane & 15, k8 = (lane >> 4) * 8;
    f4 acc[9];
    for (int t = 0; t < 9; ++t) acc[t] = (f4){0.f, 0.f, 0.f, 0.f};
    for (long g = g_lo; g < g_hi; ++g) {
        const long p0 = g * 32;
        const long row_id = p0 / across;
        const int c0 = (int)(p0 - row_id * across);
        const int row = (int)(row_id % down);
        const long img = row_id / down;
        __syncthreads();  // the previous group's operands have been read
        for (int u = lane; u < 32 * 4; u += 64) {  // 32 pixels x 4 pieces of 4 channels
            const int px = u >> 2, q = u & 3;
            *(uint2 *)&sdy[px * kWgStride + q * 4] = *(const uint2 *)(dy + (p0 + px) * cout + co0 + q * 4);
        }
        for (int u = lane; u < 3 * 34 * 4; u += 64) {
            const int r = u / (34 * 4), px = (u >> 2) % 34, q = u & 3;
            const int rr = row + r - 1;
            uint2 v = {0u, 0u};
            if (rr >= 0 && rr < down)
                v = *(const uint2 *)(xs + ((img * down + rr) * across + wrap(c0 - 1 + px, across)) * cs + cofs + q * 4);
            *(uint2 *)&sx[(r * 34 + px) * kWgStride + q * 4] = v;
        }
        __syncthreads();
        us8 av;
        for (int j = 0; j < 8; ++j) av[j] = sdy[(k8 + j) * kWgStride + m];
        const b8 a = __builtin_bit_cast(b8, av);
        for (int t = 0; t < 9; ++t) {
            const int r = t / 3, dc = t % 3;  // staged row of row + dr; staged pixel of pixel + dc - 1
            us8 bvv;
            for (int j = 0; j < 8; ++j) bvv[j] = sx[(r * 34 + k8 + j + dc) * kWgStride + m];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(b8, bvv), acc[t], 0, 0, 0);
        }
    }
    const int kq = (lane >> 4) * 4;
    float *dst = part + (long)slice * 9 * cout * cin;
    for (int t = 0; t < 9; ++t)
        for (int j = 0; j < 4; ++j) dst[((long)t * cout + co0 + kq + j) * cin + ci0 + m] = acc[t][j];
}

// The network's first convolution, whose 16-channel input holds data in channels 0..2 only (mask,
// depth, vector; the rest zero by construction): one thread per pixel, the 27 products per output
// channel in binary32 on the vector units (the matrix cores would multiply 13 zero channels), the
// weights w[9][cout][16] of channels 0..2 staged in LDS as binary32. cout <= 32.
constexpr int kIn3MaxOut = 32;
__global__ __launch_bounds__(256) void k_conv3x3_in3(const uint16_t *__restrict__ x, int n, int down, int across,
                                                     const uint16_t *__restrict__ w, int cout, int relu,
                                                     uint16_t *__restrict__ y) {
    __shared__ float sw[9 * 3 * kIn3MaxOut];  // [tap][channel][out]
    for (int i = threadIdx.x; i < 9 * 3 * cout; i += blockDim.x) {
        const int t = i / (3 * cout), c = (i / cout) % 3, o = i % cout;
        sw[(t * 3 + c) * kIn3MaxOut + o] = float_of(w[((long)t * cout + o) * 16 + c]);
    }
    __syncthreads();
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)n * down * across) return;
    const long row_id = p / across;
    const int col = (int)(p - row_id * across), row = (int)(row_id % down);
    const long img = row_id / down;
    float in[9][3];
    for (int t = 0; t < 9; ++t) {
        const int rr = row + t / 3 - 1;
        uint2 v = {0u, 0u};
        if (rr >= 0 && rr < down)
            v = *(const uint2 *)(x + ((img * down + rr) * across + wrap(col + t % 3 - 1, across)) * 16);
        in[t][0] = lo_of(v.x);
        in[t][1] = hi_of(v.x);
        in[t][2] = lo_of(v.y);
    }
    for (int o0 = 0; o0 < cout; o0 += 8) {
        float acc[8];
        for (int k = 0; k < 8; ++k) acc[k] = 0.f;
        for (int t = 0; t < 9; ++t)
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 8; ++k) acc[k] = fmaf(in[t][c], sw[(t * 3 + c) * kIn3MaxOut + o0 + k], acc[k]);
        if (relu)
            for (int k = 0; k < 8; ++k) acc[k] = acc[k] > 0.f ? acc[k] : 0.f;
        *(uint4 *)(y + p * cout + o0) =
            (uint4){pack2(acc[0], acc[1]), pack2(acc[2], acc[3]), pack2(acc[4], acc[5]), pack2(acc[6], acc[7])};
    }
}

// The first layer on the matrix cores, for output channel counts that are multiples of 16: one
// wavefront computes kIn3Groups groups of 16 consecutive pixels (of the flattened image, row, column
// order) times every output channel as D = W^T X on v_mfma_f32_16x16x32_bf16, with K = 4 tap + channel
// over taps 0..8 and channels 0..3 (channel 3 zeroed in both operands, so what the input holds there
// never enters), in two K steps of 32: taps 0..7, then tap 8 and zeros. Lane m + 16 g feeds each
// group's pixel m with taps 2g and 2g + 1 (step 0) and tap 8 + 2g (step 1, beyond tap 8 zero), one
// 8-byte load each, all groups' loads issued before the first product, and the weights of output
// channel m of each 16-channel block at the same K; it holds output channels 4g .. 4g + 3 of pixel m,
// stored as one 8-byte write. Sums are binary32, rounded once to bfloat16.
constexpr int kIn3Groups = 4;

__global__ __launch_bounds__(256) void k_conv3x3_in3_mfma(const uint16_t *__restrict__ x, long npix, int down,
                                                          int across, const uint16_t *__restrict__ w, int cout,
                                                          int relu, uint16_t *__restrict__ y) {
    // The wavefront's output tile (64 pixels x cout channels, bfloat16) is assembled in LDS and
    // written back in contiguous 16-byte pieces.
    __shared__ __attribute__((aligned(16))) uint16_t so[4][16 * kIn3Groups * kIn3MaxOut];
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long pw = ((long)blockIdx.x * 4 + wv) * 16 * kIn3Groups;
    if (pw >= npix) return;  // no barrier below: each wavefront reads back only its own tile
    // The wavefront's first pixel's image, row and column, once (uniform); each lane steps from there
    // with carries rather than dividing per pixel.
    const long rid0 = pw / across;
    const int c0 = (int)(pw - rid0 * across), r0 = (int)(rid0 % down);
    const long im0 = rid0 / down;
    b8 bx0[kIn3Groups], bx1[kIn3Groups];
#pragma unroll
    for (int q = 0; q < kIn3Groups; ++q) {
        const long p = pw + 16 * q + m;
        const bool pin = p < npix;
        int col = c0 + 16 * q + m, row = r0;
        long img = im0;
        while (col >= across) {  // at most 64 / across + 1 times
            col -= across;
            if (++row == down) {
                row = 0;
                ++img;
            }
        }
        auto tap_in = [&](int t) -> uint2 {
            uint2 v = {0u, 0u};
            const int rr = row + t / 3 - 1;
            if (pin && t < 9 && rr >= 0 && rr < down)
                v = *(const uint2 *)(x + ((img * down + rr) * across + wrap(col + t % 3 - 1, across)) * 16);
            v.y &= 0xFFFFu;
            return v;
        };
        const uint2 i0 = tap_in(2 * g), i1 = tap_in(2 * g + 1), i2 = tap_in(8 + 2 * g);
        bx0[q] = __builtin_bit_cast(b8, (u32x4){i0.x, i0.y, i1.x, i1.y});
        bx1[q] = __builtin_bit_cast(b8, (u32x4){i2.x, i2.y, 0u, 0u});
    }
    for (int ob = 0; ob < cout; ob += 16) {
        const int o = ob + m;
        auto tap_w = [&](int t) -> uint2 {
            uint2 v = {0u, 0u};
            if (t < 9) v = *(const uint2 *)(w + ((long)t * cout + o) * 16);
            v.y &= 0xFFFFu;
            return v;
        };
        const uint2 w0 = tap_w(2 * g), w1 = tap_w(2 * g + 1), w2 = tap_w(8 + 2 * g);
        const b8 aw0 = __builtin_bit_cast(b8, (u32x4){w0.x, w0.y, w1.x, w1.y});
        const b8 aw1 = __builtin_bit_cast(b8, (u32x4){w2.x, w2.y, 0u, 0u});
#pragma unroll
        for (int q = 0; q < kIn3Groups; ++q) {
            f4 acc = (f4){0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw0, bx0[q], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw1, bx1[q], acc, 0, 0, 0);
            float v[4];
            for (int j = 0; j < 4; ++j) v[j] = relu && acc[j] < 0.f ? 0.f : acc[j];
            *(uint2 *)&so[wv][(16 * q + m) * cout + ob + 4 * g] = (uint2){pack2(v[0], v[1]), pack2(v[2], v[3])};
        }
    }
    // The LDS writes above and reads below are one wavefront's, issued in order; the fence keeps the
    // compiler from moving the reads above the writes.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const long left = npix - pw;
    const int pieces = (int)((left < 16 * kIn3Groups ? left : 16 * kIn3Groups) * cout / 8);
    for (int c = lane; c < pieces; c += 64)
        *(uint4 *)(y + pw * cout + 8 * c) = *(const uint4 *)&so[wv][8 * c];
}

// dw[i] = sum over slices s of part[s][i], in slice order (deterministic).
__global__ void k_wgrad_sum(const float *__restrict__ part, int slices, long count, float *__restrict__ dw) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int k = 0; k < slices; ++k) s += part[(long)k * count + i];
    dw[i] = s;
}

// y (down/2 x across/2) = scale * mean of each 2x2 block of x.
__global__ void k_pool2(const uint16_t *__restrict__ x, int ch, int n, int down, int across, float scale,
                        uint16_t *__restrict__ y) {
    const int hd = down / 2, ha = across / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * hd * ha * ch) return;
    const int c = (int)(i % ch);
    const long q = i / ch;
    const int col = (int)(q % ha);
    const long rid = q / ha;
    const int row = (int)(rid % hd);
    const long img = rid / hd;
    const long top = ((img * down + 2 * row) * across + 2 * col) * ch + c;
    const long bot = top + (long)across * ch;
    const float s = float_of(x[top]) + float_of(x[top + ch]) + float_of(x[bot]) + float_of(x[bot + ch]);
    y[i] = bf16_of(0.25f * scale * s);
}

// y (2 down x 2 across) = scale * x at (row / 2, col / 2), where down, across are x's.
__global__ void k_upsample2(const uint16_t *__restrict__ x, int ch, int n, int down, int across, float scale,
                            uint16_t *__restrict__ y) {
    const int fd = 2 * down, fa = 2 * across;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * fd * fa * ch) return;
    const int c = (int)(i % ch);
    const long q = i / ch;
    const int col = (int)(q % fa);
    const long rid = q / fa;
    const int row = (int)(rid % fd);
    const long img = rid / fd;
    const uint16_t v = x[((img * down + row / 2) * across + col / 2) * ch + c];
    y[i] = scale == 1.f ? v : bf16_of(scale * float_of(v));
}

// The same two steps eight channels (16 bytes) a thread, for channel counts divisible by 8.

__global__ void k_pool2_v8(const uint16_t *__restrict__ x, int ch, int n, int down, int across, float scale,
                           uint16_t *__restrict__ y) {
    const int hd = down / 2, ha = across / 2, c8s = ch / 8;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * hd * ha * c8s) return;
    const int c8 = (int)(i % c8s);
    const long q = i / c8s;
    const int col = (int)(q % ha);
    const long rid = q / ha;
    const int row = (int)(rid % hd);
    const long img = rid / hd;
    const long top = ((img * down + 2 * row) * across + 2 * col) * ch + c8 * 8;
    const long bot = top + (long)across * ch;
    const uint4 a = *(const uint4 *)(x + top), b = *(const uint4 *)(x + top + ch);
    const uint4 c = *(const uint4 *)(x + bot), d = *(const uint4 *)(x + bot + ch);
    const float f = 0.25f * scale;
    uint4 o;
    o.x = pack2(f * (lo_of(a.x) + lo_of(b.x) + lo_of(c.x) + lo_of(d.x)), f * (hi_of(a.x) + hi_of(b.x) + hi_of(c.x) + hi_of(d.x)));
    o.y = pack2(f * (lo_of(a.y) + lo_of(b.y) + lo_of(c.y) + lo_of(d.y)), f * (hi_of(a.y) + hi_of(b.y) + hi_of(c.y) + hi_of(d.y)));
    o.z = pack2(f * (lo_of(a.z) + lo_of(b.z) + lo_of(c.z) + lo_of(d.z)), f * (hi_of(a.z) + hi_of(b.z) + hi_of(c.z) + hi_of(d.z)));
    o.w = pack2(f * (lo_of(a.w) + lo_of(b.w) + lo_of(c.w) + lo_of(d.w)), f * (hi_of(a.w) + hi_of(b.w) + hi_of(c.w) + hi_of(d.w)));
    *(uint4 *)(y + i * 8) = o;
}

__global__ void k_upsample2_v8(const uint16_t *__restrict__ x, int ch, int n, int down, int across, float scale,
                               uint16_t *__restrict__ y) {
    const int fd = 2 * down, fa = 2 * across, c8s = ch / 8;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * fd * fa * c8s) return;
    const int c8 = (int)(i % c8s);
    const long q = i / c8s;
    const int col = (int)(q % fa);
    const long rid = q / fa;
    const int row = (int)(rid % fd);
    const long img = rid / fd;
    uint4 v = *(const uint4 *)(x + ((img * down + row / 2) * across + col / 2) * ch + c8 * 8);
    if (scale != 1.f) {
        v.x = pack2(scale * lo_of(v.x), scale * hi_of(v.x));
        v.y = pack2(scale * lo_of(v.y), scale * hi_of(v.y));
        v.z = pack2(scale * lo_of(v.z), scale * hi_of(v.z));
        v.w = pack2(scale * lo_of(v.w), scale * hi_of(v.w));
    }
    *(uint4 *)(y + i * 8) = v;
}

// wt[8 - t][i][o] = w[t][o][i]: the weights with which the forward convolution computes the data
// gradient of the convolution by w (taps mirrored, input and output channels exchanged).
__global__ void k_flip_weights(const uint16_t *__restrict__ w, int cout, int cin, uint16_t *__restrict__ wt) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per_tap = (long)cout * cin;
    if (i >= 9 * per_tap) return;
    const int t = (int)(i / per_tap), o = (int)((i / cin) % cout), c = (int)(i % cin);
    wt[((long)(8 - t) * cin + c) * cout + o] = w[i];
}

// dy = y > 0 ? dy : 0 (the gradient through a ReLU whose output was y).
__global__ void k_relu_mask(uint16_t *__restrict__ dy, const uint16_t *__restrict__ y, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (!(float_of(y[i]) > 0.f)) dy[i] = 0;
}

// The same eight elements (16 bytes) a thread, for counts divisible by 8.
__global__ void k_relu_mask_v8(uint16_t *__restrict__ dy, const uint16_t *__restrict__ y, long count8) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count8) return;
    const uint4 yv = *(const uint4 *)(y + i * 8);
    uint4 g = *(const uint4 *)(dy + i * 8);
    const uint32_t yw[4] = {yv.x, yv.y, yv.z, yv.w};
    uint32_t gw[4] = {g.x, g.y, g.z, g.w};
    for (int k = 0; k < 4; ++k) {
        if (!(lo_of(yw[k]) > 0.f)) gw[k] &= 0xffff0000u;
        if (!(hi_of(yw[k]) > 0.f)) gw[k] &= 0x0000ffffu;
    }
    *(uint4 *)(dy + i * 8) = (uint4){gw[0], gw[1], gw[2], gw[3]};
}

__global__ void k_add_bf16_v8(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b, long count8,
                              uint16_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count8) return;
    const uint4 av = *(const uint4 *)(a + i * 8), bv = *(const uint4 *)(b + i * 8);
    const uint32_t aw[4] = {av.x, av.y, av.z, av.w}, bw[4] = {bv.x, bv.y, bv.z, bv.w};
    uint32_t o[4];
    for (int k = 0; k < 4; ++k) o[k] = pack2(lo_of(aw[k]) + lo_of(bw[k]), hi_of(aw[k]) + hi_of(bw[k]));
    *(uint4 *)(out + i * 8) = (uint4){o[0], o[1], o[2], o[3]};
}

// out = a + b, rounded once.
__global__ void k_add_bf16(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b, long count,
                           uint16_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    out[i] = bf16_of(float_of(a[i]) + float_of(b[i]));
}

// The network's input: 16 channels per cell, channels 0..1 copied from `fixed` (mask, depth),
// channel 2 = scale * v[cell] (the vector to precondition), the rest zero.
__global__ void k_pack_input(const uint16_t *__restrict__ fixed, const float *__restrict__ v, float scale, long cells,
                             uint16_t *__restrict__ out) {
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const uint32_t f = *(const uint32_t *)(fixed + cell * 2);  // channels 0 and 1
    *(uint4 *)(out + cell * 16) = (uint4){f, (uint32_t)bf16_of(scale * v[cell]), 0u, 0u};
    *(uint4 *)(out + cell * 16 + 8) = (uint4){0u, 0u, 0u, 0u};
}

// z[cell] = water ? scale * y32[cell][0] : 0 (channel 0 of the network's binary32 output).
__global__ void k_unpack_output(const float *__restrict__ y32, const uint8_t *__restrict__ water, float scale,
                                long cells, float *__restrict__ z) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    z[i] = water[i] ? scale * y32[i * 16] : 0.f;
}

inline unsigned blocks(long count, int per) { return (unsigned)((count + per - 1) / per); }

bool conv_shape_ok(int c1, int c2, int n, int down, int across, int cout) {
    return c1 > 0 && c1 % 16 == 0 && c2 >= 0 && c2 % 16 == 0 && n > 0 && down > 0 && across > 0 &&
           across % 16 == 0 && cout > 0 && cout % 16 == 0;
}

// Output-channel tiles of 64, 32 or 16 a wavefront: the widest the channel count takes, narrowed while
// the launch would hold fewer than 512 workgroups (small images: their waves, each summing all
// 9 x cin products of its tile, are the whole launch). Each output's sum is the same whatever the
// tile, so the choice changes no bits.
int pick_nb(long row_blocks, int cout) {
    int nb = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
    while (nb > 1 && row_blocks * (cout / (16 * nb)) < 512) nb >>= 1;
    return nb;
}

template <int MT>
void conv_k32_launch(const uint16_t *x1, int c1, const uint16_t *x2, int c2, int n, int down, int across,
                     const uint16_t *w, int cout, int relu, uint16_t *y1, int o1, uint16_t *y2, float *y32,
                     hipStream_t s) {
    const long tiles = (long)n * down * (across / (16 * MT));
    const int nb = pick_nb((long)blocks(tiles, 4), cout);
    const dim3 grid(blocks(tiles, 4), cout / (16 * nb));
    if (nb == 4)
        hipLaunchKernelGGL((k_conv3x3_k32<MT, 4>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w, cout,
                           relu, y1, o1, y2, y32);
    else if (nb == 2)
        hipLaunchKernelGGL((k_conv3x3_k32<MT, 2>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w, cout,
                           relu, y1, o1, y2, y32);
    else
        hipLaunchKernelGGL((k_conv3x3_k32<MT, 1>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w, cout,
                           relu, y1, o1, y2, y32);
}

// What the current device holds of k_conv3x3_tile<nb>: its CU count and the workgroups it keeps
// resident at once (the staged kernel's grid, each workgroup walking tiles until none is left).
struct TileFit {
    int cus;
    long resident;
};
TileFit tile_fit(int nb) {
    static std::atomic<long> cached[64][3];
    static std::atomic<int> cached_cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    const int v = nb == 4 ? 0 : (nb == 2 ? 1 : 2);
    long have = cached[dev][v].load(std::memory_order_relaxed);
    if (have > 0) return {cached_cus[dev].load(std::memory_order_relaxed), have};
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    hipError_t e = nb == 4   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3_tile<4>, 256, 0)
                   : nb == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3_tile<2>, 256, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3_tile<1>, 256, 0);
    if (e != hipSuccess || per_cu <= 0) per_cu = 1;
    have = (long)cus * per_cu;
    cached_cus[dev].store(cus, std::memory_order_relaxed);
    cached[dev][v].store(have, std::memory_order_relaxed);
    return {cus, have};
}

bool tile_ok(int c1, int c2, int down, int across) {
    return c1 % 32 == 0 && c2 % 32 == 0 && across % 64 == 0 && down % 4 == 0;
}

// Images of fewer pixels than this take the one-row kernel with 16-pixel tiles rather than 64.
constexpr long kSmallImage = 65536;

template <bool ADJ>
int conv_launch(const uint16_t *x1, int c1, const uint16_t *x2, int c2, int n, int down, int across,
                const uint16_t *w, int cout, int relu, uint16_t *y1, int o1, uint16_t *y2, float *y32, hipStream_t s,
                int up1 = 0) {
    // The staged kernel unless it would have fewer tiles (times output-channel blocks) than the device
    // has CUs: such layers take the one-row kernels below (more, shorter wavefronts), except where
    // the input is read through upsampling, which only the staged kernel does.
    if (!ADJ && tile_ok(c1, c2, down, across)) {
        const int nb = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
        const long tiles = (long)n * (down / 4) * (across / 64);
        const int ys = cout / (16 * nb);
        const TileFit fit = tile_fit(nb);
        if (up1 || tiles * ys >= fit.cus) {
            const long per_y = std::max<long>(1, (fit.resident + ys - 1) / ys);
            const dim3 grid((unsigned)std::min<long>(tiles, per_y), ys);
            if (nb == 4)
                hipLaunchKernelGGL((k_conv3x3_tile<4>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w,
                                   cout, relu, y1, o1, y2, y32, up1);
            else if (nb == 2)
                hipLaunchKernelGGL((k_conv3x3_tile<2>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w,
                                   cout, relu, y1, o1, y2, y32, up1);
            else
                hipLaunchKernelGGL((k_conv3x3_tile<1>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w,
                                   cout, relu, y1, o1, y2, y32, up1);
            return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
        }
    }
    const bool small = (long)n * down * across < kSmallImage;
    if (!ADJ && c1 % 32 == 0 && c2 % 32 == 0) {
        if (across % 64 == 0 && !small)
            conv_k32_launch<4>(x1, c1, x2, c2, n, down, across, w, cout, relu, y1, o1, y2, y32, s);
        else
            conv_k32_launch<1>(x1, c1, x2, c2, n, down, across, w, cout, relu, y1, o1, y2, y32, s);
        return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
    }
    const long tiles = (long)n * down * (across / 16);
    const int nb = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
    const dim3 grid(blocks(tiles, 4), cout / (16 * nb));
    if (nb == 4)
        hipLaunchKernelGGL((k_conv3x3<4, ADJ>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w, cout, relu,
                           y1, o1, y2, y32);
    else if (nb == 2)
        hipLaunchKernelGGL((k_conv3x3<2, ADJ>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w, cout, relu,
                           y1, o1, y2, y32);
    else
        hipLaunchKernelGGL((k_conv3x3<1, ADJ>), grid, dim3(256), 0, s, x1, c1, x2, c2, n, down, across, w, cout, relu,
                           y1, o1, y2, y32);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

}  // namespace

extern "C" {

int cg_conv3x3(const uint16_t *x1, int c1, const uint16_t *x2, int c2, int n, int down, int across,
               const uint16_t *w, int cout, int relu, uint16_t *y1, int o1, uint16_t *y2, float *y32,
               void *stream) {
    if (!conv_shape_ok(c1, c2, n, down, across, cout) || !x1 || (c2 && !x2) || !w) return CG_ERR_ARG;
    if (!y32 && (!y1 || o1 < 0 || o1 > cout || o1 % 16 || (o1 < cout && !y2))) return CG_ERR_ARG;
    return conv_launch<false>(x1, c1, c2 ? x2 : nullptr, c2, n, down, across, w, cout, relu, y1, y32 ? cout : o1, y2,
                              y32, (hipStream_t)stream);
}

int cg_conv3x3_in3(const uint16_t *x, int n, int down, int across, const uint16_t *w, int cout, int relu, uint16_t *y,
                   void *stream) {
    if (!x || !w || !y || n <= 0 || down <= 0 || across <= 0 || cout <= 0 || cout > kIn3MaxOut || cout % 8)
        return CG_ERR_ARG;
    // The kernel reads each input pixel's channels 0..3 as one 8-byte load and writes outputs 16 bytes
    // at a time.
    if (reinterpret_cast<uintptr_t>(x) % 8 || reinterpret_cast<uintptr_t>(y) % 16) return CG_ERR_ARG;
    const long pixels = (long)n * down * across;
    if (cout % 16 == 0 && reinterpret_cast<uintptr_t>(w) % 8 == 0)  // matrix cores: 64 pixels a wavefront
        hipLaunchKernelGGL(k_conv3x3_in3_mfma, dim3(blocks(pixels, 64 * kIn3Groups)), dim3(256), 0,
                           (hipStream_t)stream, x, pixels, down, across, w, cout, relu, y);
    else
        hipLaunchKernelGGL(k_conv3x3_in3, dim3(blocks(pixels, 256)), dim3(256), 0, (hipStream_t)stream, x, n, down,
                           across, w, cout, relu, y);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_conv3x3_up(const uint16_t *x1, int c1, const uint16_t *x2, int c2, int n, int down, int across,
                  const uint16_t *w, int cout, int relu, uint16_t *y1, int o1, uint16_t *y2, float *y32,
                  void *stream) {
    if (!conv_shape_ok(c1, c2, n, down, across, cout) || !tile_ok(c1, c2, down, across) || !x1 || (c2 && !x2) || !w)
        return CG_ERR_ARG;
    if (!y32 && (!y1 || o1 < 0 || o1 > cout || o1 % 16 || (o1 < cout && !y2))) return CG_ERR_ARG;
    return conv_launch<false>(x1, c1, c2 ? x2 : nullptr, c2, n, down, across, w, cout, relu, y1, y32 ? cout : o1, y2,
                              y32, (hipStream_t)stream, 1);
}

int cg_conv3x3_adjoint(const uint16_t *dy, int cout, int n, int down, int across, const uint16_t *w, int cin,
                       uint16_t *dx1, int c1, uint16_t *dx2, float *dx32, void *stream) {
    if (!conv_shape_ok(cout, 0, n, down, across, cin) || !dy || !w) return CG_ERR_ARG;
    if (!dx32 && (!dx1 || c1 < 0 || c1 > cin || c1 % 16 || (c1 < cin && !dx2))) return CG_ERR_ARG;
    return conv_launch<true>(dy, cout, nullptr, 0, n, down, across, w, cin, 0, dx1, dx32 ? cin : c1, dx2, dx32,
                             (hipStream_t)stream);
}

int cg_conv3x3_flip_weights(const uint16_t *w, int cout, int cin, uint16_t *wt, void *stream) {
    if (!w || !wt || cout <= 0 || cin <= 0) return CG_ERR_ARG;
    hipLaunchKernelGGL(k_flip_weights, dim3(blocks(9L * cout * cin, 256)), dim3(256), 0, (hipStream_t)stream, w, cout,
                       cin, wt);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_conv3x3_wgrad_slices(int n, int down, int across) {
    const long groups = (long)n * down * across / 16;
    long s = groups / 32;
    if (s > 256) s = 256;
    s = (s / 4) * 4;
    return s < 4 ? 4 : (int)s;
}

int cg_conv3x3_wgrad(const uint16_t *x1, int c1, const uint16_t *x2, int c2, const uint16_t *dy, int cout, int n,
                     int down, int across, float *work, size_t work_floats, float *dw, void *stream) {
    if (!conv_shape_ok(c1, c2, n, down, across, cout) || !x1 || (c2 && !x2) || !dy || !work || !dw)
        return CG_ERR_ARG;
    const int slices = cg_conv3x3_wgrad_slices(n, down, across);
    const long count = 9L * cout * (c1 + c2);
    if (work_floats < (size_t)slices * (size_t)count) return CG_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (across % 32 == 0)
        hipLaunchKernelGGL(k_conv3x3_wgrad32, dim3((cout / 16) * ((c1 + c2) / 16), slices), dim3(64), 0, s, x1, c1,
                           c2 ? x2 : nullptr, c2, dy, cout, n, down, across, slices, work);
    else
        hipLaunchKernelGGL(k_conv3x3_wgrad, dim3((cout / 16) * ((c1 + c2) / 16), slices / 4), dim3(256), 0, s, x1, c1,
                           c2 ? x2 : nullptr, c2, dy, cout, n, down, across, slices, work);
    hipLaunchKernelGGL(k_wgrad_sum, dim3(blocks(count, 256)), dim3(256), 0, s, work, slices, count, dw);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_pool2(const uint16_t *x, int channels, int n, int down, int across, float scale, uint16_t *y, void *stream) {
    if (!x || !y || channels <= 0 || n <= 0 || down <= 0 || across <= 0 || down % 2 || across % 2) return CG_ERR_ARG;
    const long count = (long)n * (down / 2) * (across / 2) * channels;
    if (channels % 8 == 0)
        hipLaunchKernelGGL(k_pool2_v8, dim3(blocks(count / 8, 256)), dim3(256), 0, (hipStream_t)stream, x, channels, n,
                           down, across, scale, y);
    else
        hipLaunchKernelGGL(k_pool2, dim3(blocks(count, 256)), dim3(256), 0, (hipStream_t)stream, x, channels, n, down,
                           across, scale, y);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_upsample2(const uint16_t *x, int channels, int n, int down, int across, float scale, uint16_t *y,
                 void *stream) {
    if (!x || !y || channels <= 0 || n <= 0 || down <= 0 || across <= 0) return CG_ERR_ARG;
    const long count = (long)n * (2 * down) * (2 * across) * channels;
    if (channels % 8 == 0)
        hipLaunchKernelGGL(k_upsample2_v8, dim3(blocks(count / 8, 256)), dim3(256), 0, (hipStream_t)stream, x, channels,
                           n, down, across, scale, y);
    else
        hipLaunchKernelGGL(k_upsample2, dim3(blocks(count, 256)), dim3(256), 0, (hipStream_t)stream, x, channels, n,
                           down, across, scale, y);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_relu_mask(uint16_t *dy, const uint16_t *y, long count, void *stream) {
    if (!dy || !y || count < 0) return CG_ERR_ARG;
    if (count == 0) return CG_OK;
    if (count % 8 == 0 && (uintptr_t)dy % 16 == 0 && (uintptr_t)y % 16 == 0)
        hipLaunchKernelGGL(k_relu_mask_v8, dim3(blocks(count / 8, 256)), dim3(256), 0, (hipStream_t)stream, dy, y,
                           count / 8);
    else
        hipLaunchKernelGGL(k_relu_mask, dim3(blocks(count, 256)), dim3(256), 0, (hipStream_t)stream, dy, y, count);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_add_bf16(const uint16_t *a, const uint16_t *b, long count, uint16_t *out, void *stream) {
    if (!a || !b || !out || count < 0) return CG_ERR_ARG;
    if (count == 0) return CG_OK;
    if (count % 8 == 0 && (uintptr_t)a % 16 == 0 && (uintptr_t)b % 16 == 0 && (uintptr_t)out % 16 == 0)
        hipLaunchKernelGGL(k_add_bf16_v8, dim3(blocks(count / 8, 256)), dim3(256), 0, (hipStream_t)stream, a, b,
                           count / 8, out);
    else
        hipLaunchKernelGGL(k_add_bf16, dim3(blocks(count, 256)), dim3(256), 0, (hipStream_t)stream, a, b, count, out);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_unet_pack_input(const uint16_t *fixed, const float *v, float scale, long cells, uint16_t *out,
                       void *stream) {
    if (!fixed || !v || !out || cells <= 0) return CG_ERR_ARG;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks(cells, 256)), dim3(256), 0, (hipStream_t)stream, fixed, v, scale,
                       cells, out);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

int cg_unet_unpack_output(const float *y32, const uint8_t *water, float scale, long cells, float *z,
                          void *stream) {
    if (!y32 || !water || !z || cells <= 0) return CG_ERR_ARG;
    hipLaunchKernelGGL(k_unpack_output, dim3(blocks(cells, 256)), dim3(256), 0, (hipStream_t)stream, y32, water,
                       scale, cells, z);
    return hipGetLastError() == hipSuccess ? CG_OK : CG_ERR_DEVICE;
}

}  // extern "C"
