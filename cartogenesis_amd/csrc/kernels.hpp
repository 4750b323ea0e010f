// Device kernels of the ocean-circulation solver for gfx950 (CDNA4, wave64).
//
// The whole library is compiled with -ffp-contract=off: every float and double operation below
// rounds once, in the order written. That is what makes the relaxation, the binary64 residual, the
// restriction and the prolongation bit-identical to the reference's JVM arithmetic
// (OceanCirculation.kt, OC below: relax/updated OC:241-274, residual OC:292-315, restrict
// OC:334-346, prolongAdd OC:354-392). Where a fused multiply-add is wanted it is written as fma().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cg {

// One grid of the ladder as the kernels see it. Weights are either one float4 (east, west, north,
// south) per row (`roww`, the circulation's row-constant weights) or one per cell (`cellw`).
// The centre weight likewise per row (`rowc`) or per cell (`cellc`).
struct LevelArgs {
    int across;
    int down;
    const uint8_t *water;
    const float4 *roww;
    const float4 *cellw;
    const double *rowc;
    const double *cellc;
};

__device__ __forceinline__ float4 weights_at(const LevelArgs &L, int row, long cell) {
    return L.roww ? L.roww[row] : L.cellw[cell];
}
__device__ __forceinline__ double centre_at(const LevelArgs &L, int row, long cell) {
    return L.rowc ? L.rowc[row] : L.cellc[cell];
}

// Binary32 update of one cell in the reference's order: ((e xE + w xW) + n xN) + s xS - f.
__device__ __forceinline__ float rb_update(float4 wt, float xe, float xw, float xn, float xs, float f) {
    return wt.x * xe + wt.y * xw + wt.z * xn + wt.w * xs - f;
}

// ------------------------------------------------------------------------------------------------
// Fused red-black smoother, row-constant weights.
//
// One wavefront owns a tile of TH rows by (128 - 2 HC) columns plus a halo of HR = NH rows and
// HC (NH rounded up to even) columns. Lane l holds the two adjacent columns 2l and 2l+1 of every
// tile row in registers (A = even column, B = odd column), so in each half-sweep every lane updates
// exactly one of its two cells per row: east/west neighbours are the lane's other register or its
// neighbour lane's (one cross-lane shuffle per row), north/south neighbours are the same lane's
// registers of the adjacent rows. NH half-sweeps (colour 0 first) run without any barrier or LDS;
// each half-sweep invalidates one more cell of the halo, so after NH the interior is exact and is
// written back. Rows beyond a pole are zero and never change, like the reference's zero beyond the
// pole; columns wrap by modular loading. Reads xin, writes xout (the two must differ: neighbouring
// tiles read the halo this tile would otherwise overwrite).
// ------------------------------------------------------------------------------------------------
template <int NH, int TH, bool ZERO_START>
__global__ __launch_bounds__(256) void k_relax_tile(LevelArgs L, const float *__restrict__ f,
                                                    const float *__restrict__ xin, float *__restrict__ xout) {
    constexpr int HR = NH;
    constexpr int HC = (NH + 1) & ~1;
    constexpr int R = TH + 2 * HR;
    constexpr int IW = 128 - 2 * HC;
    static_assert(R <= 64, "tile rows must fit a 64-bit row mask");
    const int across = L.across, down = L.down;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tiles_x = (across + IW - 1) / IW;
    const int tiles_y = (down + TH - 1) / TH;
    if (wave >= tiles_x * tiles_y) return;
    const int tx = wave % tiles_x, ty = wave / tiles_x;
    const int col0 = tx * IW, row0 = ty * TH;
    int ca = (col0 - HC + 2 * lane) % across;
    if (ca < 0) ca += across;

    float A[R], B[R], fA[R], fB[R];
    uint64_t wA = 0, wB = 0;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = row0 - HR + i;
        if (r >= 0 && r < down) {
            const long idx = (long)r * across + ca;
            const float2 fv = *reinterpret_cast<const float2 *>(f + idx);
            const uint16_t m = *reinterpret_cast<const uint16_t *>(L.water + idx);
            fA[i] = fv.x;
            fB[i] = fv.y;
            wA |= (uint64_t)((m & 0xff) != 0) << i;
            wB |= (uint64_t)((m >> 8) != 0) << i;
            if (ZERO_START) {
                A[i] = 0.0f;
                B[i] = 0.0f;
            } else {
                const float2 xv = *reinterpret_cast<const float2 *>(xin + idx);
                A[i] = xv.x;
                B[i] = xv.y;
            }
        } else {
            fA[i] = fB[i] = A[i] = B[i] = 0.0f;
        }
    }

    // Skewed schedule: half-sweep h reaches row i at step t = i + h. A colour-k update reads only
    // other-colour cells (its row's pair partner and the same column in the rows above and below),
    // which half-sweep h leaves untouched, so row i-1 being one half-sweep ahead is harmless; and
    // row i+1 has finished half-sweep h-1 earlier in the same step. The result equals NH full
    // half-sweeps in order, and each row's weights are live for only NH consecutive steps.
#pragma unroll
    for (int t = 1; t < R - 2 + NH; ++t) {
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int i = t - h;
            if (i < 1 || i > R - 2) continue;
            const int colour = h & 1;
            int r = row0 - HR + i;
            r = r < 0 ? 0 : (r >= down ? down - 1 : r);
            const float4 wt = L.roww[r];
            // Row (row0 - HR + i) has the parity of (i + HR) since row0 is even; the even column
            // of the pair has colour (row & 1).
            if (((i + HR) & 1) == colour) {
                const float xw = __shfl_up(B[i], 1, 64);
                const float v = rb_update(wt, B[i], xw, A[i - 1], A[i + 1], fA[i]);
                A[i] = ((wA >> i) & 1) ? v : 0.0f;
            } else {
                const float xe = __shfl_down(A[i], 1, 64);
                const float v = rb_update(wt, xe, A[i], B[i - 1], B[i + 1], fB[i]);
                B[i] = ((wB >> i) & 1) ? v : 0.0f;
            }
        }
    }

    const int lo = HC / 2;
    const int cabs = col0 + 2 * (lane - lo);
    if (lane >= lo && lane < lo + IW / 2 && cabs < across) {
#pragma unroll
        for (int i = HR; i < HR + TH; ++i) {
            const int r = row0 - HR + i;
            if (r < down) *reinterpret_cast<float2 *>(xout + (long)r * across + cabs) = make_float2(A[i], B[i]);
        }
    }
}

// One colour of one red-black pass, in place, any weights (per row or per cell): the general and
// the per-cell (heat) smoother. Thread t updates the t-th cell of the colour.
__global__ void k_relax_colour(LevelArgs L, const float *__restrict__ f, float *x, int colour) {
    const int across = L.across, half = across >> 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)half * L.down) return;
    const int row = (int)(t / half);
    const int col = 2 * (int)(t - (long)row * half) + ((colour + row) & 1);
    const long cell = (long)row * across + col;
    if (!L.water[cell]) {
        x[cell] = 0.0f;
        return;
    }
    const int ce = col + 1 == across ? 0 : col + 1;
    const int cw = col == 0 ? across - 1 : col - 1;
    const float xn = row > 0 ? x[cell - across] : 0.0f;
    const float xs = row + 1 < L.down ? x[cell + across] : 0.0f;
    x[cell] = rb_update(weights_at(L, row, cell), x[(long)row * across + ce], x[(long)row * across + cw], xn, xs, f[cell]);
}

// ------------------------------------------------------------------------------------------------
// Reductions: the largest magnitude of non-negative doubles, as uint64 bit patterns. A NaN's
// pattern exceeds every finite and infinite magnitude, so a NaN anywhere wins, as Kotlin's max
// (the reference's `largest`, OC:318-322) returns NaN when either side is NaN.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long abs_bits(double v) {
    return (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull;
}

__device__ __forceinline__ void block_max_atomic(unsigned long long v, unsigned long long *out) {
    // Wave reduction then one atomic per wave.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(out, v);
}

// Binary64 residual (OC:292-315): r = (x - (sum w x - f)) * centre on water, 0 on land; and its
// largest magnitude into *rmax when given.
__global__ void k_residual64(LevelArgs L, const float *__restrict__ f, const float *__restrict__ x,
                             double *__restrict__ r, unsigned long long *rmax) {
    const int across = L.across;
    const long cells = (long)across * L.down;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long m = 0;
    if (cell < cells) {
        double out = 0.0;
        if (L.water[cell]) {
            const int row = (int)(cell / across);
            const int col = (int)(cell - (long)row * across);
            const int ce = col + 1 == across ? 0 : col + 1;
            const int cw = col == 0 ? across - 1 : col - 1;
            const double vn = row > 0 ? (double)x[cell - across] : 0.0;
            const double vs = row + 1 < L.down ? (double)x[cell + across] : 0.0;
            const float4 wt = weights_at(L, row, cell);
            const double bal = (double)wt.x * (double)x[(long)row * across + ce] +
                               (double)wt.y * (double)x[(long)row * across + cw] + (double)wt.z * vn +
                               (double)wt.w * vs - (double)f[cell];
            out = ((double)x[cell] - bal) * centre_at(L, row, cell);
        }
        if (r) r[cell] = out;
        m = abs_bits(out);
    }
    if (rmax) block_max_atomic(m, rmax);
}

// The largest magnitude of f * centre (balanceOf then largest, OC:325-327, 318-322).
__global__ void k_balance_max(LevelArgs L, const float *__restrict__ f, unsigned long long *out) {
    const long cells = (long)L.across * L.down;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long m = 0;
    if (cell < cells) {
        const int row = (int)(cell / L.across);
        m = abs_bits((double)f[cell] * centre_at(L, row, cell));
    }
    block_max_atomic(m, out);
}

// restrict (OC:334-346): each coarse cell sums fine * share over the fine cells it covers in the
// reference's row-major order, so the sum is bit-identical to the sequential loop.
__global__ void k_restrict64(const double *__restrict__ fine, int fa, int fd, int ca, int cd,
                             double *__restrict__ coarse) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ca * cd) return;
    const int crow = (int)(t / ca), ccol = (int)(t - (long)crow * ca);
    const int cpc = fa / ca, rpc = fd / cd;
    const double share = 1.0 / (double)(cpc * rpc);
    double acc = 0.0;
    for (int dr = 0; dr < rpc; ++dr)
        for (int dc = 0; dc < cpc; ++dc) acc += fine[(long)(crow * rpc + dr) * fa + ccol * cpc + dc] * share;
    coarse[t] = acc;
}

// withBalance (OC:222-231): f = (balance / centre) rounded to float on water, 0 on land.
__global__ void k_with_balance(LevelArgs L, const double *__restrict__ balance, float *__restrict__ f) {
    const long cells = (long)L.across * L.down;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const int row = (int)(cell / L.across);
    f[cell] = L.water[cell] ? (float)(balance[cell] / centre_at(L, row, cell)) : 0.0f;
}

// prolongAdd (OC:354-392) in the reference's binary32 arithmetic: bilinear between coarse cell
// centres, columns wrapping, beyond a pole the edge row times `beyond` (-1 for a wall).
__global__ void k_prolong_add(const float *__restrict__ coarse, int ca, int cd, float *__restrict__ fine, int fa,
                              int fd, const uint8_t *__restrict__ water, float beyond) {
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (long)fa * fd) return;
    if (!water[cell]) return;
    const int row = (int)(cell / fa), col = (int)(cell - (long)row * fa);
    const int cpc = fa / ca, rpc = fd / cd;
    const float crow = ((float)row + 0.5f) / (float)rpc - 0.5f;
    const int above = (int)floorf(crow);
    const float db = crow - (float)above;
    const float ccol = ((float)col + 0.5f) / (float)cpc - 0.5f;
    const int left = (int)floorf(ccol);
    const float ab = ccol - (float)left;
    const int wc = (left + ca) % ca, ec = (left + 1) % ca;
    auto at = [&](int rr, int cc) -> float {
        if (rr < 0) return beyond * coarse[cc];
        if (rr >= cd) return beyond * coarse[(long)(cd - 1) * ca + cc];
        return coarse[(long)rr * ca + cc];
    };
    const float top = at(above, wc) * (1.0f - ab) + at(above, ec) * ab;
    const float bottom = at(above + 1, wc) * (1.0f - ab) + at(above + 1, ec) * ab;
    fine[cell] += top * (1.0f - db) + bottom * db;
}

__global__ void k_fill_f32(float *p, long n, float v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// relaxToConvergence (OC:784-794) on one workgroup: relax in passes of 20 until the binary64
// residual's largest magnitude is at most tol times the largest balance, or 4000 passes. Values,
// forcing and mask live in global memory (the coarsest grid is a few hundred cells, L2-resident);
// __syncthreads separates the colours. Also runs a fixed number of passes (fixed > 0) for the
// preconditioner's cycle (OC:767).
// ------------------------------------------------------------------------------------------------
__device__ void wg_relax_pass(const LevelArgs &L, const float *f, float *x) {
    const int across = L.across, half = across >> 1;
    const int n = half * L.down;
    for (int colour = 0; colour <= 1; ++colour) {
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int row = t / half;
            const int col = 2 * (t - row * half) + ((colour + row) & 1);
            const long cell = (long)row * across + col;
            if (!L.water[cell]) {
                x[cell] = 0.0f;
                continue;
            }
            const int ce = col + 1 == across ? 0 : col + 1;
            const int cw = col == 0 ? across - 1 : col - 1;
            const float xn = row > 0 ? x[cell - across] : 0.0f;
            const float xs = row + 1 < L.down ? x[cell + across] : 0.0f;
            x[cell] = rb_update(weights_at(L, row, cell), x[(long)row * across + ce], x[(long)row * across + cw], xn, xs,
                                f[cell]);
        }
        __syncthreads();
    }
}

__device__ double wg_max(double v, double *sh) {
    unsigned long long b = abs_bits(v);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off, 64);
        b = o > b ? o : b;
    }
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = __longlong_as_double((long long)b);
    __syncthreads();
    unsigned long long m = 0;
    for (int i = 0; i < nw; ++i) {
        const unsigned long long o = abs_bits(sh[i]);
        m = o > m ? o : m;
    }
    __syncthreads();
    return __longlong_as_double((long long)m);
}

__device__ double wg_residual_max(const LevelArgs &L, const float *f, const float *x, double *sh) {
    const int across = L.across;
    const long cells = (long)across * L.down;
    unsigned long long m = 0;
    for (long cell = threadIdx.x; cell < cells; cell += blockDim.x) {
        if (!L.water[cell]) continue;
        const int row = (int)(cell / across), col = (int)(cell - (long)row * across);
        const int ce = col + 1 == across ? 0 : col + 1;
        const int cw = col == 0 ? across - 1 : col - 1;
        const double vn = row > 0 ? (double)x[cell - across] : 0.0;
        const double vs = row + 1 < L.down ? (double)x[cell + across] : 0.0;
        const float4 wt = weights_at(L, row, cell);
        const double bal = (double)wt.x * (double)x[(long)row * across + ce] + (double)wt.y * (double)x[(long)row * across + cw] +
                           (double)wt.z * vn + (double)wt.w * vs - (double)f[cell];
        const unsigned long long b = abs_bits(((double)x[cell] - bal) * centre_at(L, row, cell));
        m = b > m ? b : m;
    }
    return wg_max(__longlong_as_double((long long)m), sh);
}

__global__ __launch_bounds__(1024) void k_coarsest_relax(LevelArgs L, const float *__restrict__ f, float *x,
                                                         double tol, int fixed_passes, int *passes_out) {
    __shared__ double sh[16];
    if (fixed_passes > 0) {
        for (int p = 0; p < fixed_passes; ++p) wg_relax_pass(L, f, x);
        if (threadIdx.x == 0 && passes_out) *passes_out = fixed_passes;
        return;
    }
    const int across = L.across;
    const long cells = (long)across * L.down;
    unsigned long long lb = 0;
    for (long cell = threadIdx.x; cell < cells; cell += blockDim.x) {
        const int row = (int)(cell / across);
        const unsigned long long b = abs_bits((double)f[cell] * centre_at(L, row, cell));
        lb = b > lb ? b : lb;
    }
    const double largest = wg_max(__longlong_as_double((long long)lb), sh);
    int passes = 0;
    if (largest > 0.0) {
        while (passes < 200 * 20) {
            const double res = wg_residual_max(L, f, x, sh);
            if (!(res > largest * tol)) break;
            for (int p = 0; p < 20; ++p) wg_relax_pass(L, f, x);
            passes += 20;
        }
    }
    if (threadIdx.x == 0 && passes_out) *passes_out = passes;
}

// ------------------------------------------------------------------------------------------------
// Kernels of the refined (binary64 outer, binary32 inner) solve.
// ------------------------------------------------------------------------------------------------

// The outer step: x_new = x + delta (binary64; delta may be null), u = x_new - (sum w x_new - f)
// in binary64, the correction forcing fc = (float)u on water (0 on land), and the largest
// |u * centre| into *rmax. Land values are zero in x_new.
__global__ void k_outer_residual(LevelArgs L, const float *__restrict__ f, const double *__restrict__ x,
                                 const float *__restrict__ delta, double *__restrict__ xnew, float *__restrict__ fc,
                                 unsigned long long *rmax) {
    const int across = L.across;
    const long cells = (long)across * L.down;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long m = 0;
    if (cell < cells) {
        const int row = (int)(cell / across), col = (int)(cell - (long)row * across);
        double v = 0.0, u = 0.0;
        if (L.water[cell]) {
            const int ce = col + 1 == across ? 0 : col + 1;
            const int cw = col == 0 ? across - 1 : col - 1;
            const long ie = (long)row * across + ce, iw = (long)row * across + cw;
            auto val = [&](long i) -> double { return delta ? x[i] + (double)delta[i] : x[i]; };
            v = val(cell);
            const double vn = row > 0 ? val(cell - across) : 0.0;
            const double vs = row + 1 < L.down ? val(cell + across) : 0.0;
            const float4 wt = weights_at(L, row, cell);
            const double bal = (double)wt.x * val(ie) + (double)wt.y * val(iw) + (double)wt.z * vn + (double)wt.w * vs -
                               (double)f[cell];
            u = v - bal;
            m = abs_bits(u * centre_at(L, row, cell));
        }
        xnew[cell] = v;
        fc[cell] = (float)u;
    }
    block_max_atomic(m, rmax);
}

// Binary32 residual of the correction problem restricted to the next grid, fused: each thread owns
// one coarse cell, sums u * centre_fine * share over the fine cells it covers and divides by the
// coarse centre, giving the coarse forcing (zero on coarse land).
__global__ void k_residual_restrict32(LevelArgs F, const float *__restrict__ f, const float *__restrict__ x,
                                      LevelArgs C, float *__restrict__ fcoarse) {
    const int ca = C.across, cd = C.down, fa = F.across, fd = F.down;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ca * cd) return;
    const int crow = (int)(t / ca), ccol = (int)(t - (long)crow * ca);
    if (!C.water[t]) {
        fcoarse[t] = 0.0f;
        return;
    }
    const int cpc = fa / ca, rpc = fd / cd;
    const double share = 1.0 / (double)(cpc * rpc);
    const double cc = centre_at(C, crow, t);
    float acc = 0.0f;
    for (int dr = 0; dr < rpc; ++dr) {
        const int row = crow * rpc + dr;
        for (int dc = 0; dc < cpc; ++dc) {
            const int col = ccol * cpc + dc;
            const long cell = (long)row * fa + col;
            if (!F.water[cell]) continue;
            const int ce = col + 1 == fa ? 0 : col + 1;
            const int cw = col == 0 ? fa - 1 : col - 1;
            const float xn = row > 0 ? x[cell - fa] : 0.0f;
            const float xs = row + 1 < fd ? x[cell + fa] : 0.0f;
            const float4 wt = weights_at(F, row, cell);
            const float u = x[cell] - rb_update(wt, x[(long)row * fa + ce], x[(long)row * fa + cw], xn, xs, f[cell]);
            const float g = (float)(centre_at(F, row, cell) * share / cc);
            acc += u * g;
        }
    }
    fcoarse[t] = acc;
}

// The coarsest grid solved directly: x = M f over its water cells, M = -(I - W)^-1 stored
// column-major (Mt[j * n + i] = M[i][j]) so that threads of one column read consecutive entries.
__global__ __launch_bounds__(256) void k_coarsest_direct(const int *__restrict__ idx, int n,
                                                         const float *__restrict__ Mt, const float *__restrict__ f,
                                                         float *__restrict__ x, long cells) {
    extern __shared__ float fs[];
    for (int j = threadIdx.x; j < n; j += blockDim.x) fs[j] = f[idx[j]];
    for (long c = threadIdx.x; c < cells; c += blockDim.x) x[c] = 0.0f;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        float acc = 0.0f;
        for (int j = 0; j < n; ++j) acc = fmaf(Mt[(long)j * n + i], fs[j], acc);
        x[idx[i]] = acc;
    }
}

// x64 <- (double) x32, and the reverse, elementwise.
__global__ void k_f32_to_f64(const float *__restrict__ a, double *__restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
__global__ void k_f64_to_f32(const double *__restrict__ a, float *__restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

}  // namespace cg
