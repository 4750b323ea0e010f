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

// Lane i receives lane i-1's value (wave_shr:1) / lane i+1's (wave_shl:1); the edge lane receives
// 0 (bound_ctrl), which only a tile's halo ever reads. One DPP move, with no copy of an old value.
__device__ __forceinline__ float lane_from_below(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_from_above(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
// v where bit `bit` of m is set, else +0 (one sign-extending bit extract and one AND).
__device__ __forceinline__ float keep_if(float v, uint32_t m, int bit) {
    const int sel = __builtin_amdgcn_sbfe((int)m, bit, 1);
    return __builtin_bit_cast(float, __builtin_bit_cast(int, v) & sel);
}

// The object `byte` bytes past a uniform base, the offset a 32-bit value (every grid here is far
// below 4 GiB), which the global_load/store "saddr" form takes directly, rather than a 64-bit
// address formed per lane.
template <typename T>
__device__ __forceinline__ const T &at_byte(const void *base, unsigned byte) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte);
}
template <typename T>
__device__ __forceinline__ T &at_byte(void *base, unsigned byte) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte);
}

// Largest magnitudes are reduced as the uint64 patterns of |v|: for non-negative doubles the
// order of the patterns is the order of the values, and a NaN's pattern exceeds all of them.
__device__ __forceinline__ unsigned long long abs_bits(double v) {
    return (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// A double moved between lanes by one DPP control on each 32-bit half (0x138 wave_shr:1: lane i
// receives lane i-1's; 0x130 wave_shl:1: lane i+1's). The edge lane receives 0.
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Reduction slots of a largest magnitude (see grid_max): kMaxSlots relaxed atomicMax targets, and
// the slots the next reduction will use, which the launch zeroes as it goes.
constexpr int kMaxSlots = 64;

struct MaxReduce {
    unsigned long long *out;
    unsigned long long *clear;
};

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The refinement's error estimate: the binary64 outer step's sums of the correction's squares
// (slots 0 .. kMaxSlots-1) and of the new answer's squares (the next kMaxSlots), each wave adding its
// sums into slot wave % kMaxSlots. As with the maxima, the slots start at zero and the step zeroes the
// group the next step accumulates into.
struct SumReduce {
    double *out;
    double *clear;
};

// Per instance: whether the refinement goes on, how many steps it has run, whether the step the
// gate let run computes the error estimate's sums, whether a learned solve stopped it for its
// stalling residual (fell), the sums of squared corrections of the two steps before (the
// contraction of the next one is measured against them; 0 where a step had no sums) and the
// residuals of the two steps before.
struct StepGate {
    int active;
    int steps;
    int want;
    int fell;
    double dprev2;
    double dprev2b;
    double res_prev;
    double res_prev2;
};

// Steps compute the estimate's sums only once the residual is within this factor of its threshold
// (the last few steps of a solve): before, the residual alone keeps the refinement going.
constexpr double kSumsAhead = 100.0;

// The stopping test of the binary64 refinement, run on the device before every step so that each
// instance stops at exactly the first step where it holds, however the host batches the steps: the
// previous step's residual max |u c| under thr[inst] (tol x max |f c|), and, once a correction has
// been applied, its estimated relative L2 error under etol: for a contraction rho per step,
// e_k = rho e_(k-1) and delta_k = e_(k-1) - e_k, so |e_k| = rho / (1 - rho) |delta_k|, with rho the
// larger of the last two ratios of successive correction norms (a contraction that varies from step
// to step, as a learned correction's does, is not underestimated). A residual that is not a number stops the instance too.
// Reads the previous step's groups and state `in`; the wave with `writer` set writes state `out`
// and the active flag the rest of the step reads.
struct GateArgs {
    const unsigned long long *res;  // instance 0's residual slots of the previous step
    const double *sums;             // instance 0's sums of the previous step (2 kMaxSlots)
    const StepGate *in;
    StepGate *out;
    int *active;
    int *want_sums;  // per instance: the step's outer kernel computes the sums
    const double *thr;
    const double *etol;       // per instance; <= 0: the residual alone decides
    const double *min_steps;  // per instance: no stop before this many steps of this refinement
    // A learned step's fallback: from the fourth step on, a residual above fall_rate squared times
    // the one two steps before (a rate of fall_rate a step over two steps) stops the instance with
    // `fell` set (the host goes on from its answer with the plain cycle). fall_rate <= 0: never.
    double fall_rate;
    int res_stride, sums_stride;
};

__device__ bool gate_decide(const GateArgs &g, int inst, bool writer) {
    const int lane = threadIdx.x & 63;
    const StepGate s = g.in[inst];
    bool go = false, want = false, fell = s.fell != 0;
    double d2 = 0.0, res = 0.0, etol = 0.0;
    if (s.active) {
        const unsigned long long r = wave_max_u64(g.res[inst * g.res_stride + lane]);
        double x2 = 0.0;
        if (s.want) {
            d2 = wave_sum_f64(g.sums[inst * g.sums_stride + lane]);
            x2 = wave_sum_f64(g.sums[inst * g.sums_stride + kMaxSlots + lane]);
        }
        res = __longlong_as_double((long long)r);
        const double thr = g.thr[inst];
        etol = g.etol[inst];
        bool err_ok = true;
        if (etol > 0.0 && s.steps > 0) {
            if (!s.want) {
                err_ok = false;  // the step ran without the sums: no estimate yet
            } else if (x2 > 0.0) {
                double rho = s.dprev2 > 0.0 ? sqrt(d2 / s.dprev2) : 0.5;
                if (s.dprev2 > 0.0 && s.dprev2b > 0.0) {
                    const double before = sqrt(s.dprev2 / s.dprev2b);
                    rho = before > rho ? before : rho;
                }
                rho = rho < 0.95 ? rho : 0.95;
                err_ok = sqrt(d2 / x2) * (rho / (1.0 - rho)) <= etol;
            } else {
                err_ok = d2 <= 0.0;
            }
        }
        go = !(res != res || (res < thr && err_ok && s.steps >= (int)g.min_steps[inst]));
        if (go && g.fall_rate > 0.0 && s.steps >= 3 && !(res <= g.fall_rate * g.fall_rate * s.res_prev2)) {
            go = false;
            fell = true;
        }
        want = go && etol > 0.0 && res < kSumsAhead * thr;
    }
    if (writer && lane == 0) {
        StepGate o;
        o.active = go ? 1 : 0;
        o.steps = s.steps + (go ? 1 : 0);
        o.want = want ? 1 : 0;
        o.fell = fell ? 1 : 0;
        o.dprev2 = go ? (s.want ? d2 : 0.0) : s.dprev2;
        o.dprev2b = go ? s.dprev2 : s.dprev2b;
        o.res_prev = s.active ? res : s.res_prev;
        o.res_prev2 = s.active ? s.res_prev : s.res_prev2;
        g.out[inst] = o;
        g.active[inst] = o.active;
        g.want_sums[inst] = o.want;
    }
    return go;
}

// A refinement's start: `clear` (the residual slots and sums of every instance) zeroed, and the
// gate state the first step reads set to active with no steps, every active flag 1 and no sums.
// With `only` set, an instance whose only[i] is 0 starts stopped (a solve of some instances).
__global__ void k_check_init(StepGate *state, int *active, int *want, unsigned long long *clear, long nclear, int K,
                             const int *only) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nclear) clear[i] = 0ull;
    if (i < K) {
        const int on = only ? (only[i] != 0) : 1;
        StepGate g;
        g.active = on;
        g.steps = 0;
        g.want = 0;
        g.fell = 0;
        g.dprev2 = 0.0;
        g.dprev2b = 0.0;
        g.res_prev = 0.0;
        g.res_prev2 = 0.0;
        state[i] = g;
        active[i] = on;
        want[i] = 0;
    }
}

// The stopping test as a launch of its own (one wavefront per instance, blockIdx.y), for steps whose
// first kernel is not the finest grid's fused down leg.
__global__ __launch_bounds__(64) void k_step_gate(GateArgs g) { (void)gate_decide(g, blockIdx.y, true); }

// Binary32 update of one cell in the reference's order: ((e xE + w xW) + n xN) + s xS - f.
__device__ __forceinline__ float rb_update(float4 wt, float xe, float xw, float xn, float xs, float f) {
    return wt.x * xe + wt.y * xw + wt.z * xn + wt.w * xs - f;
}

// The same update by three fused multiply-adds, for the refined method's binary32 correction
// cycle: that cycle only has to reduce its residual by a factor (its rounding never reaches the
// binary64 fixed point), so it need not round as the reference does, and it saves three of eight
// operations per update. Every kernel of that cycle takes it, so its bits do not depend on tiling.
template <bool FUSED>
__device__ __forceinline__ float rb_step(float4 wt, float xe, float xw, float xn, float xs, float f) {
    if constexpr (FUSED) return __builtin_fmaf(wt.w, xs, __builtin_fmaf(wt.z, xn, __builtin_fmaf(wt.y, xw, wt.x * xe))) - f;
    else return rb_update(wt, xe, xw, xn, xs, f);
}

// ------------------------------------------------------------------------------------------------
// Fused red-black smoother, row-constant weights.
//
// One wavefront owns a tile of TH rows by IW = 128 - 2 HC columns plus a halo of HR rows and HC
// columns. Lane l holds the two adjacent columns 2l and 2l+1 of every tile row in registers
// (A = even column, B = odd column), so in each half-sweep every lane updates exactly one of its
// two cells per row: east/west neighbours are the lane's other register or its neighbour lane's
// (one cross-lane shuffle per row), north/south neighbours are the same lane's registers of the
// adjacent rows. NH half-sweeps (colour 0 first) run without any barrier or LDS; each invalidates
// one more cell of the halo, so after NH the interior is exact and is written back. Rows beyond a
// pole are zero and never change, like the reference's zero beyond the pole (each array the
// smoother reads carries one all-zero row after its last, which such rows load); columns wrap by
// modular loading. Reads xin, writes xout (the two must differ: neighbouring tiles read the halo
// this tile would otherwise overwrite).
//
// MODE selects what is fused around the smoothing, for a ladder whose next grid halves both axes:
//   TILE_PLAIN    : xout = NH half-sweeps of xin.
//   TILE_DOWN     : the cycle's down leg: from zero (xin unread), then the binary32 residual
//                   u = x - (((e xE + w xW) + n xN) + s xS - f) of the interior, restricted as
//                   sum u * g(row) over each 2x2 block into the next grid's forcing fc (zero on its
//                   land), g(row) = centre(row) * 1/4 / centre_coarse(row / 2) in binary32.
//   TILE_UP       : the cycle's up leg: before smoothing, the coarse correction xc is added to the
//                   water cells bilinearly in prolongAdd's binary32 arithmetic (OC:354-392).
//   TILE_UP_OUTER : TILE_UP on the finest grid with the binary64 outer step fused after it: one more
//                   halo row each side, so the smoothed correction delta is exact on the interior and
//                   its ring; then, as k_outer_residual does, x64n = x64 + delta and
//                   u = x64n - (sum w x64n - f0) in binary64, written as the next cycle's binary32
//                   forcing fcn, with the largest |u c| reduced per wave into red_out. delta itself is
//                   never written.
// ------------------------------------------------------------------------------------------------
enum TileMode { TILE_PLAIN = 0, TILE_DOWN = 1, TILE_UP = 2, TILE_UP_OUTER = 3 };


// Workgroups are dealt to the MI355X's 8 XCDs round-robin by workgroup id, and each XCD has its own
// L2. Renumbering so that XCD k runs the k-th contiguous eighth of the logical blocks keeps
// neighbouring tiles, which read each other's halo rows, on one L2. A bijection on [0, gridDim.x):
// the last gridDim.x mod 8 blocks keep their ids.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int full = nb - nb % kXcds;
    if (b >= full) return b;
    return (b % kXcds) * (full / kXcds) + b / kXcds;
}

struct TileExtra {
    // TILE_DOWN
    const uint8_t *cwater;  // next grid's mask
    const float *g;         // per fine row restriction factor
    float *fc;              // next grid's forcing (out)
    // TILE_UP
    const float *xc;        // next grid's values
    int ca, cd;             // next grid's shape
    float beyond;           // -1 for a wall at the poles
    // Instances: blockIdx.y selects one; each instance's grids sit fs (this grid) and cs (the next
    // grid) floats after the previous instance's; an instance whose active flag is 0 is skipped.
    long fs, cs;
    const int *active;
    // TILE_UP_OUTER: the binary64 answer in and out (instances s64 doubles apart), the finest forcing
    // of the binary64 problem and the next cycle's forcing out (fs apart), the maxima's slots
    // (red_stride apart) and, when sums.out is set, the error estimate's sums (sums_stride apart).
    const double *x64;
    double *x64n;
    const float *f0;
    float *fcn;
    MaxReduce red;
    long s64;
    int red_stride;
    SumReduce sums;
    int sums_stride;
    const int *want_sums;  // when set: per instance, whether this step computes the sums
    // TILE_DOWN of the finest grid in a refinement step: the stopping test (gate.in set) instead of
    // the active flag; the kernel writes that flag for the rest of the step. With late set (one
    // instance, whose other kernels never read the flag: what they write is the next step's
    // scratch) only tile 0's wave runs the test, after its own work, and TILE_UP_OUTER reads the
    // flag with its loads and holds back only its outputs.
    GateArgs gate;
    int late;
};

template <int NH, int MODE>
struct TileGeom {
    static constexpr int HR = NH + ((MODE == TILE_DOWN || MODE == TILE_UP_OUTER) ? 1 : 0);
    static constexpr int HC = MODE == TILE_PLAIN ? ((NH + 1) & ~1) : ((NH + 2) & ~1);
    static constexpr int IW = 128 - 2 * HC;
};

template <int NH, int TH, bool ZERO_START, int MODE>
__global__ __launch_bounds__(256) void k_relax_tile(LevelArgs L, const float *__restrict__ f,
                                                    const float *__restrict__ xin, float *__restrict__ xout,
                                                    TileExtra ex) {
#define TILE_INST blockIdx.y
#define TILE_BLOCK() xcd_block()
#include "tile_body.inc"
#undef TILE_INST
#undef TILE_BLOCK
}

// One leg of several plans' cycles in one launch (the grids of uneven worlds at the same depth of
// their ladders, each with its own instances): job j owns the `count` workgroups from `first` on,
// `per_inst` of them per instance. first and per_inst are multiples of the XCD count, so workgroup
// b's XCD (b mod 8) is its local index's, and each XCD works on one contiguous band of every
// instance's tiles as in k_relax_tile.
struct TileJob {
    LevelArgs L;
    const float *f;
    const float *xin;
    float *xout;
    TileExtra ex;
    int first;
    int per_inst;
    int count;
};

template <int NH, int TH, bool ZERO_START, int MODE>
__global__ __launch_bounds__(256) void k_relax_tile_jobs(const TileJob *__restrict__ jobs, int njobs) {
    int j = 0;
    while (j + 1 < njobs && (int)blockIdx.x >= jobs[j + 1].first) ++j;
    const int local = (int)blockIdx.x - jobs[j].first;
    if (local < 0 || local >= jobs[j].count) return;
    const int job_inst = local / jobs[j].per_inst, lb = local % jobs[j].per_inst;
    const int job_block = (lb % kXcds) * (jobs[j].per_inst / kXcds) + lb / kXcds;
    LevelArgs L = jobs[j].L;
    const float *__restrict__ f = jobs[j].f;
    const float *__restrict__ xin = jobs[j].xin;
    float *__restrict__ xout = jobs[j].xout;
    TileExtra ex = jobs[j].ex;
#define TILE_INST job_inst
#define TILE_BLOCK() job_block
#include "tile_body.inc"
#undef TILE_INST
#undef TILE_BLOCK
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

// A grid-wide maximum: each block reduces in registers and LDS, then its thread 0 does one relaxed
// atomicMax into out[block % kMaxSlots] (no fences, so no L2 write-back per block; 64 slots keep
// the atomics uncontended). The reader takes the largest slot. The slots must be zero when the
// kernel starts; the blocks together zero all of `clear` (the slots the next reduction will use),
// which a later launch on the same stream then sees, so graph replays need no memset.
__device__ void grid_max(unsigned long long v, const MaxReduce &red) {
    __shared__ unsigned long long sh[16];
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    const int nb = gridDim.x * gridDim.y;
    if (threadIdx.x == 0 && red.clear)
        for (int i = b; i < kMaxSlots; i += nb) red.clear[i] = 0ull;
    v = wave_max_u64(v);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0;
        for (int i = 0; i < nw; ++i) m = sh[i] > m ? sh[i] : m;
        if (m) atomicMax(red.out + (b & (kMaxSlots - 1)), m);
    }
}

// The 2D launch of the per-cell binary64 kernels: blockIdx.x covers 256 columns, blockIdx.y
// covers kRowsPerBlock rows, so no thread divides.
constexpr int kRowsPerBlock = 4;

// Binary64 residual (OC:292-315): r = (x - (sum w x - f)) * centre on water, 0 on land; and its
// largest magnitude into red.out when red.out is given.
__global__ __launch_bounds__(256) void k_residual64(LevelArgs L, const float *__restrict__ f,
                                                    const float *__restrict__ x, double *__restrict__ r,
                                                    MaxReduce red) {
    const int across = L.across;
    const int col = blockIdx.x * 256 + threadIdx.x;
    unsigned long long m = 0;
    if (col < across) {
        const int ce = col + 1 == across ? 0 : col + 1;
        const int cw = col == 0 ? across - 1 : col - 1;
#pragma unroll
        for (int k = 0; k < kRowsPerBlock; ++k) {
            const int row = blockIdx.y * kRowsPerBlock + k;
            if (row >= L.down) break;
            const long cell = (long)row * across + col;
            double out = 0.0;
            if (L.water[cell]) {
                const double vn = row > 0 ? (double)x[cell - across] : 0.0;
                const double vs = row + 1 < L.down ? (double)x[cell + across] : 0.0;
                const float4 wt = weights_at(L, row, cell);
                const double bal = (double)wt.x * (double)x[(long)row * across + ce] +
                                   (double)wt.y * (double)x[(long)row * across + cw] + (double)wt.z * vn +
                                   (double)wt.w * vs - (double)f[cell];
                out = ((double)x[cell] - bal) * centre_at(L, row, cell);
            }
            if (r) r[cell] = out;
            const unsigned long long b = abs_bits(out);
            m = b > m ? b : m;
        }
    }
    if (red.out) grid_max(m, red);
}

// The largest magnitude of f * centre (balanceOf then largest, OC:325-327, 318-322).
__global__ __launch_bounds__(256) void k_balance_max(LevelArgs L, const float *__restrict__ f, MaxReduce red) {
    const int across = L.across;
    const int col = blockIdx.x * 256 + threadIdx.x;
    unsigned long long m = 0;
    if (col < across) {
#pragma unroll
        for (int k = 0; k < kRowsPerBlock; ++k) {
            const int row = blockIdx.y * kRowsPerBlock + k;
            if (row >= L.down) break;
            const long cell = (long)row * across + col;
            const unsigned long long b = abs_bits((double)f[cell] * centre_at(L, row, cell));
            m = b > m ? b : m;
        }
    }
    grid_max(m, red);
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
// dst[i] = src[i] for i < n and 0 for n <= i < total: a grid copied into a buffer that carries the
// smoother's all-zero row after the last.
__global__ void k_stage_grid(const float *__restrict__ src, float *__restrict__ dst, long n, long total) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) dst[i] = i < n ? src[i] : 0.0f;
}
__global__ void k_copy_f32(const float *__restrict__ src, float *__restrict__ dst, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
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
// |u * centre| into red. Land values are zero in x_new.
//
// Thread t of a block owns the column pair (2c, 2c+1), c = blockIdx.x * 256 + t, over the
// kRowsPerBlock rows of blockIdx.y, and loads a window of kRowsPerBlock + 2 rows once (north and
// south neighbours come from the window); east/west neighbours come from the adjacent lanes by DPP,
// and only a wave's edge lanes (and the pair at the wrap) load theirs. Needs an even width.
// Instances of the outer step: blockIdx.z selects one; its binary32 grids sit s32 floats, its
// binary64 grids s64 doubles and its reduction slots red_stride slots after the previous one's; an
// instance whose active flag is 0 is skipped.
struct OuterBatch {
    long s32, s64;
    int red_stride;
    const int *active;
    SumReduce sums;  // with out set: the error estimate's sums, as TILE_UP_OUTER adds them
    int sums_stride;
    const int *want_sums;  // when set: per instance, whether this step computes the sums
};

// A grid-wide sum of each of two values into the slots of `sums` (block b adds into slot b %
// kMaxSlots and kMaxSlots + b % kMaxSlots); the blocks together zero sums.clear.
__device__ void grid_sums(double a, double b2, const SumReduce &sums) {
    __shared__ double sh[2][16];
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    const int nb = gridDim.x * gridDim.y;
    if (threadIdx.x == 0 && sums.clear)
        for (int i = b; i < 2 * kMaxSlots; i += nb) sums.clear[i] = 0.0;
    a = wave_sum_f64(a);
    b2 = wave_sum_f64(b2);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = a;
        sh[1][w] = b2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < nw; ++i) {
            sa += sh[0][i];
            sb += sh[1][i];
        }
        if (sa != 0.0) atomicAdd(sums.out + (b & (kMaxSlots - 1)), sa);
        if (sb != 0.0) atomicAdd(sums.out + kMaxSlots + (b & (kMaxSlots - 1)), sb);
    }
}

__global__ __launch_bounds__(256) void k_outer_residual(LevelArgs L, const float *__restrict__ f,
                                                        const double *__restrict__ x, const float *__restrict__ delta,
                                                        double *__restrict__ xnew, float *__restrict__ fc,
                                                        MaxReduce red, OuterBatch ob) {
    const int inst = blockIdx.z;
    if (ob.active && !ob.active[inst]) return;
    if (inst) {
        f += inst * ob.s32;
        x += inst * ob.s64;
        if (delta) delta += inst * ob.s32;
        xnew += inst * ob.s64;
        fc += inst * ob.s32;
        red.out += inst * ob.red_stride;
        if (red.clear) red.clear += inst * ob.red_stride;
        if (ob.sums.out) {
            ob.sums.out += inst * ob.sums_stride;
            if (ob.sums.clear) ob.sums.clear += inst * ob.sums_stride;
        }
    }
    const int across = L.across, down = L.down, half = across >> 1;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const bool mine = c < half;
    const int cc = mine ? c : half - 1;  // idle lanes shadow the last pair, so DPP sources stay in range
    const int col = 2 * cc;
    const int r0 = blockIdx.y * kRowsPerBlock;
    constexpr int W = kRowsPerBlock + 2;
    double vx[W], vy[W];  // x + delta of the pair, rows r0-1 .. r0+kRowsPerBlock
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const int r = r0 - 1 + i;
        vx[i] = vy[i] = 0.0;
        if (r >= 0 && r < down) {
            const long idx = (long)r * across + col;
            const double2 xv = *reinterpret_cast<const double2 *>(x + idx);
            vx[i] = xv.x;
            vy[i] = xv.y;
            if (delta) {
                const float2 dv = *reinterpret_cast<const float2 *>(delta + idx);
                vx[i] += (double)dv.x;
                vy[i] += (double)dv.y;
            }
        }
    }
    // Columns of the pair's west and east neighbours, and whether the adjacent lane holds them.
    const int cw = col == 0 ? across - 1 : col - 1;
    const int ce = col + 2 == across ? 0 : col + 2;
    const bool lane_w = lane > 0 && col > 0;
    const bool lane_e = lane < 63 && c + 1 < half;
    unsigned long long m = 0;
    double d2 = 0.0, x2 = 0.0;
#pragma unroll
    for (int k = 0; k < kRowsPerBlock; ++k) {
        const int i = k + 1;
        const int row = r0 + k;
        // Every lane runs the shifts (DPP reads the neighbour lane's registers).
        double west = dpp_double<0x138>(vy[i]);  // lane - 1's odd column
        double east = dpp_double<0x130>(vx[i]);  // lane + 1's even column
        if (!mine || row >= down) continue;
        const long idx = (long)row * across + col;
        if (!lane_w) {
            west = x[(long)row * across + cw];
            if (delta) west += (double)delta[(long)row * across + cw];
        }
        if (!lane_e) {
            east = x[(long)row * across + ce];
            if (delta) east += (double)delta[(long)row * across + ce];
        }
        const float2 fv = *reinterpret_cast<const float2 *>(f + idx);
        const uint16_t wm = *reinterpret_cast<const uint16_t *>(L.water + idx);
        const float4 wt = weights_at(L, row, idx);
        const double cen = centre_at(L, row, idx);
        const double vn_x = row > 0 ? vx[i - 1] : 0.0, vn_y = row > 0 ? vy[i - 1] : 0.0;
        const double vs_x = row + 1 < down ? vx[i + 1] : 0.0, vs_y = row + 1 < down ? vy[i + 1] : 0.0;
        double ux = 0.0, uy = 0.0, ox = 0.0, oy = 0.0;
        if (wm & 0xff) {
            const double bal = (double)wt.x * vy[i] + (double)wt.y * west + (double)wt.z * vn_x + (double)wt.w * vs_x -
                               (double)fv.x;
            ox = vx[i];
            ux = ox - bal;
            const unsigned long long b = abs_bits(ux * cen);
            m = b > m ? b : m;
            x2 += ox * ox;
        }
        if (wm >> 8) {
            const double bal = (double)wt.x * east + (double)wt.y * vx[i] + (double)wt.z * vn_y + (double)wt.w * vs_y -
                               (double)fv.y;
            oy = vy[i];
            uy = oy - bal;
            const unsigned long long b = abs_bits(uy * cen);
            m = b > m ? b : m;
            x2 += oy * oy;
        }
        if (delta && (wm & 0xffff)) {
            const float2 dv = *reinterpret_cast<const float2 *>(delta + idx);
            if (wm & 0xff) d2 += (double)dv.x * (double)dv.x;
            if (wm >> 8) d2 += (double)dv.y * (double)dv.y;
        }
        *reinterpret_cast<double2 *>(xnew + idx) = make_double2(ox, oy);
        *reinterpret_cast<float2 *>(fc + idx) = make_float2((float)ux, (float)uy);
    }
    grid_max(m, red);
    if (ob.sums.out) {
        const bool on = !ob.want_sums || ob.want_sums[inst];
        grid_sums(on ? d2 : 0.0, on ? x2 : 0.0, ob.sums);
    }
}

// k_outer_residual from x = 0 with no delta, as the first outer step of a solve from zero: every
// value and neighbour is +0, so only the forcing, the mask and the weights are read; the
// arithmetic is the same expression with the zeros written in (x64 new = 0, u = 0 - bal on water,
// 0 on land), so the results are the same bits. Writes xnew = 0 (the binary64 answer of the next
// parity), fc, and the largest |u c| into red.out (not cleared: the solve zeroes it first).
__global__ __launch_bounds__(256) void k_first_outer(LevelArgs L, const float *__restrict__ f,
                                                     double *__restrict__ xnew, float *__restrict__ fc,
                                                     MaxReduce red, OuterBatch ob) {
    const int inst = blockIdx.z;
    if (inst) {
        f += inst * ob.s32;
        xnew += inst * ob.s64;
        fc += inst * ob.s32;
        red.out += inst * ob.red_stride;
    }
    const int across = L.across, down = L.down, half = across >> 1;
    const int c = blockIdx.x * 256 + threadIdx.x;
    unsigned long long m = 0;
    if (c < half) {
        const int col = 2 * c;
#pragma unroll
        for (int k = 0; k < kRowsPerBlock; ++k) {
            const int row = blockIdx.y * kRowsPerBlock + k;
            if (row >= down) break;
            const long idx = (long)row * across + col;
            const float2 fv = *reinterpret_cast<const float2 *>(f + idx);
            const uint16_t wm = *reinterpret_cast<const uint16_t *>(L.water + idx);
            const float4 wt = weights_at(L, row, idx);
            const double cen = centre_at(L, row, idx);
            const double z = 0.0;
            double ux = 0.0, uy = 0.0;
            if (wm & 0xff) {
                const double bal = (double)wt.x * z + (double)wt.y * z + (double)wt.z * z + (double)wt.w * z - (double)fv.x;
                ux = z - bal;
                const unsigned long long b = abs_bits(ux * cen);
                m = b > m ? b : m;
            }
            if (wm >> 8) {
                const double bal = (double)wt.x * z + (double)wt.y * z + (double)wt.z * z + (double)wt.w * z - (double)fv.y;
                uy = z - bal;
                const unsigned long long b = abs_bits(uy * cen);
                m = b > m ? b : m;
            }
            *reinterpret_cast<double2 *>(xnew + idx) = make_double2(0.0, 0.0);
            *reinterpret_cast<float2 *>(fc + idx) = make_float2((float)ux, (float)uy);
        }
    }
    grid_max(m, red);
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

// ------------------------------------------------------------------------------------------------
// The bottom of the binary32 correction cycle in one workgroup: every level from `first` down to
// the coarsest (each small enough that all of them fit in LDS together), relaxed, restricted,
// solved directly at the coarsest, prolonged and relaxed again, with __syncthreads between
// phases. Same arithmetic as the multi-workgroup kernels: binary32 updates in the reference's
// order, binary32 residual restricted with the per-row factor g, prolongAdd's bilinear weights.
// ------------------------------------------------------------------------------------------------
constexpr int kBottomMaxLevels = 8;

struct BottomLevel {
    int across, down;
    // LDS byte offsets: values and forcing, then (in the static region) the mask, the row weights
    // and the restriction factor per row to the next grid (0 at the coarsest).
    int off_x, off_f, off_w, off_rw, off_g;
};

struct BottomArgs {
    int nlev;
    BottomLevel lv[kBottomMaxLevels];
    const float *f_in;  // forcing of the first level (global)
    float *x_out;       // values of the first level (global)
    // Instances: workgroup b solves instance b, whose f_in and x_out sit b strides further; an
    // instance whose active flag is 0 is skipped.
    long f_stride, x_stride;
    const int *active;
    // The static LDS region (masks, row weights, restriction factors of every grid) as uint32 words,
    // and where it starts in LDS.
    const uint32_t *blob;
    int blob_words, static_base;
    const int *idx;     // coarsest water cells, and M = -(I - W)^-1 column-major (nw x nw)
    const float *Mt;
    int nw;
    int coarsest_passes;  // relaxation passes at the coarsest when there is no inverse
    float beyond;
};

// The bottom's levels are at most kBottomRows rows by 2 * kBottomPairs columns, so that the
// workgroup's 1024 threads, seen as 32 x 32 (tx = column pair, ty = row), give every phase at most
// one cell per thread per colour (two for the prolongation): no loops, no divisions, every LDS
// read of a phase issued before its one write. The workgroup has one CU's four SIMDs, so the
// bottom is bound by its instruction count; keeping each phase to a few instructions per cell is
// what keeps it short.
constexpr int kBottomRows = 32;
constexpr int kBottomPairs = 32;
// The register-held coarsest inverse: with nw <= 128 water cells each of the 1024 threads holds at
// most 16 entries (parts = 1024 / nw >= 8 groups of the j-sum).
constexpr int kBottomDirectRegs = 16;
constexpr int kBottomDirectRegsWater = 128;

struct BottomView {
    int across, down;
    float *x, *f;
    uint8_t *w;
    const float4 *rw;
    const float *g;
};

// `passes` red-black passes of a bottom grid in LDS. Thread (tx, ty) owns the row-ty cells of
// columns 2 tx and 2 tx + 1, one of each colour; the LDS indices of both cells and their neighbours
// are formed once for all the passes, and every read of a half-sweep is issued before its one
// write. Branch-free: rows beyond a pole and land cells are selected to +0, as the reference's
// zero beyond the pole and its land cells, which never change, hold +0.
__device__ __forceinline__ void lds_relax_passes(const BottomView &L, int passes) {
    const int tx = threadIdx.x & 31, row = threadIdx.x >> 5;
    const int across = L.across;
    const bool mine = row < L.down && tx < (across >> 1);
    const bool has_n = row > 0, has_s = row + 1 < L.down;
    int cell[2], ie[2], iw[2], in[2], is[2];
    bool wet[2] = {false, false};
#pragma unroll
    for (int colour = 0; colour <= 1; ++colour) {
        const int col = 2 * tx + ((colour + row) & 1);
        const int c = mine ? row * across + col : 0;
        const int ce = col + 1 == across ? 0 : col + 1;
        const int cw = col == 0 ? across - 1 : col - 1;
        cell[colour] = c;
        ie[colour] = mine ? row * across + ce : 0;
        iw[colour] = mine ? row * across + cw : 0;
        in[colour] = mine && has_n ? c - across : c;
        is[colour] = mine && has_s ? c + across : c;
        wet[colour] = mine && L.w[c] != 0;
    }
    const float4 wt = L.rw[mine ? row : 0];
    for (int p = 0; p < passes; ++p) {
#pragma unroll
        for (int colour = 0; colour <= 1; ++colour) {
            const float xe = L.x[ie[colour]], xw = L.x[iw[colour]];
            float xn = L.x[in[colour]], xs = L.x[is[colour]];
            const float fv = L.f[cell[colour]];
            xn = has_n ? xn : 0.0f;
            xs = has_s ? xs : 0.0f;
            const float v = rb_update(wt, xe, xw, xn, xs, fv);
            if (mine) L.x[cell[colour]] = wet[colour] ? v : 0.0f;
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(1024) void k_bottom(BottomArgs a) {
    const int inst = blockIdx.x;
    if (a.active && !a.active[inst]) return;
    const float *f_in = a.f_in + inst * a.f_stride;
    float *x_out = a.x_out + inst * a.x_stride;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ float partial[1024];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    auto view = [&](int l) {
        BottomView v;
        v.across = a.lv[l].across;
        v.down = a.lv[l].down;
        v.x = reinterpret_cast<float *>(smem + a.lv[l].off_x);
        v.f = reinterpret_cast<float *>(smem + a.lv[l].off_f);
        v.w = reinterpret_cast<uint8_t *>(smem + a.lv[l].off_w);
        v.rw = reinterpret_cast<const float4 *>(smem + a.lv[l].off_rw);
        v.g = reinterpret_cast<const float *>(smem + a.lv[l].off_g);
        return v;
    };
    // The coarsest inverse's entries this thread multiplies, and their cells, loaded now so that
    // their latency passes under the down leg (at most kBottomDirectRegs each, for nw of at most
    // kBottomDirectRegsWater water cells; larger coarsest grids read them at the solve).
    const int nw = a.nw;
    const int parts = nw > 0 && nw <= (int)blockDim.x ? (int)blockDim.x / nw : 1;
    const int chunk = (int)blockDim.x / parts;
    const bool direct_regs = nw > 0 && nw <= kBottomDirectRegsWater;
    const int di = (int)threadIdx.x % chunk, dp = (int)threadIdx.x / chunk;
    float mreg[kBottomDirectRegs];
    int jreg[kBottomDirectRegs];
    int oreg = 0;
    if (direct_regs) {
#pragma unroll
        for (int k = 0; k < kBottomDirectRegs; ++k) {
            const int j = dp + k * parts;
            const bool ok = di < nw && dp < parts && j < nw;
            mreg[k] = ok ? a.Mt[(unsigned)j * (unsigned)nw + (unsigned)di] : 0.0f;
            jreg[k] = ok ? a.idx[j] : 0;
        }
        oreg = di < nw ? a.idx[di] : 0;
    }
    // The grids' masks, row weights and restriction factors never change: the plan keeps them as
    // one blob laid out as this LDS region, copied in one coalesced pass (one load per thread).
    {
        uint32_t *dst = reinterpret_cast<uint32_t *>(smem + a.static_base);
        for (int c = threadIdx.x; c < a.blob_words; c += blockDim.x) dst[c] = a.blob[c];
    }
    for (int l = 0; l < a.nlev; ++l) {
        const int cells = a.lv[l].across * a.lv[l].down;
        float *x = reinterpret_cast<float *>(smem + a.lv[l].off_x);
        for (int c = threadIdx.x; c < cells; c += blockDim.x) x[c] = 0.0f;
    }
    {
        // Pairs (the width is even), so a 32 x 64 grid arrives in one load per thread.
        const int cells = a.lv[0].across * a.lv[0].down;
        float2 *f = reinterpret_cast<float2 *>(smem + a.lv[0].off_f);
        const float2 *src = reinterpret_cast<const float2 *>(f_in);
        for (int c = threadIdx.x; 2 * c < cells; c += blockDim.x) f[c] = src[c];
    }
    __syncthreads();
    // Down: relax twice, then the binary32 residual restricted with g (k_residual_restrict32's sums).
    for (int l = 0; l + 1 < a.nlev; ++l) {
        const BottomView L = view(l), C = view(l + 1);
        lds_relax_passes(L, 2);
        // Each grid halves both axes into the next (the plan checks it): 2 x 2 blocks.
        // Thread (tx, ty) forms coarse cell (ty, tx) from its 2 x 2 block: every LDS read issued
        // first, then the sums in the fixed order (dr, dc), a land fine cell adding nothing and a
        // land coarse cell taking +0, as k_residual_restrict32 does.
        if (ty < C.down && tx < C.across) {
            const int t = ty * C.across + tx;
            const bool cwet = C.w[t] != 0;
            float u[2][2];
            bool wet[2][2];
            float g[2];
#pragma unroll
            for (int dr = 0; dr < 2; ++dr) {
                const int row = 2 * ty + dr;
                const float4 wt = L.rw[row];
                g[dr] = L.g[row];
                const bool has_n = row > 0, has_s = row + 1 < L.down;
#pragma unroll
                for (int dc = 0; dc < 2; ++dc) {
                    const int col = 2 * tx + dc;
                    const int cell = row * L.across + col;
                    const int ce = col + 1 == L.across ? 0 : col + 1;
                    const int cw = col == 0 ? L.across - 1 : col - 1;
                    const float xn0 = L.x[has_n ? cell - L.across : cell];
                    const float xs0 = L.x[has_s ? cell + L.across : cell];
                    const float xn = has_n ? xn0 : 0.0f, xs = has_s ? xs0 : 0.0f;
                    wet[dr][dc] = L.w[cell] != 0;
                    // Colour 1 (row + col odd) was the last relaxed, from these same neighbours:
                    // its residual is exactly x - x.
                    u[dr][dc] = ((dr + dc) & 1)
                                    ? L.x[cell] - L.x[cell]
                                    : L.x[cell] - rb_update(wt, L.x[row * L.across + ce], L.x[row * L.across + cw],
                                                            xn, xs, L.f[cell]);
                }
            }
            float acc = 0.0f;
#pragma unroll
            for (int dr = 0; dr < 2; ++dr)
#pragma unroll
                for (int dc = 0; dc < 2; ++dc) {
                    const float sum = acc + u[dr][dc] * g[dr];
                    acc = wet[dr][dc] ? sum : acc;
                }
            C.f[t] = cwet ? acc : 0.0f;
        }
        __syncthreads();
    }
    // Coarsest: x = M f with the j-sum split over `parts` thread groups, then summed in LDS.
    {
        const BottomView C = view(a.nlev - 1);
        if (direct_regs) {
            // The same sums in the same order as below, from registers.
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < kBottomDirectRegs; ++k)
                if (dp + k * parts < nw) acc = fmaf(mreg[k], C.f[jreg[k]], acc);
            partial[threadIdx.x] = acc;
            __syncthreads();
            if (dp == 0 && di < nw) {
                float sum = 0.0f;
                for (int q = 0; q < parts; ++q) sum += partial[q * chunk + di];
                C.x[oreg] = sum;
            }
            __syncthreads();
        } else if (nw > 0) {
            for (int base = 0; base < nw; base += chunk) {
                const int i = base + (int)threadIdx.x % chunk;
                const int p = (int)threadIdx.x / chunk;
                float acc = 0.0f;
                if (i < nw && p < parts)
                    for (int j = p; j < nw; j += parts) acc = fmaf(a.Mt[(long)j * nw + i], C.f[a.idx[j]], acc);
                partial[threadIdx.x] = acc;
                __syncthreads();
                if (p == 0 && i < nw) {
                    float sum = 0.0f;
                    for (int q = 0; q < parts; ++q) sum += partial[q * chunk + (i - base)];
                    C.x[a.idx[i]] = sum;
                }
                __syncthreads();
            }
        } else {
            lds_relax_passes(C, a.coarsest_passes);
        }
        __syncthreads();
    }
    // Up: prolongAdd's bilinear correction (OC:354-392) on water, then relax twice. Thread (tx, ty)
    // owns fine cells (ty, 2 tx) and (ty, 2 tx + 1).
    for (int l = a.nlev - 2; l >= 0; --l) {
        const BottomView L = view(l), C = view(l + 1);
        const int ca = C.across, cd = C.down;
        const int row = ty;
        const bool mine = row < L.down && tx < (L.across >> 1);
        if (mine) {
            // prolongAdd's positions with 2 fine cells per coarse one along each axis (OC:369-388):
            // fine row r lies between coarse rows above = floor((r + 0.5) / 2 - 0.5) = (r - 1) >> 1
            // and above + 1 with weight db = 0.75 (r even) or 0.25 (r odd) on the lower one, and
            // likewise for columns; these are the reference's values exactly (dividing by 2.0f is
            // exact). Beyond a pole the edge row times `beyond`; in range the factor is 1, and
            // 1 * v is v exactly. Each thread reads only coarse values and its own fine cells, so
            // it writes at once.
            const int above = (row - 1) >> 1;
            const float db = (row & 1) ? 0.25f : 0.75f;
            const int ra = above < 0 ? 0 : above, rb = above + 1 >= cd ? cd - 1 : above + 1;
            const float fa = above < 0 ? a.beyond : 1.0f, fb = above + 1 >= cd ? a.beyond : 1.0f;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int col = 2 * tx + k;
                const int cell = row * L.across + col;
                const int left = tx - 1 + k;  // (col - 1) >> 1
                const float ab = k ? 0.25f : 0.75f;
                const int wc = left < 0 ? left + ca : left, ec = left + 1 == ca ? 0 : left + 1;
                const float top = (fa * C.x[ra * ca + wc]) * (1.0f - ab) + (fa * C.x[ra * ca + ec]) * ab;
                const float bottom = (fb * C.x[rb * ca + wc]) * (1.0f - ab) + (fb * C.x[rb * ca + ec]) * ab;
                const float v = L.x[cell] + (top * (1.0f - db) + bottom * db);
                if (L.w[cell]) L.x[cell] = v;
            }
        }
        __syncthreads();
        lds_relax_passes(L, 2);
    }
    {
        const int cells = a.lv[0].across * a.lv[0].down;
        const float2 *x = reinterpret_cast<const float2 *>(smem + a.lv[0].off_x);
        float2 *dst = reinterpret_cast<float2 *>(x_out);
        for (int c = threadIdx.x; 2 * c < cells; c += blockDim.x) dst[c] = x[c];
    }
}

// ------------------------------------------------------------------------------------------------
// The direct bottom: x = M f on a grid's water cells, M = -(I - W)^-1 dense (row-major, n x n,
// binary32, n <= kDirectMaxN), as a chip-wide matrix-vector product: one wavefront per output.
// Each workgroup gathers the water forcing of its instances into LDS; wave i sums row i of M
// against each, lane l taking columns l, l + 64, ... in order with fmaf (each step a coalesced
// 256-byte read of the row), then the lanes' sums meet in a fixed butterfly and lane 0 writes x at
// the cell. No atomics and no order that depends on timing, so the result is the same every launch
// and for every instance count. Land cells of x are not written: they hold +0.
// ------------------------------------------------------------------------------------------------
constexpr int kDirectMaxN = 2048;
constexpr int kDirectWaves = 16;  // outputs (wavefronts) per workgroup

struct DirectArgs {
    const int *idx;      // water cells
    const double *M;     // row-major, binary64
    int n;
    int K;               // instances
    const float *f;      // forcing of the grid (instances f_stride apart)
    float *x;            // values of the grid (instances x_stride apart)
    long f_stride, x_stride;
    const int *active;
};

// Workgroup (bx, by): output rows bx*kDirectWaves .. +kDirectWaves for instances
// by*NI .. +NI, so a row of M is read once per NI instances (1 for a plan of one instance, whose
// forcing then takes a quarter of the LDS).
template <int NI>
__device__ __forceinline__ void direct_bottom_body(const DirectArgs &a, const int bx, const int by) {
    constexpr int kDirectInst = NI;
    const int n = a.n;
    const int c0 = by * kDirectInst;
    const int ni = a.K - c0 < kDirectInst ? a.K - c0 : kDirectInst;
    bool on[kDirectInst];
    bool any = false;
#pragma unroll
    for (int c = 0; c < kDirectInst; ++c) {
        on[c] = c < ni && (!a.active || a.active[c0 + c]);
        any = any || on[c];
    }
    if (!any) return;
    __shared__ float fs[kDirectInst][kDirectMaxN];
    for (int c = 0; c < ni; ++c) {
        const float *f = a.f + (c0 + c) * a.f_stride;
        for (int j = threadIdx.x; j < n; j += blockDim.x) fs[c][j] = f[a.idx[j]];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int i = bx * kDirectWaves + (threadIdx.x >> 6);
    if (i >= n) return;
    const double *row = a.M + (size_t)i * n;
    double acc[kDirectInst] = {};
    if (NI == 1 || ni == 1) {
#pragma unroll 8
        for (int j = lane; j < n; j += 64) acc[0] = fma(row[j], (double)fs[0][j], acc[0]);
    } else {
#pragma unroll 4
        for (int j = lane; j < n; j += 64) {
            const double m = row[j];
#pragma unroll
            for (int c = 0; c < kDirectInst; ++c) acc[c] = fma(m, (double)fs[c][j], acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < kDirectInst; ++c) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
        if (lane == 0 && on[c]) a.x[(c0 + c) * a.x_stride + a.idx[i]] = (float)acc[c];
    }
}

template <int NI>
__global__ __launch_bounds__(64 * kDirectWaves) void k_direct_bottom(DirectArgs a) {
    direct_bottom_body<NI>(a, blockIdx.x, blockIdx.y);
}

// Several plans' direct bottoms in one launch: job j owns `count` = rows x groups workgroups from
// `first` on, workgroup first + g rows + r taking output rows r kDirectWaves .. and instance group g.
struct DirectJob {
    DirectArgs a;
    int first;
    int rows;
    int count;
};

template <int NI>
__global__ __launch_bounds__(64 * kDirectWaves) void k_direct_bottom_jobs(const DirectJob *__restrict__ jobs,
                                                                          int njobs) {
    int j = 0;
    while (j + 1 < njobs && (int)blockIdx.x >= jobs[j + 1].first) ++j;
    const int local = (int)blockIdx.x - jobs[j].first;
    if (local < 0 || local >= jobs[j].count) return;
    const DirectArgs a = jobs[j].a;
    direct_bottom_body<NI>(a, local % jobs[j].rows, local / jobs[j].rows);
}

// ------------------------------------------------------------------------------------------------
// Flexible GMRES (Saad 1993, SIAM J. Sci. Comput. 14, 461-469) on the binary32 correction problem,
// right-preconditioned by one binary32 V-cycle, inside the binary64 refinement. Unnormalised
// Arnoldi: basis vectors q_j are kept with their norms nu_j, so that no pass only rescales. The
// small dense work (Hessenberg, Givens rotations, back-substitution) runs in one-block kernels on
// device scalars, so an inner solve needs no host round trip.
// ------------------------------------------------------------------------------------------------
constexpr int kFgmresMax = 16;

struct FgmresState {
    double H[(kFgmresMax + 1) * kFgmresMax];  // column-major, ld = kFgmresMax + 1
    double cs[kFgmresMax], sn[kFgmresMax];
    double g[kFgmresMax + 1];
    double nu[kFgmresMax + 1];
    double coef[kFgmresMax + 1];  // k_orth's: coef[i] = h_ij / nu_i; scale = 1 / nu_j
    double scale;
    double y[kFgmresMax];
    double resid;                 // |g_{j+1}|: the least-squares residual estimate
};

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += sh[i];
    return t;
}

// partials[b] = sum over the block's cells of a * b (binary64), 2D launch as k_residual64.
__global__ __launch_bounds__(256) void k_dot_partials(int across, int down, const float *__restrict__ a,
                                                      const float *__restrict__ b, double *__restrict__ partials) {
    __shared__ double sh[16];
    const int col = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (col < across) {
#pragma unroll
        for (int k = 0; k < kRowsPerBlock; ++k) {
            const int row = blockIdx.y * kRowsPerBlock + k;
            if (row >= down) break;
            const long cell = (long)row * across + col;
            acc += (double)a[cell] * (double)b[cell];
        }
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x + blockIdx.y * gridDim.x] = t;
}

// a = (W - I) z on water (0 on land), and partials[i * nb + b] = sum a * q_i for i <= j.
__global__ __launch_bounds__(256) void k_apply_dots(LevelArgs L, const float *__restrict__ z, float *__restrict__ a,
                                                    const float *__restrict__ Q, long qstride, int j,
                                                    double *__restrict__ partials) {
    __shared__ double sh[16];
    const int across = L.across;
    const int col = blockIdx.x * 256 + threadIdx.x;
    double acc[kFgmresMax + 1];
#pragma unroll
    for (int i = 0; i <= kFgmresMax; ++i) acc[i] = 0.0;
    if (col < across) {
        const int ce = col + 1 == across ? 0 : col + 1;
        const int cw = col == 0 ? across - 1 : col - 1;
#pragma unroll
        for (int k = 0; k < kRowsPerBlock; ++k) {
            const int row = blockIdx.y * kRowsPerBlock + k;
            if (row >= L.down) break;
            const long cell = (long)row * across + col;
            float v = 0.0f;
            if (L.water[cell]) {
                const float zn = row > 0 ? z[cell - across] : 0.0f;
                const float zs = row + 1 < L.down ? z[cell + across] : 0.0f;
                v = rb_update(weights_at(L, row, cell), z[(long)row * across + ce], z[(long)row * across + cw], zn, zs,
                              z[cell]);
            }
            a[cell] = v;
#pragma unroll
            for (int i = 0; i <= kFgmresMax; ++i)
                if (i <= j) acc[i] += (double)v * (double)Q[i * qstride + cell];
        }
    }
    const int nb = gridDim.x * gridDim.y, b = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
    for (int i = 0; i <= kFgmresMax; ++i) {
        if (i > j) break;
        const double t = block_sum(acc[i], sh);
        if (threadIdx.x == 0) partials[(long)i * nb + b] = t;
    }
}

// q_{j+1} = a * scale - sum_{i<=j} coef_i q_i, and partials[b] = sum q_{j+1}^2.
__global__ __launch_bounds__(256) void k_orth(int across, int down, const float *__restrict__ a, float *Q, long qstride,
                                              int j, const FgmresState *__restrict__ st, double *__restrict__ partials) {
    __shared__ double sh[16];
    const int col = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (col < across) {
        const float scale = (float)st->scale;
        float c[kFgmresMax + 1];
#pragma unroll
        for (int i = 0; i <= kFgmresMax; ++i) c[i] = i <= j ? (float)st->coef[i] : 0.0f;
        float *out = Q + (long)(j + 1) * qstride;
#pragma unroll
        for (int k = 0; k < kRowsPerBlock; ++k) {
            const int row = blockIdx.y * kRowsPerBlock + k;
            if (row >= down) break;
            const long cell = (long)row * across + col;
            float v = a[cell] * scale;
#pragma unroll
            for (int i = 0; i <= kFgmresMax; ++i)
                if (i <= j) v -= c[i] * Q[i * qstride + cell];
            out[cell] = v;
            acc += (double)v * (double)v;
        }
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x + blockIdx.y * gridDim.x] = t;
}

// The inner solve's start: nu_0 = beta = |q_0| from the partials, g = (beta, 0, ...).
__global__ __launch_bounds__(256) void k_fgmres_start(const double *__restrict__ partials, int nb, FgmresState *st) {
    __shared__ double sh[16];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partials[b];
    const double t = block_sum(v, sh);
    if (threadIdx.x == 0) {
        const double beta = sqrt(t);
        st->nu[0] = beta;
        for (int i = 0; i <= kFgmresMax; ++i) st->g[i] = 0.0;
        st->g[0] = beta;
        st->resid = beta;
    }
}

// h_ij = <a, q_i> / (nu_j nu_i); k_orth's coefficients h_ij / nu_i and scale 1 / nu_j.
__global__ __launch_bounds__(256) void k_fgmres_dots(const double *__restrict__ partials, int nb, int j,
                                                     FgmresState *st) {
    __shared__ double sh[16];
    for (int i = 0; i <= j; ++i) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partials[(long)i * nb + b];
        const double t = block_sum(v, sh);
        if (threadIdx.x == 0) {
            const double nuj = st->nu[j];
            const double h = nuj > 0.0 && st->nu[i] > 0.0 ? t / (nuj * st->nu[i]) : 0.0;
            st->H[j * (kFgmresMax + 1) + i] = h;
            st->coef[i] = st->nu[i] > 0.0 ? h / st->nu[i] : 0.0;
        }
    }
    if (threadIdx.x == 0) st->scale = st->nu[j] > 0.0 ? 1.0 / st->nu[j] : 0.0;
}

// nu_{j+1} = |q_{j+1}| = h_{j+1,j}; the previous rotations applied to column j, a new one chosen
// to annihilate h_{j+1,j}, and g rotated with it.
__global__ __launch_bounds__(256) void k_fgmres_norm(const double *__restrict__ partials, int nb, int j,
                                                     FgmresState *st) {
    __shared__ double sh[16];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partials[b];
    const double t = block_sum(v, sh);
    if (threadIdx.x != 0) return;
    const double hn = sqrt(t);
    st->nu[j + 1] = hn;
    double *h = st->H + j * (kFgmresMax + 1);
    h[j + 1] = hn;
    for (int i = 0; i < j; ++i) {
        const double a = st->cs[i] * h[i] + st->sn[i] * h[i + 1];
        const double b = -st->sn[i] * h[i] + st->cs[i] * h[i + 1];
        h[i] = a;
        h[i + 1] = b;
    }
    const double r = hypot(h[j], h[j + 1]);
    const double c = r > 0.0 ? h[j] / r : 1.0, s = r > 0.0 ? h[j + 1] / r : 0.0;
    st->cs[j] = c;
    st->sn[j] = s;
    h[j] = r;
    h[j + 1] = 0.0;
    st->g[j + 1] = -s * st->g[j];
    st->g[j] = c * st->g[j];
    st->resid = fabs(st->g[j + 1]);
}

// y = R^-1 g over the first m columns, and the combination's coefficients y_j / nu_j.
__global__ void k_fgmres_solve(int m, FgmresState *st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = m - 1; i >= 0; --i) {
        double v = st->g[i];
        for (int k = i + 1; k < m; ++k) v -= st->H[k * (kFgmresMax + 1) + i] * st->y[k];
        const double d = st->H[i * (kFgmresMax + 1) + i];
        st->y[i] = d != 0.0 ? v / d : 0.0;
    }
    for (int i = 0; i < m; ++i) st->coef[i] = st->nu[i] > 0.0 ? st->y[i] / st->nu[i] : 0.0;
}

// delta = sum_{i<m} coef_i z_i.
__global__ __launch_bounds__(256) void k_combine(long cells, const float *__restrict__ Z, long zstride, int m,
                                                 const FgmresState *__restrict__ st, float *__restrict__ delta) {
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    double v = 0.0;
    for (int i = 0; i < m; ++i) v += st->coef[i] * (double)Z[i * zstride + cell];
    delta[cell] = (float)v;
}

// ------------------------------------------------------------------------------------------------
// The Krylov rescue's binary64 vector kernels (solveByKrylov, OC:548-607; applyOperator OC:637-659).
// ------------------------------------------------------------------------------------------------

// applyOperator: into = ((e zE + w zW + n zN + s zS) - z) * centre on water, 0 on land.
__global__ __launch_bounds__(256) void k_apply64(LevelArgs L, const double *__restrict__ z, double *__restrict__ into) {
    const int across = L.across;
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= across) return;
    const int ce = col + 1 == across ? 0 : col + 1;
    const int cw = col == 0 ? across - 1 : col - 1;
#pragma unroll
    for (int k = 0; k < kRowsPerBlock; ++k) {
        const int row = blockIdx.y * kRowsPerBlock + k;
        if (row >= L.down) break;
        const long cell = (long)row * across + col;
        double out = 0.0;
        if (L.water[cell]) {
            const double zn = row > 0 ? z[cell - across] : 0.0;
            const double zs = row + 1 < L.down ? z[cell + across] : 0.0;
            const float4 wt = weights_at(L, row, cell);
            const double nb = (double)wt.x * z[(long)row * across + ce] + (double)wt.y * z[(long)row * across + cw] +
                              (double)wt.z * zn + (double)wt.w * zs;
            out = (nb - z[cell]) * centre_at(L, row, cell);
        }
        into[cell] = out;
    }
}

// balanceOf (OC:325-327) into a binary64 array: f * centre.
__global__ void k_balance64(LevelArgs L, const float *__restrict__ f, double *__restrict__ out) {
    const long cells = (long)L.across * L.down;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    out[cell] = (double)f[cell] * centre_at(L, (int)(cell / L.across), cell);
}

// r = water ? balance - r : 0 (OC:565).
__global__ void k_rinit64(const uint8_t *__restrict__ water, const double *__restrict__ balance, double *r, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = water[i] ? balance[i] - r[i] : 0.0;
}

// Binary64 dot partials, one per block (the host adds them in block order).
__global__ __launch_bounds__(256) void k_dot64(const double *__restrict__ a, const double *__restrict__ b, long n,
                                               double *__restrict__ partials) {
    __shared__ double sh[16];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc += a[i] * b[i];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Largest magnitude of a binary64 vector into red.
__global__ __launch_bounds__(256) void k_maxabs64(const double *__restrict__ a, long n, MaxReduce red) {
    unsigned long long m = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const unsigned long long b = abs_bits(a[i]);
        m = b > m ? b : m;
    }
    grid_max(m, red);
}

// p = r + beta (p - omega v)  (OC:578).
__global__ void k_bicg_p(double *p, const double *__restrict__ r, const double *__restrict__ v, double beta,
                         double omega, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// y += a * x.
__global__ void k_axpy64(double *y, const double *__restrict__ x, double a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// x += omega shat; r -= omega t  (OC:595-598, one loop as the reference's).
__global__ void k_bicg_xr(double *x, double *r, const double *__restrict__ shat, const double *__restrict__ t,
                          double omega, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] += omega * shat[i];
        r[i] -= omega * t[i];
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

// ---- the round's input and output steps ------------------------------------------------------

// OceanStage.waterOn (OS:551-573): solve cell (row, col) is water when every map cell it overlaps by
// more than 1e-6 of a map cell (TOUCHING_SHARE, OS:129) is water. One thread per solve cell, grid
// (across / 256, down); the map is read straight from HBM (the overlap is a few cells).
__global__ __launch_bounds__(256) void k_water_on(const uint8_t *__restrict__ land, int map_across, int map_down,
                                                  int across, int down, uint8_t *__restrict__ water) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    if (col >= across) return;
    const double touching = 1e-6;
    const double cols_per = (double)map_across / across;
    const double rows_per = (double)map_down / down;
    const int first_row = (int)floor(row * rows_per + touching);
    const int last_row = min((int)ceil((row + 1) * rows_per - touching) - 1, map_down - 1);
    const int first_col = (int)floor(col * cols_per + touching);
    const int last_col = min((int)ceil((col + 1) * cols_per - touching) - 1, map_across - 1);
    int any_land = 0;
    for (int mr = first_row; mr <= last_row; ++mr)
        for (int mc = first_col; mc <= last_col; ++mc) any_land |= land[(long)mr * map_across + mc];
    water[(long)row * across + col] = any_land ? 0 : 1;
}

// OceanStage.curlForcing (OS:689-708): F = (d tau_n/dx - d tau_e/dy) / (rho H), binary64, central
// differences, columns wrapping, edge rows one-sided, the row above the northern one. rho = 1025
// (OS:64); H the 500 m wind-driven layer (OS:74) when depth is null, else depth per cell. F is
// written when non-null; f, when non-null, receives withBalance (OC:222-231): (float)(F / c) on
// water, 0 on land, c the stencil L's centre weight.
__global__ __launch_bounds__(256) void k_curl_forcing(int across, int down, double width_m, double height_m,
                                                      const double *__restrict__ east,
                                                      const double *__restrict__ north,
                                                      const double *__restrict__ depth, double *__restrict__ F,
                                                      LevelArgs L, float *__restrict__ f) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    if (col >= across) return;
    const int rn = row > 0 ? row - 1 : 0;
    const int rs = row + 1 < down ? row + 1 : down - 1;
    const double ns_m = (double)(rs - rn) * height_m;
    const int ce = col + 1 == across ? 0 : col + 1;
    const int cw = col == 0 ? across - 1 : col - 1;
    const long base = (long)row * across;
    const long cell = base + col;
    const double dtn_dx = (north[base + ce] - north[base + cw]) / (2.0 * width_m);
    const double dte_dy = (east[(long)rn * across + col] - east[(long)rs * across + col]) / ns_m;
    const double scale = depth ? 1.0 / (1025.0 * depth[cell]) : 1.0 / (1025.0 * 500.0);
    const double v = (dtn_dx - dte_dy) * scale;
    if (F) F[cell] = v;
    if (f) f[cell] = L.water[cell] ? (float)(v / centre_at(L, row, cell)) : 0.0f;
}

// OceanCirculation.velocities (OC:801-828): eastward -(psi_N - psi_S) / 2dy and northward
// (psi_E - psi_W) / 2dx at cell centres, the differences in binary32, the quotients in binary64 then
// rounded; zero on land; beyond a pole psi is minus the edge row's. psi is s32, or s64 rounded to
// binary32 per load when s32 is null.
__global__ __launch_bounds__(256) void k_velocities(int across, int down, const uint8_t *__restrict__ water,
                                                    const float *__restrict__ s32, const double *__restrict__ s64,
                                                    double width_m, double height_m, float *__restrict__ east,
                                                    float *__restrict__ north) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    if (col >= across) return;
    const long base = (long)row * across;
    const long cell = base + col;
    if (!water[cell]) {
        east[cell] = 0.0f;
        north[cell] = 0.0f;
        return;
    }
    auto psi = [&](long i) -> float { return s32 ? s32[i] : (float)s64[i]; };
    const int ce = col + 1 == across ? 0 : col + 1;
    const int cw = col == 0 ? across - 1 : col - 1;
    const float here = psi(cell);
    const float sn = row > 0 ? psi(cell - across) : -here;
    const float ss = row + 1 < down ? psi(cell + across) : -here;
    east[cell] = (float)(-(double)(sn - ss) / (2.0 * height_m));
    north[cell] = (float)((double)(psi(base + ce) - psi(base + cw)) / (2.0 * width_m));
}

// ---- correctBodyMeans (OC:709-731) --------------------------------------------------------------

constexpr int kBodySegment = 4096;  // most cells of one body one workgroup sums

// The binary64 residual of one water cell, in k_residual64's order.
__device__ __forceinline__ double residual64_at(const LevelArgs &L, const float *__restrict__ f,
                                                const float *__restrict__ x, int row, int col) {
    const int across = L.across;
    const long cell = (long)row * across + col;
    const int ce = col + 1 == across ? 0 : col + 1;
    const int cw = col == 0 ? across - 1 : col - 1;
    const double vn = row > 0 ? (double)x[cell - across] : 0.0;
    const double vs = row + 1 < L.down ? (double)x[cell + across] : 0.0;
    const float4 wt = weights_at(L, row, cell);
    const double bal = (double)wt.x * (double)x[(long)row * across + ce] + (double)wt.y * (double)x[(long)row * across + cw] +
                       (double)wt.z * vn + (double)wt.w * vs - (double)f[cell];
    return ((double)x[cell] - bal) * centre_at(L, row, cell);
}

// One workgroup per segment of the body-grouped cell list: the sum of the residuals of its cells,
// each thread summing a strided share in order, then a fixed tree: deterministic.
__global__ __launch_bounds__(256) void k_body_segment_sums(LevelArgs L, const float *__restrict__ f,
                                                           const float *__restrict__ x,
                                                           const int *__restrict__ order,
                                                           const int *__restrict__ seg_start,
                                                           const int *__restrict__ seg_len,
                                                           double *__restrict__ seg_sum) {
    __shared__ double part[256];
    const int seg = blockIdx.x;
    const int start = seg_start[seg], len = seg_len[seg];
    double acc = 0.0;
    for (int i = threadIdx.x; i < len; i += 256) {
        const int cell = order[start + i];
        const int row = cell / L.across;
        acc += residual64_at(L, f, x, row, cell - row * L.across);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) seg_sum[seg] = part[0];
}

// One thread per body: its segments' sums in order, over its margin sum, rounded to binary32; a body
// with no margin is left alone (shift 0).
__global__ __launch_bounds__(256) void k_body_shift(const double *__restrict__ seg_sum,
                                                    const int *__restrict__ body_seg0,
                                                    const int *__restrict__ body_nseg,
                                                    const double *__restrict__ margin, int bodies,
                                                    float *__restrict__ shift) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bodies) return;
    double sum = 0.0;
    const int s0 = body_seg0[b], ns = body_nseg[b];
    for (int k = 0; k < ns; ++k) sum += seg_sum[s0 + k];
    const double m = margin[b];
    shift[b] = m <= 0.0 ? 0.0f : (float)(sum / m);
}

// x -= shift[body] on every cell of a body.
__global__ __launch_bounds__(256) void k_body_apply(const int *__restrict__ body_of, const float *__restrict__ shift,
                                                    float *__restrict__ x, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = body_of[i];
    if (b >= 0) x[i] -= shift[b];
}

}  // namespace cg
