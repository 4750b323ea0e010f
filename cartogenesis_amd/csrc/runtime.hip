// The ocean solver's runtime and C ABI: device contexts, resident stencils, the relaxation seam,
// device-resident solve plans (reference-exact multigrid, binary32-in-binary64 refinement, FGMRES),
// hipGraph replay and kernel timing. See include/cgocean.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <string>
#include <vector>

#include "cgocean.h"
#include "coarse_net.hpp"
#include "kernels.hpp"

using namespace cg;

namespace {

constexpr int kMaxLevels = 24;
constexpr int kMostCycles = 200;          // OC:116
constexpr double kCoarsestShare = 0.1;    // OC:110
constexpr int kPrecondCoarsestPasses = 200;  // OC:629
constexpr int kKrylovRestarts = 3;        // OC:535
constexpr int kMaxInstances = 64;          // right-hand sides one plan solves together
constexpr int kMaxGroupPlans = 16;         // plans one group solves together
constexpr size_t kMaxMergedSets = 32;       // merged step graphs a group keeps (sets of running plans)
constexpr int kDirectMaxWater = 2048;     // largest coarsest grid solved by a dense inverse
constexpr int kRefinedMostSteps = 1000;
// The learned correction's fallback: a residual falling by less than this factor a step (over two
// steps). Learned solves that pay contract by about 0.42 a step, the plain cycle by 0.40; a learned
// instance at 0.65-0.7 a step (measured: a 1024 x 2048 world whose learned solve took 39-45 steps
// against 20 plain under a factor of 0.8) goes on with the plain cycle from its fourth step.
constexpr double kNetFallbackRate = 0.6;
// Plain multigrid judges a learned answer's error: the estimate rho / (1 - rho) |delta| assumes one
// contraction for every component of the error, which the plain cycle's nearly uniform smoothing
// and exact coarse solve give, but a learned correction does not (a smooth component it gets
// slightly wrong decays slowly while |delta| shrinks with the fast ones). So the learned steps stop
// on the residual alone, and every instance then goes on from its learned answer with plain steps,
// at least kJudgeSteps of them (so that the plain estimate's contraction is the larger of two ratios
// of plain corrections: after two steps its one ratio still carries the learned steps' fast
// transient), until the plain method's own stopping rule holds.
constexpr int kJudgeSteps = 3;
// ... and under this share of the error bound: warm from a learned answer, the plain estimate can
// still fall a percent or so short of the error at the step where it first holds (measured at
// 512 x 1024: 1.01e-10 under the bound 1e-10; one more step gives 3e-11).
constexpr double kJudgeErrorShare = 0.5;

struct Stencil {
    int across = 0, down = 0;
    long cells = 0;
    bool rowconst = false;
    bool has_centre = false;
    uint8_t *water = nullptr;
    float4 *roww = nullptr;
    float4 *cellw = nullptr;
    double *rowc = nullptr;
    double *cellc = nullptr;
    float *fbuf = nullptr;  // the seam's forcing
    float *xbuf[2] = {nullptr, nullptr};
    // Plans built on this stencil; a stencil released while plans still use it is freed when the
    // last of them is destroyed, so a release never pulls memory out from under a live plan.
    int plan_refs = 0;
    bool released = false;
    // Host copies, for the coarsest grid's direct solve.
    std::vector<uint8_t> h_water;
    std::vector<float4> h_w;   // per row or per cell
    std::vector<double> h_c;   // per row or per cell

    LevelArgs args() const {
        LevelArgs a;
        a.across = across;
        a.down = down;
        a.water = water;
        a.roww = rowconst ? roww : nullptr;
        a.cellw = rowconst ? nullptr : cellw;
        a.rowc = rowconst ? rowc : nullptr;
        a.cellc = rowconst ? nullptr : cellc;
        return a;
    }
};

struct TimedLaunch {
    hipEvent_t a, b;
    int kind;
};

// Kernel kinds of the timing record: the finest grid's fused smoother from zero (the cycle's
// down leg) and from given values (its up leg).
constexpr int kKindSmoothFromZero = 0;
constexpr int kKindSmooth = 1;

}  // namespace

// The allocation-free relaxation's scratch (cg_ocean_relax_noalloc): the forcing and the two value
// buffers of the ping-pong, each a grid of up to kScratchCells cells plus one all-zero row of up to
// kScratchAcross columns, in one device array that every GPU the library runs on holds until the
// process ends (25.2 MB). Host code never passes `scratch` itself to a kernel: its address on each
// device comes from hipGetSymbolAddress when a context opens on that device.
constexpr long kScratchCells = 1L << 21;   // 1024 x 2048
constexpr long kScratchAcross = 4096;
constexpr long kScratchGrid = kScratchCells + kScratchAcross;
static __device__ __align__(256) float scratch[3 * kScratchGrid];

struct cg_ctx {
    int device = 0;
    float *scratch = nullptr;  // this device's address of `scratch`
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::string err;
    std::string name;
    long min_cells = 1L << 16;
    int tile_rows = 16;
    int use_graph = 1;
    int fgmres_m = 20;
    int time_kernels = 0;
    int fused = 1;
    int fuse_outer = 1;  // the finest up leg and the binary64 outer step in one launch
    int net_fallback = 1;  // a learned solve switches to the plain cycle when its residual stalls
    int direct_bottom = 1;  // the binary32 cycle's bottom grid solved by its dense inverse
    int group_merged = 1;   // a group's refined steps as one graph whose launches cover every plan
    // The binary64-refined methods stop once the residual is under tol x stop_permille / 1000
    int stop_permille = 1000;
    // and, once a correction has been applied, their estimated relative L2 error (the step gate's)
    // is at most error_tol (0: the residual alone decides).
    double error_tol = 1e-10;
    // Passes per level of the fused binary32 cycle, 4 bits per level from level 0 up (2, 4 or 6;
    // levels past the eighth and the bottom's run 2). Four on the three grids below the finest: at
    // 1024 x 2048 the refinement takes 15 steps instead of 19 for 4-5 us more a step (823 against
    // 750 solves/s); six passes cost far more than they save.
    long level_passes = 0x22224442;
    size_t lds_limit = 0;       // dynamic LDS a workgroup may use (k_bottom)
    bool bottom_attr_set = false;
    std::map<uint64_t, Stencil *> stencils;
    std::set<Stencil *> orphans;  // released stencils that plans still use
    uint64_t next_handle = 1;
    unsigned long long *h_pinned = nullptr;  // host-visible scalars
    std::vector<TimedLaunch> timed;
    size_t timed_used = 0;
    long timed_dropped = 0;
};

namespace {

void set_err(cg_ctx *ctx, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
}

#define HIPCHK(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_err(ctx, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return CG_ERR_DEVICE;                                                              \
        }                                                                                      \
    } while (0)

inline unsigned blocks_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

template <typename T>
int dalloc(cg_ctx *ctx, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        set_err(ctx, "hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
        *p = nullptr;
        return CG_ERR_NOMEM;
    }
    return CG_OK;
}

// A grid array the smoother reads: cells plus one all-zero row after the last (row `down`),
// which rows beyond a pole load. Zero-filled.
template <typename T>
int dalloc_grid(cg_ctx *ctx, T **p, int across, int down) {
    const size_t n = (size_t)across * (down + 1);
    int rc = dalloc(ctx, p, n);
    if (rc) return rc;
    hipError_t e = hipMemset(*p, 0, sizeof(T) * n);
    if (e != hipSuccess) {
        set_err(ctx, "hipMemset failed: %s", hipGetErrorString(e));
        return CG_ERR_DEVICE;
    }
    return CG_OK;
}

void dfree(void *p) {
    if (p) (void)hipFree(p);
}

void free_stencil(Stencil *s) {
    if (!s) return;
    dfree(s->water);
    dfree(s->roww);
    dfree(s->cellw);
    dfree(s->rowc);
    dfree(s->cellc);
    dfree(s->fbuf);
    dfree(s->xbuf[0]);
    dfree(s->xbuf[1]);
    delete s;
}

Stencil *find_stencil(cg_ctx *ctx, cg_stencil h) {
    auto it = ctx->stencils.find(h);
    return it == ctx->stencils.end() ? nullptr : it->second;
}

// Event-timed launch bracket for kernel-duration records on the context's stream.
struct TimeScope {
    cg_ctx *ctx;
    TimedLaunch *tl = nullptr;
    TimeScope(cg_ctx *c, bool on, int kind) : ctx(c) {
        if (!on || !ctx->time_kernels) return;
        if (ctx->timed_used >= 200000) {
            ctx->timed_dropped++;
            return;
        }
        if (ctx->timed_used == ctx->timed.size()) {
            TimedLaunch t;
            if (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess) return;
            ctx->timed.push_back(t);
        }
        tl = &ctx->timed[ctx->timed_used++];
        tl->kind = kind;
        (void)hipEventRecord(tl->a, ctx->stream);
    }
    ~TimeScope() {
        if (tl) (void)hipEventRecord(tl->b, ctx->stream);
    }
};

// The fused smoother: `passes` red-black passes of a row-constant stencil from xin into xout,
// in launches of two passes (and one of one pass for an odd count), ping-ponging between xin and
// xout so that the final values land in xout. xin is read only when zero_start is false. For
// per-cell weights the colour kernel runs in place on xout after copying xin.
template <int TH, int MODE, int NH>
hipError_t launch_tile_mode(hipStream_t s, const LevelArgs &L, const float *f, const float *xin, float *xout, bool zero,
                            const TileExtra &ex, int instances) {
    const int iw = TileGeom<NH, MODE>::IW;
    const long waves = (long)((L.across + iw - 1) / iw) * ((L.down + TH - 1) / TH);
    unsigned blocks = (unsigned)((waves + 3) / 4);
    // With several instances (grid y), a row of blocks a multiple of the XCD count keeps each
    // block's XCD its x index modulo 8, which xcd_block's renumbering assumes.
    if (instances > 1) blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    const dim3 grid(blocks, instances > 1 ? instances : 1);
    if (zero) hipLaunchKernelGGL((k_relax_tile<NH, TH, true, MODE>), grid, dim3(256), 0, s, L, f, xin, xout, ex);
    else hipLaunchKernelGGL((k_relax_tile<NH, TH, false, MODE>), grid, dim3(256), 0, s, L, f, xin, xout, ex);
    return hipGetLastError();
}

template <int TH>
hipError_t launch_tile(hipStream_t s, const LevelArgs &L, const float *f, const float *xin, float *xout, int nh,
                       bool zero, int mode, const TileExtra &ex, int k) {
    // The fused down and up legs run 2, 4 or 6 passes (4, 8 or 12 half-sweeps) in their one launch.
    switch (mode) {
        case TILE_DOWN:
            if (nh == 12) return launch_tile_mode<TH, TILE_DOWN, 12>(s, L, f, xin, xout, true, ex, k);
            if (nh == 8) return launch_tile_mode<TH, TILE_DOWN, 8>(s, L, f, xin, xout, true, ex, k);
            return launch_tile_mode<TH, TILE_DOWN, 4>(s, L, f, xin, xout, true, ex, k);
        case TILE_UP:
            if (nh == 12) return launch_tile_mode<TH, TILE_UP, 12>(s, L, f, xin, xout, false, ex, k);
            if (nh == 8) return launch_tile_mode<TH, TILE_UP, 8>(s, L, f, xin, xout, false, ex, k);
            return launch_tile_mode<TH, TILE_UP, 4>(s, L, f, xin, xout, false, ex, k);
        case TILE_UP_OUTER:
            return launch_tile_mode<TH, TILE_UP_OUTER, 4>(s, L, f, xin, xout, false, ex, k);
        default:
            if (nh == 4) return launch_tile_mode<TH, TILE_PLAIN, 4>(s, L, f, xin, xout, zero, ex, k);
            return launch_tile_mode<TH, TILE_PLAIN, 2>(s, L, f, xin, xout, zero, ex, k);
    }
}

hipError_t tile_dispatch(cg_ctx *ctx, const LevelArgs &L, const float *f, const float *xin, float *xout, int nh,
                         bool zero, int mode = TILE_PLAIN, const TileExtra &ex = TileExtra{}, int rows = 0,
                         int instances = 1, hipStream_t stream = nullptr) {
    hipStream_t s = stream ? stream : ctx->stream;
    switch (rows ? rows : ctx->tile_rows) {
        case 2: return launch_tile<2>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
        case 4: return launch_tile<4>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
        case 6: return launch_tile<6>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
        case 8: return launch_tile<8>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
        case 24: return launch_tile<24>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
        case 32: return launch_tile<32>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
        default: return launch_tile<16>(s, L, f, xin, xout, nh, zero, mode, ex, instances);
    }
}

// Launches of k_relax_tile_jobs: the legs of several plans (one mode, tile height and pass count for
// all) in one launch, their job table in device memory. Tile heights 2, 4, 6 and 8; two or four
// passes (the outer step's leg: two).
constexpr int kJobRows[4] = {2, 4, 6, 8};

template <int TH, int MODE, int NH>
hipError_t launch_jobs_th(hipStream_t s, const TileJob *jobs, int njobs, unsigned blocks) {
    constexpr bool zero = MODE == TILE_DOWN;
    hipLaunchKernelGGL((k_relax_tile_jobs<NH, TH, zero, MODE>), dim3(blocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}

template <int MODE, int NH>
hipError_t launch_jobs_mode(hipStream_t s, int rows, const TileJob *jobs, int njobs, unsigned blocks) {
    switch (rows) {
        case 2: return launch_jobs_th<2, MODE, NH>(s, jobs, njobs, blocks);
        case 4: return launch_jobs_th<4, MODE, NH>(s, jobs, njobs, blocks);
        case 6: return launch_jobs_th<6, MODE, NH>(s, jobs, njobs, blocks);
        default: return launch_jobs_th<8, MODE, NH>(s, jobs, njobs, blocks);
    }
}

hipError_t launch_jobs(hipStream_t s, int mode, int nh, int rows, const TileJob *jobs, int njobs, unsigned blocks) {
    if (mode == TILE_UP_OUTER) return launch_jobs_mode<TILE_UP_OUTER, 4>(s, rows, jobs, njobs, blocks);
    if (mode == TILE_DOWN)
        return nh == 8 ? launch_jobs_mode<TILE_DOWN, 8>(s, rows, jobs, njobs, blocks)
                       : launch_jobs_mode<TILE_DOWN, 4>(s, rows, jobs, njobs, blocks);
    return nh == 8 ? launch_jobs_mode<TILE_UP, 8>(s, rows, jobs, njobs, blocks)
                   : launch_jobs_mode<TILE_UP, 4>(s, rows, jobs, njobs, blocks);
}

// The interior width of a tile of the given mode and half-sweeps (TileGeom's IW).
int jobs_tile_width(int mode, int nh) {
    if (mode == TILE_UP_OUTER) return TileGeom<4, TILE_UP_OUTER>::IW;
    if (mode == TILE_DOWN) return nh == 8 ? TileGeom<8, TILE_DOWN>::IW : TileGeom<4, TILE_DOWN>::IW;
    return nh == 8 ? TileGeom<8, TILE_UP>::IW : TileGeom<4, TILE_UP>::IW;
}

// Runs `passes` passes on level values held in a ping-pong pair x[0], x[1] with x[*cur] current;
// on return *cur names the buffer holding the result. zero_start: the current values are all zero
// and need not be read. timed: record the launches under kernel kind 0. stream: where to enqueue
// (nullptr: the context's stream).
int relax_pp(cg_ctx *ctx, const Stencil *st, const float *f, float *x[2], int *cur, int passes, bool zero_start,
             bool timed, hipStream_t stream = nullptr) {
    if (passes <= 0) return CG_OK;
    hipStream_t s = stream ? stream : ctx->stream;
    const LevelArgs L = st->args();
    if (st->rowconst && st->across >= 2 && (st->across & 1) == 0) {
        int left = passes;
        bool zero = zero_start;
        while (left > 0) {
            const int nh = left >= 2 ? 4 : 2;
            TimeScope ts(ctx, timed && nh == 4, zero ? kKindSmoothFromZero : kKindSmooth);
            hipError_t e = tile_dispatch(ctx, L, f, x[*cur], x[1 - *cur], nh, zero, TILE_PLAIN, TileExtra{}, 0, 1, s);
            if (e != hipSuccess) {
                set_err(ctx, "smoother launch failed: %s", hipGetErrorString(e));
                return CG_ERR_DEVICE;
            }
            *cur = 1 - *cur;
            left -= nh / 2;
            zero = false;
        }
        return CG_OK;
    }
    if (zero_start) {
        hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(st->cells, 256)), dim3(256), 0, s, x[*cur], st->cells, 0.0f);
    }
    const long half = (long)(st->across / 2) * st->down;
    for (int p = 0; p < passes; ++p)
        for (int colour = 0; colour <= 1; ++colour)
            hipLaunchKernelGGL(k_relax_colour, dim3(blocks_for(half, 256)), dim3(256), 0, s, L, f, x[*cur], colour);
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

}  // namespace

// ================================================================================================
// Plans
// ================================================================================================

struct PlanLevel {
    Stencil *st = nullptr;
    float *f = nullptr;         // this level's forcing (coarse levels: restricted; level 0: correction)
    float *x[2] = {nullptr, nullptr};
    int cur = 0;
    double *bal = nullptr;      // restricted balance (reference method), this level's size
    float *g = nullptr;         // per row: centre * share / centre of the next level (binary32)
};

struct cg_plan {
    cg_ctx *ctx = nullptr;
    int n = 0;
    int pole_is_wall = 1;
    // Instances: K right-hand sides on one ladder. Every per-instance grid array holds K copies,
    // each a grid with its zero row (vstride(l) floats); the binary64 answers K copies of the finest
    // cells. The refined method's residual slots of instance i and binary64 parity q are group
    // 2 i + q of d_res: the step writing binary64 parity 1 - q (reading q) reduces into group q. Its
    // error-estimate sums are group 2 i + q of d_sums (2 kMaxSlots doubles each; groups 2K and 2K + 1
    // serve kernel timing). The gate states, d_res and d_sums are consecutive pieces of one block
    // (d_check), so one copy brings a check's data to the host (h_check).
    int K = 1;
    int red_base = 0;
    int first_base = 4;
    unsigned long long *h_red = nullptr;  // pinned copy of d_u64
    int *d_active = nullptr, *h_active = nullptr;  // per instance: still solving
    // The refined method's step gate (gate_decide): state per binary64 parity and instance, the
    // residual thresholds per instance, the sums; `gated` while a refined solve (or the autotuner)
    // runs steps, so that every kernel of a step honours the active flags.
    unsigned char *d_check = nullptr, *h_check = nullptr;
    size_t check_bytes = 0;  // the piece a check copies (gate states, residual groups, sums)
    StepGate *d_gate = nullptr, *h_gate = nullptr;
    unsigned long long *d_res = nullptr, *h_res = nullptr;
    double *d_sums = nullptr, *h_sums = nullptr;
    // Per instance the residual threshold, then the error bound, then the fewest steps before a stop.
    double *d_thr = nullptr, *h_thr = nullptr;
    std::vector<double> thr_up;  // the thresholds and bounds last uploaded
    int *d_want = nullptr;
    bool gated = false;
    bool gate_in_down = false;  // the next finest down leg runs the gate (set by refine_step)
    std::vector<double> largest_k;        // per instance max |f c|, < 0 when unknown
    std::vector<int> ans_par;             // per instance: the binary64 parity holding its answer
    std::vector<int> last_steps;          // per instance: steps of its last solve from zero (0: none)
    PlanLevel lv[kMaxLevels];
    float *f0 = nullptr;          // the finest forcing
    float *f_alt = nullptr;       // with the fused outer step: level 0's other correction forcing
    uint32_t *d_bblob = nullptr;  // k_bottom's static LDS region
    int tuned_rows[kMaxLevels] = {0};  // per level: tile rows the autotuner chose (0: the default rule)
    double *r64 = nullptr;        // finest-size binary64 scratch
    double *x64[2] = {nullptr, nullptr};
    int x64cur = 0;
    bool x64_valid = false;
    float *x32ref = nullptr;      // reference-method answer lives in lv[0].x[lv[0].cur]
    // Maxima, in groups of kMaxSlots: group 0 and 1 the refined method's residual (by binary64
    // parity), 2 the largest balance, 3 the reference method's residual.
    unsigned long long *d_u64 = nullptr;
    int *d_passes = nullptr;
    double largest = -1.0;        // max |f0 * c0|, cached per forcing
    // Coarsest direct solve.
    int nw = 0;
    int *d_idx = nullptr;
    float *d_Mt = nullptr;
    // Graphs of one refinement step, for each parity of the binary64 ping-pong, and the context
    // settings they were captured under (a change of those rebuilds them).
    hipGraphExec_t step_exec[2] = {nullptr, nullptr};
    hipGraph_t step_graph[2] = {nullptr, nullptr};  // their graphs, which a group of plans composes
    int group_refs = 0;                              // groups (cg_ocean_group_create) holding the plan
    long graph_key = -1;
    double graph_etol = -1.0;
    // The learned coarse-grid correction (cg_ocean_plan_set_coarse_net): the network, whether the
    // steps being built or run use it, their graphs (same keys), and the last learned solve's
    // fallbacks to the plain cycle.
    CoarseNet net;
    bool net_on = false;
    hipGraphExec_t net_exec[2] = {nullptr, nullptr};
    long net_graph_key = -1;
    double net_graph_etol = -1.0;
    int net_fallbacks = 0;
    // The direct bottom (CG_OPT_DIRECT_BOTTOM): level `bottom` solved by M = -(I - W)^-1 over its
    // bnw water cells (row-major, binary64); bnw = 0 when the plan has none.
    int bnw = 0;
    int *d_bidx = nullptr;
    double *d_bM = nullptr;  // binary64
    // The fused cycle: tile levels 0..bottom-1 run fused down/up kernels, levels bottom..n-1 run
    // in one workgroup (k_bottom) with bottom_lds bytes of LDS; fusable is false when a level
    // has per-cell weights or does not halve both axes, and the unfused cycle runs instead.
    bool fusable = false;
    int bottom = 0;
    size_t bottom_lds = 0;
    // FGMRES workspace (allocated on first use): basis Q (kFgmresMax + 1 vectors) and
    // preconditioned vectors Z (kFgmresMax), each `vstride` floats (a grid with its zero row).
    // Krylov rescue workspace (binary64 vectors, allocated on first use).
    double *kx = nullptr, *kr = nullptr, *ksh = nullptr, *kp = nullptr, *kv = nullptr, *kph = nullptr, *kshat = nullptr,
           *kt = nullptr, *kbal = nullptr, *kpart = nullptr;
    float *kf32 = nullptr;
    std::vector<double> h_part;
    float *fq = nullptr, *fz = nullptr, *fa = nullptr, *fdelta = nullptr;
    // The caller's preconditioner for CG_METHOD_LEARNED (cg_ocean_plan_set_preconditioner).
    cg_precond_fn pre_fn = nullptr;
    void *pre_user = nullptr;
    long vstride = 0;
    double *fpart = nullptr;
    FgmresState *fst = nullptr;
    int last_answer = 0;  // 0: reference (binary32 in lv[0]); 1: binary64 in x64[x64cur]
    // The round's input and output steps (allocated on first use): the wind stress and depth on the
    // finest grid, and the velocities of the answer.
    double *tau_e = nullptr, *tau_n = nullptr, *depth = nullptr;
    float *vel = nullptr;
    // The finest grid's water bodies (correctBodyMeans, OC:709-731): body per cell; the cells grouped
    // by body in cell order, cut into segments of at most kBodySegment cells; per body its first
    // segment, its segment count and its margin sum c (1 - (e + w + n + s)), summed on the host in
    // cell order as the reference does.
    int nbodies = 0, nseg = 0;
    int *d_body_of = nullptr, *d_order = nullptr, *d_seg_start = nullptr, *d_seg_len = nullptr;
    int *d_body_seg0 = nullptr, *d_body_nseg = nullptr;
    double *d_margin = nullptr, *d_seg_sum = nullptr;
    float *d_shift = nullptr;
};

// Several plans of one context solved together (cg_ocean_group_*): worlds of uneven size, each a
// batched plan of its months, stepped by one host loop. Each plan's step graph is launched on a
// stream of the group's own (one per plan, default priority), so the plans' steps run concurrently
// on the device; events join those streams to the context's before each read of the gates.
struct cg_group {
    cg_ctx *ctx = nullptr;
    std::vector<cg_plan *> plans;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;  // per plan: its stream's last launch
    hipEvent_t start = nullptr;    // the context's stream after the solves' first outer steps
    // The merged steps (CG_OPT_GROUP_MERGED), one per set of plans still running (bit k: plan k),
    // built when a solve first needs it, and the settings they were built under.
    struct Merged {
        hipGraphExec_t exec[2] = {nullptr, nullptr};  // by binary64 parity
        TileJob *d_jobs = nullptr;                    // the tile launches' job tables
        DirectJob *d_djobs = nullptr;                 // the direct bottoms' (per parity)
    };
    std::map<unsigned, Merged> merged;
    std::vector<long> merged_key;
};

namespace {

void free_plan(cg_plan *p) {
    if (!p) return;
    for (int i = 0; i < p->n; ++i) {
        Stencil *st = p->lv[i].st;
        if (st && --st->plan_refs == 0 && st->released) {
            p->ctx->orphans.erase(st);
            free_stencil(st);
        }
        p->lv[i].st = nullptr;
        dfree(p->lv[i].f);
        dfree(p->lv[i].x[0]);
        dfree(p->lv[i].x[1]);
        dfree(p->lv[i].bal);
        dfree(p->lv[i].g);
    }
    dfree(p->f0);
    dfree(p->f_alt);
    dfree(p->d_bblob);
    dfree(p->r64);
    for (double *v : {p->kx, p->kr, p->ksh, p->kp, p->kv, p->kph, p->kshat, p->kt, p->kbal, p->kpart}) dfree(v);
    dfree(p->kf32);
    dfree(p->fq);
    dfree(p->fz);
    dfree(p->fa);
    dfree(p->fdelta);
    dfree(p->fpart);
    dfree(p->fst);
    dfree(p->x64[0]);
    dfree(p->x64[1]);
    dfree(p->d_u64);
    dfree(p->d_active);
    dfree(p->d_check);
    dfree(p->d_thr);
    dfree(p->d_want);
    if (p->h_red) (void)hipHostFree(p->h_red);
    if (p->h_active) (void)hipHostFree(p->h_active);
    for (void *h : {(void *)p->h_check, (void *)p->h_thr})
        if (h) (void)hipHostFree(h);
    dfree(p->d_passes);
    dfree(p->d_idx);
    dfree(p->d_Mt);
    dfree(p->d_bidx);
    dfree(p->d_bM);
    dfree(p->tau_e);
    dfree(p->tau_n);
    dfree(p->depth);
    dfree(p->vel);
    for (void *v : {(void *)p->d_body_of, (void *)p->d_order, (void *)p->d_seg_start, (void *)p->d_seg_len,
                    (void *)p->d_body_seg0, (void *)p->d_body_nseg, (void *)p->d_margin, (void *)p->d_seg_sum,
                    (void *)p->d_shift})
        dfree(v);
    for (int i = 0; i < 2; ++i)
        if (p->step_exec[i]) (void)hipGraphExecDestroy(p->step_exec[i]);
    for (int i = 0; i < 2; ++i)
        if (p->net_exec[i]) (void)hipGraphExecDestroy(p->net_exec[i]);
    coarse_net_free(p->net);
    delete p;
}

// The coarsest grid's dense inverse, M = -(I - W)^-1 over its water cells, by Gauss-Jordan
// elimination with partial pivoting in binary64, stored column-major in binary32.
int build_coarsest_inverse(cg_ctx *ctx, cg_plan *p) {
    const Stencil *st = p->lv[p->n - 1].st;
    const int across = st->across, down = st->down;
    std::vector<int> idx;
    std::vector<int> pos(st->cells, -1);
    for (long c = 0; c < st->cells; ++c)
        if (st->h_water[c]) {
            pos[c] = (int)idx.size();
            idx.push_back((int)c);
        }
    const int n = (int)idx.size();
    p->nw = n;
    if (n == 0 || n > kDirectMaxWater) return CG_OK;
    std::vector<double> a((size_t)n * n, 0.0), inv((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int cell = idx[i];
        const int row = cell / across, col = cell % across;
        const float4 w = st->rowconst ? st->h_w[row] : st->h_w[cell];
        a[(size_t)i * n + i] = 1.0;
        auto sub = [&](int nb, double wt) {
            if (nb >= 0 && pos[nb] >= 0) a[(size_t)i * n + pos[nb]] -= wt;
        };
        const int ce = col + 1 == across ? 0 : col + 1;
        const int cw = col == 0 ? across - 1 : col - 1;
        sub(row * across + ce, w.x);
        sub(row * across + cw, w.y);
        if (row > 0) sub(cell - across, w.z);
        if (row + 1 < down) sub(cell + across, w.w);
        inv[(size_t)i * n + i] = 1.0;
    }
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a[(size_t)i * n + k]) > std::fabs(a[(size_t)piv * n + k])) piv = i;
        if (a[(size_t)piv * n + k] == 0.0) {
            p->nw = 0;  // singular: fall back to relaxation
            return CG_OK;
        }
        if (piv != k)
            for (int j = 0; j < n; ++j) {
                std::swap(a[(size_t)k * n + j], a[(size_t)piv * n + j]);
                std::swap(inv[(size_t)k * n + j], inv[(size_t)piv * n + j]);
            }
        const double d = 1.0 / a[(size_t)k * n + k];
        for (int j = 0; j < n; ++j) {
            a[(size_t)k * n + j] *= d;
            inv[(size_t)k * n + j] *= d;
        }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const double m = a[(size_t)i * n + k];
            if (m == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                a[(size_t)i * n + j] -= m * a[(size_t)k * n + j];
                inv[(size_t)i * n + j] -= m * inv[(size_t)k * n + j];
            }
        }
    }
    std::vector<float> mt((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) mt[(size_t)j * n + i] = (float)(-inv[(size_t)i * n + j]);
    int rc;
    if ((rc = dalloc(ctx, &p->d_idx, n))) return rc;
    if ((rc = dalloc(ctx, &p->d_Mt, (size_t)n * n))) return rc;
    HIPCHK(ctx, hipMemcpy(p->d_idx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(p->d_Mt, mt.data(), sizeof(float) * mt.size(), hipMemcpyHostToDevice));
    return CG_OK;
}

// The direct bottom's inverse (cg_bottom_inverse) of level `bottom`, uploaded for k_direct_bottom.
int build_direct_bottom(cg_ctx *ctx, cg_plan *p) {
    p->bnw = 0;
    if (p->bottom <= 0) return CG_OK;
    const Stencil *st = p->lv[p->bottom].st;
    const long cells = st->cells;
    std::vector<float> w[4];
    for (int k = 0; k < 4; ++k) w[k].resize(cells);
    for (long c = 0; c < cells; ++c) {
        const float4 v = st->rowconst ? st->h_w[c / st->across] : st->h_w[c];
        w[0][c] = v.x;
        w[1][c] = v.y;
        w[2][c] = v.z;
        w[3][c] = v.w;
    }
    std::vector<int32_t> idx(kDirectMaxN);
    std::vector<double> m((size_t)kDirectMaxN * kDirectMaxN);
    const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int n = cg_bottom_inverse(st->across, st->down, st->h_water.data(), w[0].data(), w[1].data(),
                                    w[2].data(), w[3].data(), kDirectMaxN, threads, idx.data(), m.data());
    if (n <= 0) return CG_OK;
    int rc;
    if ((rc = dalloc(ctx, &p->d_bidx, n))) return rc;
    if ((rc = dalloc(ctx, &p->d_bM, (size_t)n * n))) return rc;
    HIPCHK(ctx, hipMemcpy(p->d_bidx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(p->d_bM, m.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice));
    p->bnw = n;
    return CG_OK;
}

// The dynamic LDS k_bottom may use on this device (its static LDS is 4 KiB of partial sums).
size_t bottom_lds_limit(cg_ctx *ctx) {
    if (ctx->bottom_attr_set) return ctx->lds_limit;
    ctx->bottom_attr_set = true;
    int optin = 0, plain = 0;
    (void)hipDeviceGetAttribute(&plain, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device);
    size_t total = (size_t)std::max(optin, plain);
    if (total < 16384) total = 65536;
    size_t dyn = total - 4096 - 256;
    if (hipFuncSetAttribute((const void *)k_bottom, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) {
        (void)hipGetLastError();
        dyn = (size_t)plain > 4096 + 256 ? (size_t)plain - 4096 - 256 : 0;
    }
    ctx->lds_limit = dyn;
    return dyn;
}

// Restriction factors per row, and which levels the fused cycle runs in one workgroup.
// k_bottom's LDS layout for levels bottom .. n-1: every grid's values and forcing, then the static
// region (each grid's mask, row weights, restriction factors), 16-byte aligned pieces.
void bottom_layout(const cg_plan *p, BottomArgs *ba) {
    const int n = p->n, b = p->bottom;
    ba->nlev = n - b;
    for (int l = b; l < n; ++l) {
        BottomLevel &bl = ba->lv[l - b];
        bl.across = p->lv[l].st->across;
        bl.down = p->lv[l].st->down;
    }
    size_t off = 0;
    for (int i = 0; i < ba->nlev; ++i) {
        const size_t cells = (size_t)ba->lv[i].across * ba->lv[i].down;
        ba->lv[i].off_x = (int)off;
        off += (cells * 4 + 15) & ~(size_t)15;
        ba->lv[i].off_f = (int)off;
        off += (cells * 4 + 15) & ~(size_t)15;
    }
    ba->static_base = (int)off;
    for (int i = 0; i < ba->nlev; ++i) {
        const size_t cells = (size_t)ba->lv[i].across * ba->lv[i].down;
        ba->lv[i].off_w = (int)off;
        off += (cells + 15) & ~(size_t)15;
        ba->lv[i].off_rw = (int)off;
        off += (size_t)ba->lv[i].down * 16;
        ba->lv[i].off_g = (int)off;
        off += ((size_t)ba->lv[i].down * 4 + 15) & ~(size_t)15;
    }
    ba->blob_words = (int)((off - (size_t)ba->static_base) / 4);
}

// The static region's bytes, from the stencils' host copies: mask (0/1), row weights, and
// g(row) = centre(row) / 4 / centre_next(row / 2) in binary32 (the tiles' restriction factor).
int build_bottom_blob(cg_ctx *ctx, cg_plan *p) {
    BottomArgs ba{};
    bottom_layout(p, &ba);
    std::vector<uint32_t> words((size_t)ba.blob_words, 0u);
    unsigned char *base = reinterpret_cast<unsigned char *>(words.data()) - ba.static_base;
    for (int i = 0; i < ba.nlev; ++i) {
        const Stencil *st = p->lv[p->bottom + i].st;
        for (long c = 0; c < st->cells; ++c) base[ba.lv[i].off_w + c] = st->h_water[c] ? 1 : 0;
        std::memcpy(base + ba.lv[i].off_rw, st->h_w.data(), sizeof(float4) * (size_t)st->down);
        float *g = reinterpret_cast<float *>(base + ba.lv[i].off_g);
        if (i + 1 < ba.nlev) {
            const Stencil *C = p->lv[p->bottom + i + 1].st;
            for (int row = 0; row < st->down; ++row) g[row] = (float)(st->h_c[row] * 0.25 / C->h_c[row / 2]);
        }
    }
    int rc;
    if ((rc = dalloc(ctx, &p->d_bblob, words.size()))) return rc;
    HIPCHK(ctx, hipMemcpy(p->d_bblob, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice));
    return CG_OK;
}

int plan_fusion_setup(cg_ctx *ctx, cg_plan *p) {
    const int n = p->n;
    bool rowconst = true;
    for (int l = 0; l < n; ++l) rowconst = rowconst && p->lv[l].st->rowconst;
    for (int l = 0; l + 1 < n; ++l) {
        const Stencil *F = p->lv[l].st, *C = p->lv[l + 1].st;
        if (!F->rowconst || !C->rowconst) continue;
        const int rpc = F->down / C->down, cpc = F->across / C->across;
        const double share = 1.0 / (double)(cpc * rpc);
        std::vector<float> g(F->down);
        for (int row = 0; row < F->down; ++row) g[row] = (float)(F->h_c[row] * share / C->h_c[row / rpc]);
        int rc;
        if ((rc = dalloc(ctx, &p->lv[l].g, F->down))) return rc;
        HIPCHK(ctx, hipMemcpy(p->lv[l].g, g.data(), sizeof(float) * F->down, hipMemcpyHostToDevice));
    }
    p->fusable = false;
    if (!rowconst) return CG_OK;
    const size_t limit = bottom_lds_limit(ctx);
    auto level_bytes = [&](int l) {
        const size_t cells = (size_t)p->lv[l].st->cells;
        const size_t down = (size_t)p->lv[l].st->down;
        return 2 * ((cells * 4 + 15) & ~(size_t)15) + ((cells + 15) & ~(size_t)15) + down * 16 +
               ((down * 4 + 15) & ~(size_t)15);
    };
    int b = n - 1;
    size_t bytes = level_bytes(b);
    if (bytes > limit || n - b > kBottomMaxLevels) return CG_OK;
    // k_bottom's levels are at most kBottomRows rows by 2 kBottomPairs columns.
    auto fits = [&](int l) {
        return p->lv[l].st->down <= kBottomRows && p->lv[l].st->across <= 2 * kBottomPairs;
    };
    if (!fits(b)) return CG_OK;
    while (b > 0 && n - (b - 1) <= kBottomMaxLevels && bytes + level_bytes(b - 1) <= limit && fits(b - 1)) {
        --b;
        bytes += level_bytes(b);
    }
    // Every grid of the fused cycle halves both axes into the next (the bottom's transfers and the
    // tiles' restriction assume 2 x 2 blocks).
    for (int l = 0; l + 1 < n; ++l) {
        const Stencil *F = p->lv[l].st, *C = p->lv[l + 1].st;
        if (F->across != 2 * C->across || F->down != 2 * C->down || (F->down & 1)) return CG_OK;
    }
    p->bottom = b;
    p->bottom_lds = bytes;
    int rc;
    if ((rc = build_bottom_blob(ctx, p))) return rc;
    if ((rc = build_direct_bottom(ctx, p))) return rc;
    p->fusable = true;
    return CG_OK;
}

// The largest of a group's kMaxSlots maxima, once copied to host memory.
unsigned long long group_max(const unsigned long long *slots) {
    unsigned long long m = 0;
    for (int i = 0; i < kMaxSlots; ++i) m = slots[i] > m ? slots[i] : m;
    return m;
}

// Reads the maximum held in instance 0's residual group of binary64 parity q (d_res).
int read_res(cg_plan *p, int q, unsigned long long *v) {
    cg_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipMemcpyAsync(p->h_res + q * kMaxSlots, p->d_res + q * kMaxSlots, kMaxSlots * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *v = group_max(p->h_res + q * kMaxSlots);
    return CG_OK;
}

// Reads the maximum held in slot group `group` (kMaxSlots slots at d_u64 + group * kMaxSlots).
int read_u64(cg_plan *p, int group, unsigned long long *v) {
    cg_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipMemcpyAsync(p->h_red + group * kMaxSlots, p->d_u64 + group * kMaxSlots,
                               kMaxSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *v = group_max(p->h_red + group * kMaxSlots);
    return CG_OK;
}

double bits_to_double(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// The 2D launch grid of the per-cell binary64 kernels.
inline dim3 grid2d(const Stencil *st) {
    return dim3(blocks_for(st->across, 256), blocks_for(st->down, kRowsPerBlock));
}

// Floats between two instances' copies of a grid: the grid and its zero row.
inline long vstride(const Stencil *st) { return (long)st->across * (st->down + 1); }

// The per-instance flags the kernels of a step honour: several instances' (a stopped instance's
// kernels return at once). One instance's step kernels read none, except the binary64 outer step
// while the refined method's gate decides (outer_flags): what the others write is the next step's
// scratch, so an extra step after the gate stopped the instance changes nothing but time.
inline const int *active_flags(const cg_plan *p) { return p->K > 1 ? p->d_active : nullptr; }
inline const int *outer_flags(const cg_plan *p) { return (p->K > 1 || p->gated) ? p->d_active : nullptr; }

// Instance 0's error-estimate sums for binary64 parity q (instances 4 kMaxSlots apart).
inline double *sums_group(const cg_plan *p, int q) { return p->d_sums + (size_t)q * 2 * kMaxSlots; }
constexpr int kSumsStride = 4 * kMaxSlots;

// The gate of the step that reads binary64 parity q = x64cur: the previous step's groups (parity
// 1 - q), state in from parity 1 - q, out to parity q.
GateArgs gate_args(const cg_plan *p) {
    const int q = p->x64cur;
    GateArgs g{};
    g.res = p->d_res + (size_t)(1 - q) * kMaxSlots;
    g.res_stride = 2 * kMaxSlots;
    g.sums = sums_group(p, 1 - q);
    g.sums_stride = kSumsStride;
    g.in = p->d_gate + (size_t)(1 - q) * p->K;
    g.out = p->d_gate + (size_t)q * p->K;
    g.active = p->d_active;
    g.want_sums = p->d_want;
    g.thr = p->d_thr;
    g.etol = p->d_thr + p->K;           // set with the thresholds by gate_reset
    g.min_steps = p->d_thr + 2 * p->K;  // likewise
    g.fall_rate = (p->net_on && p->ctx->net_fallback) ? kNetFallbackRate : 0.0;
    return g;
}

// The host's replay of gate_decide on the values read back, in the same arithmetic (the wave sums'
// butterfly, correctly rounded division and square root), so the host knows when the next gate
// would stop an instance without running that step.
double butterfly_sum(const double *v) {
    double a[kMaxSlots];
    for (int l = 0; l < kMaxSlots; ++l) a[l] = v[l];
    for (int off = 32; off > 0; off >>= 1) {
        double b[kMaxSlots];
        for (int l = 0; l < kMaxSlots; ++l) b[l] = a[l] + a[l ^ off];
        for (int l = 0; l < kMaxSlots; ++l) a[l] = b[l];
    }
    return a[0];
}

struct GateView {
    double res;   // max |u c| of the last step
    double d2, x2;
    double est;   // estimated relative L2 error (< 0 when the step applied no correction)
    double rho;   // contraction of the correction's norm
    bool stop;    // what the next gate decides
};

GateView gate_view(const unsigned long long *res_slots, const double *sums, const StepGate &st, double thr,
                   double etol, int min_steps) {
    GateView v{};
    unsigned long long m = 0;
    for (int l = 0; l < kMaxSlots; ++l) m = res_slots[l] > m ? res_slots[l] : m;
    std::memcpy(&v.res, &m, sizeof v.res);
    v.d2 = st.want ? butterfly_sum(sums) : 0.0;
    v.x2 = st.want ? butterfly_sum(sums + kMaxSlots) : 0.0;
    v.est = -1.0;
    v.rho = 0.5;
    bool err_ok = true;
    if (etol > 0.0 && st.steps > 0) {
        if (!st.want) {
            v.d2 = v.x2 = 0.0;
            err_ok = false;
        } else if (v.x2 > 0.0) {
            double rho = st.dprev2 > 0.0 ? std::sqrt(v.d2 / st.dprev2) : 0.5;
            if (st.dprev2 > 0.0 && st.dprev2b > 0.0) {
                const double before = std::sqrt(st.dprev2 / st.dprev2b);
                rho = before > rho ? before : rho;
            }
            rho = rho < 0.95 ? rho : 0.95;
            v.rho = rho;
            v.est = std::sqrt(v.d2 / v.x2) * (rho / (1.0 - rho));
            err_ok = v.est <= etol;
        } else {
            err_ok = v.d2 <= 0.0;
        }
    }
    v.stop = !st.active || v.res != v.res || (v.res < thr && err_ok && st.steps >= min_steps);
    return v;
}

// A reduction into slot group `group`, zeroed first by a fill kernel.
MaxReduce max_into(cg_plan *p, int group) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kMaxSlots), 0, p->ctx->stream, p->d_u64 + group * kMaxSlots, kMaxSlots,
                       0ull);
    MaxReduce r;
    r.out = p->d_u64 + group * kMaxSlots;
    r.clear = nullptr;
    return r;
}

// max |f0 * c0|: the reference's largest balance (OC:483).
int plan_largest(cg_plan *p, double *out) {
    if (p->largest >= 0.0) {
        *out = p->largest;
        return CG_OK;
    }
    cg_ctx *ctx = p->ctx;
    const Stencil *st = p->lv[0].st;
    const MaxReduce red = max_into(p, 2);
    hipLaunchKernelGGL(k_balance_max, grid2d(st), dim3(256), 0, ctx->stream, st->args(), p->f0, red);
    HIPCHK(ctx, hipGetLastError());
    unsigned long long b;
    int rc = read_u64(p, 2, &b);
    if (rc) return rc;
    p->largest = bits_to_double(b);
    *out = p->largest;
    return CG_OK;
}

// max |f c| of instance i's finest forcing, cached until its forcing changes.
int plan_largest_at(cg_plan *p, int i, double *out) {
    if (p->K == 1) return plan_largest(p, out);
    if (p->largest_k[i] >= 0.0) {
        *out = p->largest_k[i];
        return CG_OK;
    }
    cg_ctx *ctx = p->ctx;
    const Stencil *st = p->lv[0].st;
    const MaxReduce red = max_into(p, 2);
    hipLaunchKernelGGL(k_balance_max, grid2d(st), dim3(256), 0, ctx->stream, st->args(), p->f0 + i * vstride(st), red);
    HIPCHK(ctx, hipGetLastError());
    unsigned long long b;
    int rc = read_u64(p, 2, &b);
    if (rc) return rc;
    p->largest_k[i] = bits_to_double(b);
    *out = p->largest_k[i];
    return CG_OK;
}


// ---- the reference method ---------------------------------------------------------------------

// The binary64 residual's largest magnitude of level 0's current binary32 values.
int ref_residual(cg_plan *p, const float *f, double *out) {
    cg_ctx *ctx = p->ctx;
    const PlanLevel &l0 = p->lv[0];
    const MaxReduce red = max_into(p, 3);
    hipLaunchKernelGGL(k_residual64, grid2d(l0.st), dim3(256), 0, ctx->stream, l0.st->args(), f,
                       l0.x[l0.cur], (double *)nullptr, red);
    HIPCHK(ctx, hipGetLastError());
    unsigned long long b;
    int rc = read_u64(p, 3, &b);
    if (rc) return rc;
    *out = bits_to_double(b);
    return CG_OK;
}

// vCycle (OC:740-781) with level 0's forcing f0 and its values in lv[0]; fixed > 0 relaxes the
// coarsest that many passes instead of to convergence.
int ref_vcycle(cg_plan *p, const float *f0, double tol, int fixed) {
    cg_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const int n = p->n;
    int rc;
    for (int l = 0; l < n - 1; ++l) {
        PlanLevel &L = p->lv[l];
        PlanLevel &C = p->lv[l + 1];
        const float *f = l == 0 ? f0 : L.f;
        if ((rc = relax_pp(ctx, L.st, f, L.x, &L.cur, 2, false, l == 0))) return rc;
        MaxReduce none{};
        hipLaunchKernelGGL(k_residual64, grid2d(L.st), dim3(256), 0, s, L.st->args(), f, L.x[L.cur],
                           p->r64, none);
        hipLaunchKernelGGL(k_restrict64, dim3(blocks_for(C.st->cells, 256)), dim3(256), 0, s, p->r64, L.st->across,
                           L.st->down, C.st->across, C.st->down, C.bal);
        hipLaunchKernelGGL(k_with_balance, dim3(blocks_for(C.st->cells, 256)), dim3(256), 0, s, C.st->args(), C.bal, C.f);
        hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(C.st->cells, 256)), dim3(256), 0, s, C.x[C.cur], C.st->cells, 0.0f);
    }
    if (n > 1) {
        PlanLevel &C = p->lv[n - 1];
        hipLaunchKernelGGL(k_coarsest_relax, dim3(1), dim3(1024), 0, s, C.st->args(), C.f, C.x[C.cur],
                           tol * kCoarsestShare, fixed, p->d_passes);
    }
    for (int l = n - 2; l >= 0; --l) {
        PlanLevel &L = p->lv[l];
        PlanLevel &C = p->lv[l + 1];
        const float *f = l == 0 ? f0 : L.f;
        hipLaunchKernelGGL(k_prolong_add, dim3(blocks_for(L.st->cells, 256)), dim3(256), 0, s, C.x[C.cur], C.st->across,
                           C.st->down, L.x[L.cur], L.st->across, L.st->down, L.st->water,
                           p->pole_is_wall ? -1.0f : 1.0f);
        if ((rc = relax_pp(ctx, L.st, f, L.x, &L.cur, 2, false, l == 0))) return rc;
    }
    if (n == 1) {
        if ((rc = relax_pp(ctx, p->lv[0].st, f0, p->lv[0].x, &p->lv[0].cur, 2, false, true))) return rc;
    }
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

// correctBodyMeans (OC:709-731) of x against forcing f on the finest grid, when the plan has bodies.
int correct_bodies(cg_plan *p, const float *f, float *x) {
    if (p->nbodies == 0) return CG_OK;
    cg_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const Stencil *st = p->lv[0].st;
    if (p->nseg == 0) return CG_OK;  // bodies without cells: nothing to shift
    hipLaunchKernelGGL(k_body_segment_sums, dim3(p->nseg), dim3(256), 0, s, st->args(), f, x, p->d_order,
                       p->d_seg_start, p->d_seg_len, p->d_seg_sum);
    hipLaunchKernelGGL(k_body_shift, dim3(blocks_for(p->nbodies, 256)), dim3(256), 0, s, p->d_seg_sum, p->d_body_seg0,
                       p->d_body_nseg, p->d_margin, p->nbodies, p->d_shift);
    hipLaunchKernelGGL(k_body_apply, dim3(blocks_for(st->cells, 256)), dim3(256), 0, s, p->d_body_of, p->d_shift, x,
                       st->cells);
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

// solve (OC:473-493) from the binary32 values in lv[0], with correctBodyMeans after each cycle when
// the plan has bodies.
int ref_solve(cg_plan *p, double tol, cg_solution *info, const volatile int *cancel) {
    double largest;
    int rc = plan_largest(p, &largest);
    if (rc) return rc;
    double res = 0.0;
    if (largest > 0.0) {
        if ((rc = ref_residual(p, p->f0, &res))) return rc;
        res /= largest;
    }
    int cycles = 0;
    while (res >= tol && cycles < kMostCycles) {
        if (cancel && *cancel) {
            info->status = CG_CANCELLED;
            break;
        }
        if ((rc = ref_vcycle(p, p->f0, tol, 0))) return rc;
        if ((rc = correct_bodies(p, p->f0, p->lv[0].x[p->lv[0].cur]))) return rc;
        cycles++;
        if ((rc = ref_residual(p, p->f0, &res))) return rc;
        res /= largest;
    }
    info->cycles = cycles;
    info->relative_residual = res;
    p->last_answer = 0;
    return CG_OK;
}

// ---- the Krylov rescue (solveWithKrylovRescue / solveByKrylov, OC:507-607) --------------------

constexpr int kDotBlocks = 1024;

// dot (OC:661-665) on the device: per-block partial sums added on the host in block order.
int krylov_dot(cg_plan *p, const double *a, const double *b, double *out) {
    cg_ctx *ctx = p->ctx;
    const long n = p->lv[0].st->cells;
    hipLaunchKernelGGL(k_dot64, dim3(kDotBlocks), dim3(256), 0, ctx->stream, a, b, n, p->kpart);
    HIPCHK(ctx, hipMemcpyAsync(p->h_part.data(), p->kpart, sizeof(double) * kDotBlocks, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double sum = 0.0;
    for (int i = 0; i < kDotBlocks; ++i) sum += p->h_part[i];
    *out = sum;
    return CG_OK;
}

int krylov_largest(cg_plan *p, const double *a, double *out) {
    cg_ctx *ctx = p->ctx;
    const MaxReduce red = max_into(p, 3);
    hipLaunchKernelGGL(k_maxabs64, dim3(kDotBlocks), dim3(256), 0, ctx->stream, a, p->lv[0].st->cells, red);
    unsigned long long b;
    int rc = read_u64(p, 3, &b);
    if (rc) return rc;
    *out = bits_to_double(b);
    return CG_OK;
}

// precondition (OC:610-622): one reference V-cycle from zero on L z = balance, the coarsest relaxed
// a fixed 200 passes, then the bodies' means; answer widened to binary64.
int krylov_precondition(cg_plan *p, const double *balance, double *out, double tol) {
    cg_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    PlanLevel &l0 = p->lv[0];
    const long n = l0.st->cells;
    hipLaunchKernelGGL(k_with_balance, dim3(blocks_for(n, 256)), dim3(256), 0, s, l0.st->args(), balance, p->kf32);
    HIPCHK(ctx, hipMemsetAsync(l0.x[l0.cur], 0, sizeof(float) * n, s));
    int rc = ref_vcycle(p, p->kf32, tol, kPrecondCoarsestPasses);
    if (rc) return rc;
    if ((rc = correct_bodies(p, p->kf32, l0.x[l0.cur]))) return rc;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0, s, l0.x[l0.cur], out, n);
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

// solveByKrylov from the binary32 values in lv[0] (in and out), binary64 throughout.
int krylov_solve(cg_plan *p, double tol, int most, int *iterations, double *relative) {
    cg_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    PlanLevel &l0 = p->lv[0];
    const Stencil *st = l0.st;
    const long n = st->cells;
    const unsigned nb = blocks_for(n, 256);
    int rc;
    if (!p->kx) {
        for (double **v : {&p->kx, &p->kr, &p->ksh, &p->kp, &p->kv, &p->kph, &p->kshat, &p->kt, &p->kbal})
            if ((rc = dalloc(ctx, v, n))) return rc;
        if ((rc = dalloc(ctx, &p->kpart, kDotBlocks))) return rc;
        if ((rc = dalloc_grid(ctx, &p->kf32, st->across, st->down))) return rc;
        p->h_part.assign(kDotBlocks, 0.0);
    }
    double largest;
    if ((rc = plan_largest(p, &largest))) return rc;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(nb), dim3(256), 0, s, l0.x[l0.cur], p->kx, n);
    hipLaunchKernelGGL(k_balance64, dim3(nb), dim3(256), 0, s, st->args(), p->f0, p->kbal);
    hipLaunchKernelGGL(k_apply64, grid2d(st), dim3(256), 0, s, st->args(), p->kx, p->kr);
    hipLaunchKernelGGL(k_rinit64, dim3(nb), dim3(256), 0, s, st->water, p->kbal, p->kr, n);
    double r_max;
    if ((rc = krylov_largest(p, p->kr, &r_max))) return rc;
    double res = largest > 0.0 ? r_max / largest : 0.0;
    HIPCHK(ctx, hipMemcpyAsync(p->ksh, p->kr, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(ctx, hipMemsetAsync(p->kp, 0, sizeof(double) * n, s));
    HIPCHK(ctx, hipMemsetAsync(p->kv, 0, sizeof(double) * n, s));
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    while (res >= tol && it < most) {
        double rho_next;
        if ((rc = krylov_dot(p, p->ksh, p->kr, &rho_next))) return rc;
        if (rho_next == 0.0 || omega == 0.0) break;
        const double beta = (rho_next / rho) * (alpha / omega);
        hipLaunchKernelGGL(k_bicg_p, dim3(nb), dim3(256), 0, s, p->kp, p->kr, p->kv, beta, omega, n);
        if ((rc = krylov_precondition(p, p->kp, p->kph, tol))) return rc;
        hipLaunchKernelGGL(k_apply64, grid2d(st), dim3(256), 0, s, st->args(), p->kph, p->kv);
        double sv;
        if ((rc = krylov_dot(p, p->ksh, p->kv, &sv))) return rc;
        if (sv == 0.0) break;
        alpha = rho_next / sv;
        hipLaunchKernelGGL(k_axpy64, dim3(nb), dim3(256), 0, s, p->kr, p->kv, -alpha, n);
        hipLaunchKernelGGL(k_axpy64, dim3(nb), dim3(256), 0, s, p->kx, p->kph, alpha, n);
        rho = rho_next;
        it++;
        if ((rc = krylov_largest(p, p->kr, &r_max))) return rc;
        res = r_max / largest;
        if (res < tol) break;
        if ((rc = krylov_precondition(p, p->kr, p->kshat, tol))) return rc;
        hipLaunchKernelGGL(k_apply64, grid2d(st), dim3(256), 0, s, st->args(), p->kshat, p->kt);
        double tt, tr;
        if ((rc = krylov_dot(p, p->kt, p->kt, &tt))) return rc;
        if (tt == 0.0) omega = 0.0;
        else {
            if ((rc = krylov_dot(p, p->kt, p->kr, &tr))) return rc;
            omega = tr / tt;
        }
        hipLaunchKernelGGL(k_bicg_xr, dim3(nb), dim3(256), 0, s, p->kx, p->kr, p->kshat, p->kt, omega, n);
        if ((rc = krylov_largest(p, p->kr, &r_max))) return rc;
        res = r_max / largest;
    }
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nb), dim3(256), 0, s, p->kx, l0.x[l0.cur], n);
    HIPCHK(ctx, hipGetLastError());
    *iterations = it;
    *relative = res;
    return CG_OK;
}

// solveWithKrylovRescue (OC:507-528) from the binary32 values in lv[0].
int rescue_solve(cg_plan *p, double tol, cg_solution *info, const volatile int *cancel) {
    int rc = ref_solve(p, tol, info, cancel);
    if (rc || info->status == CG_CANCELLED) return rc;
    if (info->relative_residual < tol || !std::isfinite(info->relative_residual)) return CG_OK;
    // Values not all finite also return as the cycles left them (OC:516).
    const long n = p->lv[0].st->cells;
    std::vector<float> v(n);
    HIPCHK(p->ctx, hipMemcpy(v.data(), p->lv[0].x[p->lv[0].cur], sizeof(float) * n, hipMemcpyDeviceToHost));
    for (float a : v)
        if (!std::isfinite(a)) return CG_OK;
    int its = 0, total = 0, restarts = 0;
    double rr = info->relative_residual;
    if ((rc = krylov_solve(p, tol, kMostCycles, &its, &rr))) return rc;
    total = its;
    while (rr >= tol && restarts < kKrylovRestarts) {
        HIPCHK(p->ctx, hipMemcpy(v.data(), p->lv[0].x[p->lv[0].cur], sizeof(float) * n, hipMemcpyDeviceToHost));
        bool finite = true;
        for (float a : v) finite = finite && std::isfinite(a);
        if (!finite) break;
        if ((rc = krylov_solve(p, tol, kMostCycles, &its, &rr))) return rc;
        total += its;
        restarts++;
    }
    info->cycles += total;
    info->relative_residual = rr;
    return CG_OK;
}

// solveByKrylov from the binary32 values in lv[0], started afresh up to kKrylovRestarts times while
// it stops short of tol with finite values: the heat's loop (OS:486-497). cycles counts the
// iterations of every start, refinements the starts.
int krylov_restarted(cg_plan *p, double tol, cg_solution *info, const volatile int *cancel) {
    const long n = p->lv[0].st->cells;
    std::vector<float> v(n);
    int its = 0, total = 0, starts = 0;
    double rr = 0.0;
    int rc;
    while (true) {
        if (cancel && *cancel) {
            info->status = CG_CANCELLED;
            break;
        }
        if ((rc = krylov_solve(p, tol, kMostCycles, &its, &rr))) return rc;
        total += its;
        starts++;
        if (rr < tol || starts > kKrylovRestarts) break;
        HIPCHK(p->ctx, hipMemcpy(v.data(), p->lv[0].x[p->lv[0].cur], sizeof(float) * n, hipMemcpyDeviceToHost));
        bool finite = true;
        for (float a : v) finite = finite && std::isfinite(a);
        if (!finite) break;
    }
    info->cycles = total;
    info->refinements = starts;
    info->relative_residual = rr;
    p->last_answer = 0;
    return CG_OK;
}

// ---- the refined method ----------------------------------------------------------------------

// Tile rows for a grid of the fused cycle: the context's choice where it gives at least one
// wavefront per SIMD (1024 on the 256-CU MI355X), else 4, so that a small grid's launch, bound by
// one wavefront's serial path, runs the shortest path.
int level_tile_rows(const cg_ctx *ctx, const Stencil *st) {
    const int th = ctx->tile_rows;
    const long waves = (long)((st->across + 115) / 116) * ((st->down + th - 1) / th);
    return waves >= 1024 ? th : 4;
}

// Half-sweeps of level l's fused legs: twice its passes in the context's schedule.
int level_half_sweeps(const cg_ctx *ctx, int l) {
    if (l >= 8) return 4;
    return 2 * (int)((ctx->level_passes >> (4 * l)) & 0xf);
}

// The bottom grids (levels bottom .. n-1) in one workgroup per instance: k_bottom.
// Whether the binary32 cycle solves level `bottom` by its dense inverse (k_direct_bottom) instead
// of running k_bottom's levels.
inline bool direct_bottom(const cg_plan *p) { return p->ctx->direct_bottom && p->bnw > 0 && p->bottom > 0; }

// Level `bottom`'s values from its forcing by the dense inverse, every instance in one launch.
DirectArgs direct_args(cg_plan *p) {
    PlanLevel &B0 = p->lv[p->bottom];
    DirectArgs da{};
    da.idx = p->d_bidx;
    da.M = p->d_bM;
    da.n = p->bnw;
    da.K = p->K;
    da.f = B0.f;
    da.x = B0.x[B0.cur];
    da.f_stride = vstride(B0.st);
    da.x_stride = vstride(B0.st);
    da.active = active_flags(p);
    return da;
}

void direct_bottom_launch(cg_plan *p) {
    const DirectArgs da = direct_args(p);
    const unsigned rows = (p->bnw + kDirectWaves - 1) / kDirectWaves;
    if (p->K == 1)
        hipLaunchKernelGGL(k_direct_bottom<1>, dim3(rows, 1), dim3(64 * kDirectWaves), 0, p->ctx->stream, da);
    else
        hipLaunchKernelGGL(k_direct_bottom<4>, dim3(rows, (p->K + 3) / 4), dim3(64 * kDirectWaves), 0, p->ctx->stream,
                           da);
}

void bottom_launch(cg_plan *p, const float *in0, float *out0) {
    if (direct_bottom(p)) {
        direct_bottom_launch(p);
        return;
    }
    hipStream_t s = p->ctx->stream;
    const int b = p->bottom;
    const float beyond = p->pole_is_wall ? -1.0f : 1.0f;
    BottomArgs ba{};
    bottom_layout(p, &ba);
    ba.blob = p->d_bblob;
    PlanLevel &B0 = p->lv[b];
    ba.f_in = (b == 0 && in0) ? in0 : B0.f;
    ba.x_out = (b == 0 && out0) ? out0 : B0.x[B0.cur];
    ba.idx = p->d_idx;
    ba.Mt = p->d_Mt;
    // A coarsest grid with no water has the answer zero (its values start at zero); one with an
    // inverse is solved directly; otherwise it is relaxed a fixed number of passes.
    ba.nw = p->d_Mt ? p->nw : 0;
    ba.coarsest_passes = (p->nw == 0) ? 0 : kPrecondCoarsestPasses;
    ba.beyond = beyond;
    ba.f_stride = vstride(B0.st);
    ba.x_stride = vstride(B0.st);
    ba.active = active_flags(p);
    hipLaunchKernelGGL(k_bottom, dim3(p->K), dim3(1024), p->bottom_lds, s, ba);
}

// Level 0's correction forcing for binary64 parity q when the outer step is fused into the finest
// up leg: the two buffers alternate, since that launch writes the next forcing while its
// neighbouring waves still read this one.
inline float *level0_forcing(cg_plan *p, int q) { return q == 0 ? p->lv[0].f : p->f_alt; }

// Whether the refined method's steps fuse the outer step into the finest up leg.
inline bool outer_fused(const cg_plan *p) {
    const cg_ctx *ctx = p->ctx;
    return ctx->fuse_outer && ctx->fused && p->fusable && p->bottom > 0 && level_half_sweeps(ctx, 0) == 4;
}

// The fused cycle: per tile level one down launch (smooth from zero, residual, restriction) and
// one up launch (prolongation, smooth); the bottom levels in one k_bottom launch.
// With fuse_outer the finest up leg also runs the binary64 outer step (TILE_UP_OUTER): it reads
// x64[x64cur] and writes x64[1 - x64cur], the next cycle's forcing (level0_forcing(p, 1 - x64cur)) and
// the maxima of parity x64cur; the caller flips x64cur.
// Level l's tile rows: the autotuner's choice for this plan, else the context's rule.
int plan_tile_rows(const cg_plan *p, int l) {
    return p->tuned_rows[l] ? p->tuned_rows[l] : level_tile_rows(p->ctx, p->lv[l].st);
}

// One tile launch of a plan's cycle as data: what tile_dispatch takes, so that the legs of several
// plans can also share one launch (a group's merged step, launch_jobs).
struct LegSpec {
    LevelArgs L;
    const float *f = nullptr;
    const float *xin = nullptr;
    float *xout = nullptr;
    TileExtra ex{};
    int nh = 4;
    bool zero = false;
    int mode = TILE_PLAIN;
    int rows = 0;
    int K = 1;
};

hipError_t launch_spec(cg_ctx *ctx, const LegSpec &s) {
    return tile_dispatch(ctx, s.L, s.f, s.xin, s.xout, s.nh, s.zero, s.mode, s.ex, s.rows, s.K);
}

// Level l's down leg (smooth from zero, residual, restriction into level l + 1's forcing) with
// `rows` tile rows; level 0 reads in0 when given. Leaves the result in L.x[1 - L.cur].
LegSpec down_leg_spec(cg_plan *p, int l, const float *in0, int rows) {
    PlanLevel &L = p->lv[l];
    PlanLevel &C = p->lv[l + 1];
    LegSpec s;
    s.ex.cwater = C.st->water;
    s.ex.g = L.g;
    s.ex.fc = C.f;
    s.ex.fs = vstride(L.st);
    s.ex.cs = vstride(C.st);
    s.ex.active = active_flags(p);
    if (l == 0 && p->gate_in_down) {
        s.ex.gate = gate_args(p);
        s.ex.late = p->K == 1;
    }
    s.L = L.st->args();
    s.f = (l == 0 && in0) ? in0 : L.f;
    s.xin = L.x[L.cur];
    s.xout = L.x[1 - L.cur];
    s.nh = level_half_sweeps(p->ctx, l);
    s.zero = true;
    s.mode = TILE_DOWN;
    s.rows = rows;
    s.K = p->K;
    return s;
}

hipError_t down_leg(cg_plan *p, int l, const float *in0, int rows) {
    return launch_spec(p->ctx, down_leg_spec(p, l, in0, rows));
}

// Level l's up leg (prolongation of level l + 1's values, smoothing) with `rows` tile rows; at
// level 0 with fuse_outer, the outer step too (TILE_UP_OUTER, see vcycle32_fused).
LegSpec up_leg_spec(cg_plan *p, int l, const float *in0, float *out0, bool fuse_outer, int rows) {
    PlanLevel &L = p->lv[l];
    PlanLevel &C = p->lv[l + 1];
    LegSpec s;
    s.ex.xc = C.x[C.cur];
    s.ex.ca = C.st->across;
    s.ex.cd = C.st->down;
    s.ex.beyond = p->pole_is_wall ? -1.0f : 1.0f;
    s.ex.fs = vstride(L.st);
    s.ex.cs = vstride(C.st);
    s.ex.active = active_flags(p);
    s.mode = TILE_UP;
    if (l == 0 && fuse_outer) {
        const int q = p->x64cur;
        s.mode = TILE_UP_OUTER;
        s.ex.x64 = p->x64[q];
        s.ex.x64n = p->x64[1 - q];
        s.ex.f0 = p->f0;
        s.ex.fcn = level0_forcing(p, 1 - q);
        s.ex.red.out = p->d_res + q * kMaxSlots;
        s.ex.red.clear = p->d_res + (1 - q) * kMaxSlots;
        s.ex.s64 = L.st->cells;
        s.ex.red_stride = 2 * kMaxSlots;
        s.ex.sums.out = sums_group(p, q);
        s.ex.sums.clear = sums_group(p, 1 - q);
        s.ex.sums_stride = kSumsStride;
        s.ex.want_sums = p->gated ? p->d_want : nullptr;
        s.ex.active = outer_flags(p);
        s.ex.late = p->K == 1;
    }
    s.L = L.st->args();
    s.f = (l == 0 && in0) ? in0 : L.f;
    s.xin = L.x[L.cur];
    s.xout = (l == 0 && out0) ? out0 : L.x[1 - L.cur];
    s.nh = level_half_sweeps(p->ctx, l);
    s.zero = false;
    s.rows = rows;
    s.K = p->K;
    return s;
}

hipError_t up_leg(cg_plan *p, int l, const float *in0, float *out0, bool fuse_outer, int rows) {
    return launch_spec(p->ctx, up_leg_spec(p, l, in0, out0, fuse_outer, rows));
}

int vcycle32_fused(cg_plan *p, bool timed, const float *in0, float *out0, bool fuse_outer) {
    cg_ctx *ctx = p->ctx;
    // With the learned correction the cycle goes down to the network's grid, whose correction the
    // network proposes from its forcing, in place of the grids below it.
    const bool net = p->net_on && p->net.level > 0;
    const int b = net ? p->net.level : p->bottom;
    for (int l = 0; l < b; ++l) {
        TimeScope ts(ctx, timed && l == 0, kKindSmoothFromZero);
        hipError_t e = down_leg(p, l, in0, plan_tile_rows(p, l));
        if (e != hipSuccess) {
            set_err(ctx, "down-leg launch failed: %s", hipGetErrorString(e));
            return CG_ERR_DEVICE;
        }
        p->lv[l].cur = 1 - p->lv[l].cur;
    }
    if (net) {
        PlanLevel &N = p->lv[b];
        const int rc = coarse_net_forward(p->net, N.f, N.x[N.cur], ctx->stream);
        if (rc) {
            set_err(ctx, "the coarse network's launches failed (%d)", rc);
            return rc;
        }
    } else {
        bottom_launch(p, in0, out0);
    }
    for (int l = b - 1; l >= 0; --l) {
        TimeScope ts(ctx, timed && l == 0, kKindSmooth);
        hipError_t e = up_leg(p, l, in0, out0, fuse_outer, plan_tile_rows(p, l));
        if (e != hipSuccess) {
            set_err(ctx, "up-leg launch failed: %s", hipGetErrorString(e));
            return CG_ERR_DEVICE;
        }
        p->lv[l].cur = 1 - p->lv[l].cur;
    }
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

// One binary32 V(2,2) cycle on the correction problem from zero, with level 0's forcing lv[0].f.
// The answer is left in lv[0].x[lv[0].cur].
int vcycle32_unfused(cg_plan *p, bool timed);

// One binary32 V(2,2) cycle on the correction problem from zero. Level 0's forcing is in0 (or
// lv[0].f) and its answer goes to out0 (or stays in lv[0].x[lv[0].cur]).
int vcycle32(cg_plan *p, bool timed, const float *in0 = nullptr, float *out0 = nullptr) {
    if (p->ctx->fused && p->fusable) return vcycle32_fused(p, timed, in0, out0, false);
    cg_ctx *ctx = p->ctx;
    PlanLevel &L0 = p->lv[0];
    if (in0) HIPCHK(ctx, hipMemcpyAsync(L0.f, in0, sizeof(float) * L0.st->cells, hipMemcpyDeviceToDevice, ctx->stream));
    int rc = vcycle32_unfused(p, timed);
    if (rc) return rc;
    if (out0)
        HIPCHK(ctx, hipMemcpyAsync(out0, L0.x[L0.cur], sizeof(float) * L0.st->cells, hipMemcpyDeviceToDevice,
                                   ctx->stream));
    return CG_OK;
}

int vcycle32_unfused(cg_plan *p, bool timed) {
    cg_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    // With the direct bottom the cycle ends at level `bottom`, solved by its dense inverse, as the
    // fused cycle does; otherwise at the coarsest grid.
    const bool direct = direct_bottom(p);
    const int n = direct ? p->bottom + 1 : p->n;
    int rc;
    for (int l = 0; l < n - 1; ++l) {
        PlanLevel &L = p->lv[l];
        PlanLevel &C = p->lv[l + 1];
        if ((rc = relax_pp(ctx, L.st, L.f, L.x, &L.cur, 2, true, timed && l == 0))) return rc;
        hipLaunchKernelGGL(k_residual_restrict32, dim3(blocks_for(C.st->cells, 256)), dim3(256), 0, s, L.st->args(), L.f,
                           L.x[L.cur], C.st->args(), C.f);
    }
    PlanLevel &B = p->lv[n - 1];
    if (direct) {
        direct_bottom_launch(p);
    } else if (n == 1) {
        if ((rc = relax_pp(ctx, B.st, B.f, B.x, &B.cur, 2, true, timed))) return rc;
    } else if (p->nw == 0) {
        hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(B.st->cells, 256)), dim3(256), 0, s, B.x[B.cur], B.st->cells, 0.0f);
    } else if (p->d_Mt) {
        hipLaunchKernelGGL(k_coarsest_direct, dim3(1), dim3(256), sizeof(float) * p->nw, s, p->d_idx, p->nw, p->d_Mt, B.f,
                           B.x[B.cur], B.st->cells);
    } else {
        hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(B.st->cells, 256)), dim3(256), 0, s, B.x[B.cur], B.st->cells, 0.0f);
        hipLaunchKernelGGL(k_coarsest_relax, dim3(1), dim3(1024), 0, s, B.st->args(), B.f, B.x[B.cur], 0.0,
                           kPrecondCoarsestPasses, p->d_passes);
    }
    for (int l = n - 2; l >= 0; --l) {
        PlanLevel &L = p->lv[l];
        PlanLevel &C = p->lv[l + 1];
        hipLaunchKernelGGL(k_prolong_add, dim3(blocks_for(L.st->cells, 256)), dim3(256), 0, s, C.x[C.cur], C.st->across,
                           C.st->down, L.x[L.cur], L.st->across, L.st->down, L.st->water,
                           p->pole_is_wall ? -1.0f : 1.0f);
        if ((rc = relax_pp(ctx, L.st, L.f, L.x, &L.cur, 2, false, timed && l == 0))) return rc;
    }
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

// The outer binary64 step: x64[next] = x64[cur] + delta, its residual into lv[0].f, max into slot 0.
int outer_step(cg_plan *p, const float *delta, float *fc = nullptr) {
    cg_ctx *ctx = p->ctx;
    const Stencil *st = p->lv[0].st;
    // Slot x64cur receives the maximum (zeroed by the previous step); the other slot is zeroed for
    // the next step.
    MaxReduce red;
    red.out = p->d_res + p->x64cur * kMaxSlots;
    red.clear = p->d_res + (1 - p->x64cur) * kMaxSlots;
    OuterBatch ob{};
    ob.s32 = vstride(st);
    ob.s64 = st->cells;
    ob.red_stride = 2 * kMaxSlots;
    ob.active = outer_flags(p);
    ob.sums.out = sums_group(p, p->x64cur);
    ob.sums.clear = sums_group(p, 1 - p->x64cur);
    ob.sums_stride = kSumsStride;
    ob.want_sums = p->gated ? p->d_want : nullptr;
    const dim3 grid(blocks_for(st->across / 2, 256), blocks_for(st->down, kRowsPerBlock), p->K);
    hipLaunchKernelGGL(k_outer_residual, grid, dim3(256), 0, ctx->stream, st->args(), p->f0,
                       p->x64[p->x64cur], delta, p->x64[1 - p->x64cur], fc ? fc : p->lv[0].f, red, ob);
    HIPCHK(ctx, hipGetLastError());
    p->x64cur = 1 - p->x64cur;
    return CG_OK;
}

// One refinement step: a binary32 cycle on the current correction forcing and the binary64 outer
// step, fused into the finest up leg where outer_fused, else as its own launch.
int refine_step(cg_plan *p, bool timed) {
    // The step gate: inside the finest down leg when the step starts with one, else a launch of its
    // own first.
    const bool down_first = p->ctx->fused && p->fusable && p->bottom > 0;
    if (p->gated && !down_first) {
        hipLaunchKernelGGL(k_step_gate, dim3(1, p->K), dim3(64), 0, p->ctx->stream, gate_args(p));
        HIPCHK(p->ctx, hipGetLastError());
    }
    p->gate_in_down = p->gated && down_first;
    int rc;
    if (outer_fused(p)) {
        rc = vcycle32_fused(p, timed, level0_forcing(p, p->x64cur), nullptr, true);
        p->gate_in_down = false;
        if (rc) return rc;
        p->x64cur = 1 - p->x64cur;
        return CG_OK;
    }
    rc = vcycle32(p, timed);
    p->gate_in_down = false;
    if (rc) return rc;
    return outer_step(p, p->lv[0].x[p->lv[0].cur]);
}

// One refinement step as a graph: the V-cycle on the current correction forcing, then the outer
// step applying its answer. The residual maxima are copied to the host once per batch of steps,
// outside the graph. The binary32 level buffers' ping-pong parities are the same at the start of
// every step (each level runs an even number of launches per cycle), so one graph per binary64
// parity suffices.
void drop_graphs(cg_plan *p) {
    for (int i = 0; i < 2; ++i) {
        if (p->step_exec[i]) (void)hipGraphExecDestroy(p->step_exec[i]);
        if (p->net_exec[i]) (void)hipGraphExecDestroy(p->net_exec[i]);
        if (p->step_graph[i]) (void)hipGraphDestroy(p->step_graph[i]);
        p->step_exec[i] = p->net_exec[i] = nullptr;
        p->step_graph[i] = nullptr;
    }
}

// The step graphs of the plain cycle, or with net_on those of the learned one.
int build_step_graphs(cg_plan *p) {
    cg_ctx *ctx = p->ctx;
    // The settings that shape a step, the context's tile rows among them (grids the plan has no
    // tuned height for take theirs); graphs captured under others are rebuilt.
    const long key = ((long)ctx->level_passes << 12) | ((long)ctx->tile_rows << 4) | (ctx->fused << 3) |
                     (ctx->fuse_outer << 2) | (ctx->direct_bottom << 1) | (p->net_on ? ctx->net_fallback : 0);
    hipGraphExec_t *exec = p->net_on ? p->net_exec : p->step_exec;
    long &gkey = p->net_on ? p->net_graph_key : p->graph_key;
    double &getol = p->net_on ? p->net_graph_etol : p->graph_etol;
    if (gkey != key || getol != ctx->error_tol) {
        getol = ctx->error_tol;
        for (int i = 0; i < 2; ++i) {
            if (exec[i]) {
                (void)hipGraphExecDestroy(exec[i]);
                exec[i] = nullptr;
            }
            if (!p->net_on && p->step_graph[i]) {
                (void)hipGraphDestroy(p->step_graph[i]);
                p->step_graph[i] = nullptr;
            }
        }
        gkey = key;
    }
    for (int par = 0; par < 2; ++par) {
        if (exec[par]) continue;
        int saved_x64 = p->x64cur;
        int saved_cur[kMaxLevels];
        for (int l = 0; l < p->n; ++l) saved_cur[l] = p->lv[l].cur;
        p->x64cur = par;
        hipGraph_t g = nullptr;
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = refine_step(p, false);
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        for (int l = 0; l < p->n; ++l) {
            if (p->lv[l].cur != saved_cur[l]) {
                set_err(ctx, "level %d ping-pong parity changed over a cycle", l);
                rc = rc ? rc : CG_ERR_STATE;
            }
            p->lv[l].cur = saved_cur[l];
        }
        p->x64cur = saved_x64;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            set_err(ctx, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
            return CG_ERR_DEVICE;
        }
        e = hipGraphInstantiate(&exec[par], g, nullptr, nullptr, 0);
        // The plain step's graph is kept: a group of plans composes its steps from them.
        if (!p->net_on && e == hipSuccess) {
            if (p->step_graph[par]) (void)hipGraphDestroy(p->step_graph[par]);
            p->step_graph[par] = g;
        } else {
            (void)hipGraphDestroy(g);
        }
        if (e != hipSuccess) {
            set_err(ctx, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            exec[par] = nullptr;
            return CG_ERR_DEVICE;
        }
    }
    return CG_OK;
}

// Sets up the step gate for a refinement whose first graph step reads binary64 parity first_par:
// every instance active with no steps run, residual thresholds thr[i] (absolute: tol x max |f c|),
// the error bound etol and the fewest steps min_steps of every instance, and every instance's
// residual and sums groups zeroed (the first outer step reduces into parity 1 - first_par's, which
// the first gate reads). Bounds and counts live in device memory beside the thresholds, so one
// step graph serves every setting.
int gate_reset(cg_plan *p, int first_par, const std::vector<double> &thr, double etol, int min_steps,
               const std::vector<char> *only = nullptr) {
    cg_ctx *ctx = p->ctx;
    const int K = p->K;
    // They cross the bus only when they change (a new forcing, tol or method).
    std::vector<double> up(thr);
    up.resize(2 * K, etol);
    up.resize(3 * K, (double)min_steps);
    if (p->thr_up != up) {
        for (int i = 0; i < 3 * K; ++i) p->h_thr[i] = up[i];
        HIPCHK(ctx, hipMemcpyAsync(p->d_thr, p->h_thr, sizeof(double) * 3 * K, hipMemcpyHostToDevice, ctx->stream));
        p->thr_up = up;
    }
    const long nclear = (long)2 * K * kMaxSlots + (long)4 * K * kMaxSlots;  // residual groups, then sums
    if (only) {
        for (int i = 0; i < K; ++i) p->h_active[i] = (*only)[i] ? 1 : 0;
        HIPCHK(ctx, hipMemcpyAsync(p->d_active, p->h_active, sizeof(int) * K, hipMemcpyHostToDevice, ctx->stream));
    }
    hipLaunchKernelGGL(k_check_init, dim3(blocks_for(nclear, 256)), dim3(256), 0, ctx->stream,
                       p->d_gate + (size_t)(1 - first_par) * K, p->d_active, p->d_want, p->d_res, nclear, K,
                       only ? p->d_active : nullptr);
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

// One refined solve's host side on one plan: thresholds, the gate's view of each instance, the
// batch prediction, the results. refined_solve_core runs one; a group of plans runs several, one
// step of every plan per launch.
struct RefineRun {
    cg_plan *p = nullptr;
    int K = 0;
    double tol = 0.0, etol = 0.0;
    int min_steps = 0;
    bool cold = true, learned = false;
    int first_par = 0;
    const std::vector<char> *only = nullptr;
    std::vector<double> largest, thr, res_prev, rate;
    std::vector<char> done;
    int launched = 0;    // steps launched on this plan
    int last_batch = 0;  // steps of its last batch

    // What the host knows of instance i after `launched` steps: the state the last gate wrote (or
    // the initial one), and the gate's view of the last step the instance ran.
    StepGate state_of(int i, int launched) const {
        const int par = launched == 0 ? 1 - first_par : (first_par + launched - 1) & 1;
        return p->h_gate[(size_t)par * K + i];
    }
    GateView view_of(int i, const StepGate &sg) const {
        const int g = (first_par + sg.steps + 1) & 1;  // the group its last step (or the first outer) reduced into
        return gate_view(p->h_res + (size_t)(2 * i + g) * kMaxSlots, p->h_sums + (size_t)(2 * i + g) * 2 * kMaxSlots,
                         sg, thr[i], etol, min_steps);
    }
    // Enqueues the copy of the gate states of both parities, the residual groups and the sums of
    // every instance (the caller synchronizes the stream).
    int read_async() const {
        HIPCHK(p->ctx, hipMemcpyAsync(p->h_check, p->d_check, p->check_bytes, hipMemcpyDeviceToHost, p->ctx->stream));
        return CG_OK;
    }
};

// Starts a refined solve on plan p: thresholds, the gate, the first outer step (from zero, or from
// the plan's answer when warm); the plan's steps gated (and learned) from here on, its step graphs
// built when `graph`.
int run_begin(RefineRun &r, cg_plan *p, double tol_asked, bool warm, bool learned, const std::vector<char> *only,
              bool graph, double error_share = 1.0, int min_steps = 0) {
    cg_ctx *ctx = p->ctx;
    r.p = p;
    r.K = p->K;
    r.tol = tol_asked * (ctx->stop_permille / 1000.0);
    r.etol = ctx->error_tol * error_share;
    r.min_steps = min_steps;
    r.learned = learned;
    r.only = only;
    const int K = p->K;
    const Stencil *st = p->lv[0].st;
    r.largest.assign(K, 0.0);
    r.thr.assign(K, 0.0);
    int rc;
    for (int i = 0; i < K; ++i) {
        if ((rc = plan_largest_at(p, i, &r.largest[i]))) return rc;
        r.thr[i] = r.largest[i] > 0.0 ? r.tol * r.largest[i] : INFINITY;
    }
    r.cold = !warm || !p->x64_valid;
    p->x64_valid = true;
    // The first outer step reads parity q0 (zero, from cold) and writes 1 - q0, where the steps start.
    const int q0 = p->x64cur;
    r.first_par = 1 - q0;
    if ((rc = gate_reset(p, r.first_par, r.thr, r.etol, r.min_steps, only))) return rc;
    float *const first_fc = outer_fused(p) ? level0_forcing(p, r.first_par) : p->lv[0].f;
    if (r.cold) {
        // From zero: x = 0, u = 0 - (0 - f) on water; its maxima in parity q0's groups.
        MaxReduce red;
        red.out = p->d_res + (size_t)q0 * kMaxSlots;
        red.clear = nullptr;
        OuterBatch ob{};
        ob.s32 = vstride(st);
        ob.s64 = st->cells;
        ob.red_stride = 2 * kMaxSlots;
        const dim3 grid(blocks_for(st->across / 2, 256), blocks_for(st->down, kRowsPerBlock), K);
        hipLaunchKernelGGL(k_first_outer, grid, dim3(256), 0, ctx->stream, st->args(), p->f0, p->x64[r.first_par],
                           first_fc, red, ob);
        HIPCHK(ctx, hipGetLastError());
        p->x64cur = r.first_par;
    } else if ((rc = outer_step(p, nullptr, first_fc))) {
        return rc;
    }
    p->gated = true;
    p->net_on = learned;
    if (learned) p->net_fallbacks = 0;
    if (graph && (rc = build_step_graphs(p))) return rc;
    // From zero the relative residual before the first step is 1 (the zero answer's), which dates the
    // first batch's rate.
    r.res_prev.assign(K, r.cold ? 1.0 : -1.0);
    r.rate.assign(K, -1.0);
    r.done.assign(K, 0);
    if (only)
        for (int i = 0; i < K; ++i) r.done[i] = !(*only)[i];
    return CG_OK;
}

// The steps run `r` wants next: one fewer than the most steps any of its instances still running
// is predicted to need by its residual's rate and its correction's contraction (the device stops
// each one at its own step, so a batch sized for the slowest costs the others only skipped
// launches), at most 16; 8 before any rate is known. The last read of each instance replays its
// next gate, so a solve ends without launching the step that would only stop it.
int run_want(RefineRun &r, int launched, int last_batch) {
    cg_plan *p = r.p;
    int batch = 0;
    if (launched == 0 && r.cold) {
        // From zero the residual is max |f c| itself: every instance with a forcing needs steps,
        // about as many as its last solve from zero took (one fewer, so the batch ends before the
        // stop), else 8.
        for (int i = 0; i < r.K; ++i) {
            if (r.done[i]) continue;
            if (r.largest[i] > 0.0) batch = std::max(batch, p->last_steps[i] > 1 ? p->last_steps[i] - 1 : 8);
            else r.done[i] = 1;
        }
        return std::min(batch, 32);
    }
    for (int i = 0; i < r.K; ++i) {
        if (r.done[i]) continue;
        const StepGate sg = r.state_of(i, launched);
        const GateView v = r.view_of(i, sg);
        if (v.stop) {
            r.done[i] = 1;
            continue;
        }
        const double rr = r.largest[i] > 0.0 ? v.res / r.largest[i] : 0.0;
        if (last_batch > 0 && r.res_prev[i] > 0.0 && rr > 0.0) r.rate[i] = std::pow(rr / r.res_prev[i], 1.0 / last_batch);
        r.res_prev[i] = rr;
        int need_res = 1, need_err = 0;
        if (v.res < r.thr[i]) need_res = 0;
        else if (r.rate[i] > 0.0 && r.rate[i] < 0.95)
            need_res = (int)std::ceil(std::log(r.tol / rr) / std::log(r.rate[i]));
        if (r.etol > 0.0) {
            if (v.est < 0.0) need_err = need_res + 1;
            else if (v.est > r.etol && v.rho < 0.95)
                need_err = (int)std::ceil(std::log(r.etol / v.est) / std::log(v.rho));
            else if (v.est > r.etol) need_err = 1;
        }
        // One step short of the prediction: the early steps' rate is slower than the last ones',
        // so predictions run long, and a step past the stop costs more than a read.
        const int want = std::max(1, std::max(need_res, need_err) - 1);
        batch = std::max(batch, want);
    }
    return std::min(batch, 16);
}

// After a read: marks the instances that have stopped; returns how many still run.
int run_left(RefineRun &r, int launched) {
    int left = 0;
    for (int i = 0; i < r.K; ++i) {
        if (r.done[i]) continue;
        const StepGate sg = r.state_of(i, launched);
        if (!sg.active || r.view_of(i, sg).stop) r.done[i] = 1;
        else ++left;
    }
    return left;
}

void run_finish(RefineRun &r, int launched, int status, cg_solution *infos) {
    cg_plan *p = r.p;
    for (int i = 0; i < r.K; ++i) {
        if (r.only && !(*r.only)[i]) continue;
        const StepGate sg = r.state_of(i, launched);
        const GateView v = r.view_of(i, sg);
        infos[i] = cg_solution{};
        infos[i].cycles = sg.steps;
        infos[i].refinements = sg.steps + 1;
        infos[i].relative_residual = r.largest[i] > 0.0 ? v.res / r.largest[i] : 0.0;
        infos[i].status = status;
        infos[i].refinements = sg.fell;  // the caller (refined_solve) reads and resets it
        p->ans_par[i] = (r.first_par + sg.steps) & 1;
        if (r.cold && status == CG_OK && !r.learned) p->last_steps[i] = sg.steps;
    }
    if (r.K == 1) p->x64cur = p->ans_par[0];
    p->last_answer = 1;
}

// Steps the runs (plans of one context) together until every instance of every run has stopped:
// each batch is as long as the longest any run wants, and `step(on)` enqueues one refinement step
// of every run k with on[k] (the runs with an instance still running). `join` orders the
// context's stream after every launch before the gates are read. Returns the status.
template <typename Step, typename Join>
int run_steps(std::vector<RefineRun> &runs, const volatile int *cancel, Step step, Join join, int *status_out) {
    cg_ctx *ctx = runs[0].p->ctx;
    auto read_back = [&]() -> int {
        int rc = join();
        if (rc) return rc;
        for (const RefineRun &r : runs) {
            rc = r.read_async();
            if (rc) return rc;
        }
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return CG_OK;
    };
    int status = CG_OK, rc;
    // A warm start reads its first outer step's residual before any step: it may already be under tol.
    bool warm = false;
    for (const RefineRun &r : runs) warm = warm || !r.cold;
    if (warm && (rc = read_back())) return rc;
    std::vector<char> on(runs.size(), 0);
    int total = 0;
    while (total < kRefinedMostSteps) {
        int batch = 0;
        for (size_t k = 0; k < runs.size(); ++k) {
            const int want = run_want(runs[k], runs[k].launched, runs[k].last_batch);
            on[k] = want > 0;
            batch = std::max(batch, want);
        }
        if (batch <= 0) break;  // every instance's next gate would stop it
        if (cancel && *cancel) {
            status = CG_CANCELLED;
            break;
        }
        batch = std::min(batch, kRefinedMostSteps - total);
        for (int k = 0; k < batch; ++k)
            if ((rc = step(on))) return rc;
        total += batch;
        int left = 0;
        for (size_t k = 0; k < runs.size(); ++k) {
            runs[k].last_batch = on[k] ? batch : 0;
            if (on[k]) runs[k].launched += batch;
        }
        if ((rc = read_back())) return rc;
        for (size_t k = 0; k < runs.size(); ++k)
            if (on[k]) left += run_left(runs[k], runs[k].launched);
        if (left == 0) break;
    }
    if (total == 0 && (rc = read_back())) return rc;
    *status_out = status;
    return CG_OK;
}

// The refined method over every instance of the plan; infos[i] receives instance i's result.
// Every step starts with the gate (gate_decide) that stops an instance at the first step where its
// residual is under tol and its estimated error under the context's error bound, so each instance
// runs exactly the steps it would alone, however the host groups the steps into batches between
// its reads; a stopped instance's kernels return at once and its answer stays where it stopped.
int refined_solve_core(cg_plan *p, double tol_asked, bool warm, cg_solution *infos, const volatile int *cancel,
                       bool learned, const std::vector<char> *only, double error_share = 1.0, int min_steps = 0) {
    cg_ctx *ctx = p->ctx;
    const bool graph = ctx->use_graph && !ctx->time_kernels;
    struct Ungate {
        cg_plan *p;
        ~Ungate() {
            p->gated = false;
            p->net_on = false;
        }
    } ungate{p};
    std::vector<RefineRun> runs(1);
    int rc = run_begin(runs[0], p, tol_asked, warm, learned, only, graph, error_share, min_steps);
    if (rc) return rc;
    auto step = [&](const std::vector<char> &) -> int {
        if (graph) {
            hipGraphExec_t *const exec = p->net_on ? p->net_exec : p->step_exec;
            HIPCHK(ctx, hipGraphLaunch(exec[p->x64cur], ctx->stream));
            p->x64cur = 1 - p->x64cur;
            return CG_OK;
        }
        return refine_step(p, true);
    };
    int status = CG_OK;
    if ((rc = run_steps(runs, cancel, step, []() { return CG_OK; }, &status))) return rc;
    run_finish(runs[0], runs[0].launched, status, infos);
    return CG_OK;
}

// After a learned solve's steps (infos[i].refinements holding whether instance i's gate stopped it
// for its stalling residual: a fallback): plain multigrid judges every answer. Each instance goes on
// from its learned answer (or, if one is not a number, the whole solve starts again plainly) with the
// plain cycle, at least kJudgeSteps steps and then until the plain method's residual and error rule
// holds under kJudgeErrorShare of the bound; the steps of both count. net_fallbacks counts the
// instances whose residual stalled.
int learned_fallback(cg_plan *p, double tol, cg_solution *infos, const volatile int *cancel) {
    cg_ctx *ctx = p->ctx;
    const int K = p->K;
    int rc;
    std::vector<char> judged(K, 0);
    bool bad = false;
    int stalled = 0;
    for (int i = 0; i < K; ++i) {
        stalled += infos[i].refinements != 0 && infos[i].status == CG_OK;
        judged[i] = infos[i].status == CG_OK;
        bad = bad || !(infos[i].relative_residual == infos[i].relative_residual);
        infos[i].refinements = infos[i].cycles + 1;
    }
    if (bad && ctx->net_fallback) {
        p->net_fallbacks = K;
        rc = refined_solve_core(p, tol, false, infos, cancel, false, nullptr);
        for (int i = 0; i < K; ++i) infos[i].refinements = infos[i].cycles + 1;
        return rc;
    }
    p->net_fallbacks = stalled;
    if (!ctx->net_fallback) return CG_OK;  // the network measured alone: no plain steps at all
    // Each answer into the parity the plain steps read, then those steps from the answers.
    const long n = p->lv[0].st->cells;
    const int q = p->x64cur;
    bool any = false;
    for (int i = 0; i < K; ++i) {
        if (!judged[i]) continue;
        any = true;
        if (p->ans_par[i] != q)
            HIPCHK(ctx, hipMemcpyAsync(p->x64[q] + i * n, p->x64[p->ans_par[i]] + i * n, sizeof(double) * n,
                                       hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (!any) return CG_OK;
    std::vector<cg_solution> more(infos, infos + K);
    if ((rc = refined_solve_core(p, tol, true, more.data(), cancel, false, &judged, kJudgeErrorShare, kJudgeSteps)))
        return rc;
    for (int i = 0; i < K; ++i) {
        if (!judged[i]) continue;
        infos[i].cycles += more[i].cycles;
        infos[i].relative_residual = more[i].relative_residual;
        infos[i].status = more[i].status;
        infos[i].refinements = infos[i].cycles + 1;
    }
    return CG_OK;
}

// The refined method, or with `learned` its learned coarse-grid correction and the plain judge of its
// answers, over every instance.
int refined_solve(cg_plan *p, double tol, bool warm, cg_solution *infos, const volatile int *cancel,
                  bool learned = false) {
    const int K = p->K;
    if (learned) p->net_fallbacks = 0;
    int rc = refined_solve_core(p, tol, warm, infos, cancel, learned, nullptr, learned ? 0.0 : 1.0);
    if (rc || !learned) {
        for (int i = 0; i < K; ++i) infos[i].refinements = infos[i].cycles + 1;
        return rc;
    }
    return learned_fallback(p, tol, infos, cancel);
}

// The binary64 refinement whose inner solve is FGMRES(m) right-preconditioned by the binary32
// V-cycle. Each outer step: m inner iterations from the current residual (q_0 = u, the correction
// forcing the outer kernel writes), delta from the least-squares combination, then the outer
// binary64 residual. m is the context's restart length for the first step, then the number of
// iterations the observed rate predicts are still needed (plus one), within 2..restart.
// learned: the first inner iteration of each outer step takes the caller's preconditioner instead
// of the cycle (it sees the outer residual itself, the kind of input it was trained on; the cycles
// of the later iterations and the least-squares step build on its direction), until an outer step
// fails to halve the residual; from then on every iteration is a plain cycle.
int fgmres_solve(cg_plan *p, double tol_asked, bool warm, cg_solution *info, const volatile int *cancel,
                 bool learned = false) {
    cg_ctx *ctx = p->ctx;
    const double tol = tol_asked * (ctx->stop_permille / 1000.0);
    hipStream_t s = ctx->stream;
    const Stencil *st = p->lv[0].st;
    const int M = std::min(ctx->fgmres_m, kFgmresMax);
    int rc;
    if (!p->fq) {
        p->vstride = (long)st->across * (st->down + 1);
        if ((rc = dalloc(ctx, &p->fq, (size_t)(kFgmresMax + 1) * p->vstride))) return rc;
        if ((rc = dalloc(ctx, &p->fz, (size_t)kFgmresMax * p->vstride))) return rc;
        if ((rc = dalloc(ctx, &p->fa, st->cells))) return rc;
        if ((rc = dalloc(ctx, &p->fdelta, st->cells))) return rc;
        const dim3 g = grid2d(st);
        if ((rc = dalloc(ctx, &p->fpart, (size_t)(kFgmresMax + 1) * g.x * g.y))) return rc;
        if ((rc = dalloc(ctx, &p->fst, 1))) return rc;
        HIPCHK(ctx, hipMemset(p->fq, 0, sizeof(float) * (kFgmresMax + 1) * p->vstride));
        HIPCHK(ctx, hipMemset(p->fz, 0, sizeof(float) * kFgmresMax * p->vstride));
        HIPCHK(ctx, hipMemset(p->fst, 0, sizeof(FgmresState)));
    }
    double largest;
    if ((rc = plan_largest(p, &largest))) return rc;
    if (!warm || !p->x64_valid) HIPCHK(ctx, hipMemsetAsync(p->x64[p->x64cur], 0, sizeof(double) * st->cells, s));
    p->x64_valid = true;
    hipLaunchKernelGGL(k_fill_u64, dim3(2), dim3(kMaxSlots), 0, s, p->d_res, 2 * kMaxSlots, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(4), dim3(kMaxSlots), 0, s, reinterpret_cast<unsigned long long *>(p->d_sums),
                       4 * kMaxSlots, 0ull);
    int slot = p->x64cur;
    if ((rc = outer_step(p, nullptr, p->fq))) return rc;
    unsigned long long bits;
    if ((rc = read_res(p, slot, &bits))) return rc;
    double res = largest > 0.0 ? bits_to_double(bits) / largest : 0.0;
    const dim3 g2 = grid2d(st);
    const int nb = (int)(g2.x * g2.y);
    int cycles = 0, outer = 0;
    double rate = -1.0;
    bool net_on = learned;
    // The error estimate of the refined method's gate, per outer step: its correction's norm against
    // the answer's, times rho / (1 - rho) for the contraction of successive corrections.
    const double etol = ctx->error_tol;
    double d2prev = 0.0, d2prev2 = 0.0;
    bool err_ok = true;  // no correction applied yet
    while ((res >= tol || !err_ok) && cycles < kRefinedMostSteps) {
        if (cancel && *cancel) {
            info->status = CG_CANCELLED;
            break;
        }
        int m = M;
        if (rate > 0.0 && rate < 0.95 && res > 0.0)
            m = (int)std::max(2.0, std::min((double)M, std::ceil(std::log(tol / res) / std::log(rate)) + 1.0));
        hipLaunchKernelGGL(k_dot_partials, g2, dim3(256), 0, s, st->across, st->down, p->fq, p->fq, p->fpart);
        hipLaunchKernelGGL(k_fgmres_start, dim3(1), dim3(256), 0, s, p->fpart, nb, p->fst);
        for (int j = 0; j < m; ++j) {
            float *qj = p->fq + (long)j * p->vstride;
            float *zj = p->fz + (long)j * p->vstride;
            if (net_on && j == 0) {
                const int prc = p->pre_fn(p->pre_user, qj, zj, (void *)s);
                if (prc != CG_OK) {
                    set_err(ctx, "the learned preconditioner returned %d", prc);
                    return prc < 0 ? prc : CG_ERR_STATE;
                }
            } else if ((rc = vcycle32(p, false, qj, zj))) {
                return rc;
            }
            hipLaunchKernelGGL(k_apply_dots, g2, dim3(256), 0, s, st->args(), zj, p->fa, p->fq, p->vstride, j, p->fpart);
            hipLaunchKernelGGL(k_fgmres_dots, dim3(1), dim3(256), 0, s, p->fpart, nb, j, p->fst);
            hipLaunchKernelGGL(k_orth, g2, dim3(256), 0, s, st->across, st->down, p->fa, p->fq, p->vstride, j, p->fst,
                               p->fpart);
            hipLaunchKernelGGL(k_fgmres_norm, dim3(1), dim3(256), 0, s, p->fpart, nb, j, p->fst);
        }
        hipLaunchKernelGGL(k_fgmres_solve, dim3(1), dim3(64), 0, s, m, p->fst);
        hipLaunchKernelGGL(k_combine, dim3(blocks_for(st->cells, 256)), dim3(256), 0, s, st->cells, p->fz, p->vstride,
                           m, p->fst, p->fdelta);
        HIPCHK(ctx, hipGetLastError());
        slot = p->x64cur;
        if ((rc = outer_step(p, p->fdelta, p->fq))) return rc;
        if ((rc = read_res(p, slot, &bits))) return rc;
        {
            double sums[2 * kMaxSlots];
            HIPCHK(ctx, hipMemcpy(sums, sums_group(p, slot), sizeof sums, hipMemcpyDeviceToHost));
            const double d2 = butterfly_sum(sums), x2 = butterfly_sum(sums + kMaxSlots);
            if (etol <= 0.0) {
                err_ok = true;
            } else if (x2 > 0.0) {
                double rho = d2prev > 0.0 ? std::sqrt(d2 / d2prev) : 0.5;
                if (d2prev > 0.0 && d2prev2 > 0.0) rho = std::max(rho, std::sqrt(d2prev / d2prev2));
                rho = rho < 0.95 ? rho : 0.95;
                err_ok = std::sqrt(d2 / x2) * (rho / (1.0 - rho)) <= etol;
            } else {
                err_ok = d2 <= 0.0;
            }
            d2prev2 = d2prev;
            d2prev = d2;
        }
        cycles += m;
        outer++;
        const double prev = res;
        const double r = bits_to_double(bits);
        res = largest > 0.0 ? r / largest : 0.0;
        if (std::isnan(r)) break;
        if (prev > 0.0 && res > 0.0) rate = std::pow(res / prev, 1.0 / m);
        if (net_on && !(res <= 0.5 * prev)) net_on = false;  // plain cycles from here on
    }
    info->cycles = cycles;
    info->refinements = outer + 1;
    info->relative_residual = res;
    p->last_answer = 1;
    return CG_OK;
}

// ------------------------------------------------------------------------------------------------
// A group's merged step: the refinement step of every plan still running in one graph, each tile
// launch covering the grids of one depth of those plans' ladders (k_relax_tile_jobs), then every
// plan's direct bottom in one launch, then the up legs the same way, the finest one with the binary64
// outer step fused in. One graph per set of running plans (a plan whose instances have all stopped
// drops out, as it gets no more launches on the streams path); each instance stops at its own step
// by its gate.
// ------------------------------------------------------------------------------------------------

// Every plan fuses the outer step into its finest up leg (two passes there), ends its cycle at the
// direct bottom and runs two or four passes a leg on its other grids, and the learned correction is
// off.
bool group_mergeable(const cg_group *gr) {
    const cg_ctx *ctx = gr->ctx;
    if (!ctx->group_merged || !ctx->use_graph || ctx->time_kernels) return false;
    for (const cg_plan *p : gr->plans) {
        if (!outer_fused(p) || !direct_bottom(p) || p->net_on) return false;
        for (int l = 1; l < p->bottom; ++l)
            if (level_half_sweeps(ctx, l) != 4 && level_half_sweeps(ctx, l) != 8) return false;
    }
    return true;
}

// The merged launches' tile height at depth l: that of the group's plan with the largest grid there
// (its tuned height, else the context's rule), within the job kernels' 2..8. Tiles only partition the
// work, so the answers do not depend on it.
int merged_rows(const cg_group *gr, int l) {
    const cg_plan *best = nullptr;
    for (const cg_plan *p : gr->plans)
        if (p->bottom > l && (!best || p->lv[l].st->cells > best->lv[l].st->cells)) best = p;
    const int r = best ? plan_tile_rows(best, l) : 8;
    for (int k : kJobRows)
        if (r <= k) return k;
    return 8;
}

std::vector<long> merged_key(const cg_group *gr) {
    const cg_ctx *ctx = gr->ctx;
    std::vector<long> k{ctx->level_passes, ctx->tile_rows, ctx->fused, ctx->fuse_outer, ctx->direct_bottom,
                        (long)std::llround(ctx->error_tol * 1e15)};
    for (int l = 0; l < kMaxLevels; ++l) k.push_back(merged_rows(gr, l));
    return k;
}

void drop_merged(cg_group *gr) {
    for (auto &kv : gr->merged) {
        for (hipGraphExec_t &e : kv.second.exec)
            if (e) (void)hipGraphExecDestroy(e);
        dfree(kv.second.d_jobs);
        dfree(kv.second.d_djobs);
    }
    gr->merged.clear();
}

// The merged step of the plans in `mask` (after run_begin has gated every plan, with the stream idle):
// built for both binary64 parities on first use, the job tables filled on the host while the launches
// that read them are captured and copied to the device before the graphs first run.
int merged_for(cg_group *gr, unsigned mask, cg_group::Merged **out) {
    cg_ctx *ctx = gr->ctx;
    auto it = gr->merged.find(mask);
    if (it != gr->merged.end()) {
        *out = &it->second;
        return CG_OK;
    }
    // The sets of running plans that solves meet are few (each solve's sets shrink one plan at a
    // time); a cache that has grown past kMaxMergedSets is emptied, with no step of it in flight.
    if (gr->merged.size() >= kMaxMergedSets) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        drop_merged(gr);
    }
    std::vector<cg_plan *> plans;
    for (size_t k = 0; k < gr->plans.size(); ++k)
        if (mask >> k & 1u) plans.push_back(gr->plans[k]);
    const size_t P = plans.size();
    int maxB = 0;
    for (const cg_plan *p : plans) maxB = std::max(maxB, p->bottom);
    const size_t tables = 2 * 2 * (size_t)maxB;  // per parity: maxB down launches, maxB up launches
    cg_group::Merged m;
    int rc = dalloc(ctx, &m.d_jobs, tables * P);
    if (!rc) rc = dalloc(ctx, &m.d_djobs, 2 * P);
    if (rc) {
        dfree(m.d_jobs);
        return rc;
    }
    auto fail = [&](int code) {
        for (hipGraphExec_t &e : m.exec)
            if (e) (void)hipGraphExecDestroy(e);
        dfree(m.d_jobs);
        dfree(m.d_djobs);
        return code;
    };
    std::vector<TileJob> host(tables * P);
    std::vector<DirectJob> dhost(2 * P);
    for (int par = 0; par < 2; ++par) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_err(ctx, "hipStreamBeginCapture failed");
            return fail(CG_ERR_DEVICE);
        }
        std::vector<int> saved_x64(P);
        std::vector<std::vector<int>> saved_cur(P);
        for (size_t k = 0; k < P; ++k) {
            cg_plan *p = plans[k];
            saved_x64[k] = p->x64cur;
            for (int l = 0; l < p->n; ++l) saved_cur[k].push_back(p->lv[l].cur);
            p->x64cur = par;
            p->gate_in_down = p->gated;
        }
        hipError_t e = hipSuccess;
        size_t t = (size_t)par * 2 * maxB;
        // One launch of the given legs (each plan's instances, its own grid and ladder).
        auto emit = [&](const std::vector<LegSpec> &specs, int rows, int mode, int nh) {
            TileJob *tab = host.data() + t * P;
            const TileJob *dtab = m.d_jobs + t * P;
            ++t;
            const int iw = jobs_tile_width(mode, nh);
            int first = 0, n = 0;
            for (const LegSpec &sp : specs) {
                const long waves = (long)((sp.L.across + iw - 1) / iw) * ((sp.L.down + rows - 1) / rows);
                const int per_inst = (int)((waves + 4 * kXcds - 1) / (4 * kXcds)) * kXcds;
                TileJob &jb = tab[n++];
                jb.L = sp.L;
                jb.f = sp.f;
                jb.xin = sp.xin;
                jb.xout = sp.xout;
                jb.ex = sp.ex;
                jb.first = first;
                jb.per_inst = per_inst;
                jb.count = per_inst * sp.K;
                first += jb.count;
            }
            if (n > 0 && e == hipSuccess) e = launch_jobs(ctx->stream, mode, nh, rows, dtab, n, (unsigned)first);
        };
        for (int l = 0; l < maxB; ++l) {
            std::vector<LegSpec> specs;
            const int rows = merged_rows(gr, l);
            for (cg_plan *p : plans)
                if (p->bottom > l) {
                    specs.push_back(down_leg_spec(p, l, l == 0 ? level0_forcing(p, p->x64cur) : nullptr, rows));
                    p->lv[l].cur = 1 - p->lv[l].cur;
                }
            emit(specs, rows, TILE_DOWN, level_half_sweeps(ctx, l));
        }
        {
            // Every plan's direct bottom in one launch, four instances a workgroup row.
            DirectJob *tab = dhost.data() + (size_t)par * P;
            int first = 0;
            for (size_t k = 0; k < P; ++k) {
                cg_plan *p = plans[k];
                tab[k].a = direct_args(p);
                tab[k].first = first;
                tab[k].rows = (p->bnw + kDirectWaves - 1) / kDirectWaves;
                tab[k].count = tab[k].rows * ((p->K + 3) / 4);
                first += tab[k].count;
            }
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_direct_bottom_jobs<4>, dim3(first), dim3(64 * kDirectWaves), 0, ctx->stream,
                                   m.d_djobs + (size_t)par * P, (int)P);
                e = hipGetLastError();
            }
        }
        for (int l = maxB - 1; l >= 0; --l) {
            std::vector<LegSpec> specs;
            const int rows = merged_rows(gr, l);
            for (cg_plan *p : plans)
                if (p->bottom > l) {
                    specs.push_back(up_leg_spec(p, l, l == 0 ? level0_forcing(p, p->x64cur) : nullptr, nullptr, true,
                                                rows));
                    p->lv[l].cur = 1 - p->lv[l].cur;
                }
            emit(specs, rows, l == 0 ? TILE_UP_OUTER : TILE_UP, level_half_sweeps(ctx, l));
        }
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        for (size_t k = 0; k < P; ++k) {
            cg_plan *p = plans[k];
            p->x64cur = saved_x64[k];
            for (int l = 0; l < p->n; ++l) p->lv[l].cur = saved_cur[k][l];
            p->gate_in_down = false;
        }
        if (e != hipSuccess || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            set_err(ctx, "capturing the group's merged step failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : ec));
            return fail(CG_ERR_DEVICE);
        }
        e = hipGraphInstantiate(&m.exec[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            m.exec[par] = nullptr;
            set_err(ctx, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            return fail(CG_ERR_DEVICE);
        }
    }
    if (hipMemcpy(m.d_jobs, host.data(), sizeof(TileJob) * host.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m.d_djobs, dhost.data(), sizeof(DirectJob) * dhost.size(), hipMemcpyHostToDevice) != hipSuccess) {
        set_err(ctx, "copying the merged step's job tables failed");
        return fail(CG_ERR_DEVICE);
    }
    *out = &(gr->merged[mask] = m);
    return CG_OK;
}

}  // namespace

// ================================================================================================
// C ABI
// ================================================================================================

extern "C" {

int cg_ocean_open(int device, cg_ctx **out, char *reason, size_t reason_len) {
    auto fail = [&](const char *msg) {
        if (reason && reason_len) snprintf(reason, reason_len, "%s", msg);
        if (out) *out = nullptr;
        return CG_ERR_DEVICE;
    };
    if (!out) return CG_ERR_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail("no HIP device is visible");
    if (device < 0 || device >= count) return fail("no such HIP device");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail("the device's properties could not be read");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        char msg[256];
        snprintf(msg, sizeof msg, "device %d is %s, not gfx950 (MI355X)", device, prop.gcnArchName);
        return fail(msg);
    }
    if (hipSetDevice(device) != hipSuccess) return fail("hipSetDevice failed");
    // The scratch array's address on this device, looked up once per device (each device holds its
    // own copy of the array).
    static std::mutex scratch_mu;
    static std::map<int, float *> scratch_at;
    float *scratch_ptr = nullptr;
    {
        std::lock_guard<std::mutex> g(scratch_mu);
        auto it = scratch_at.find(device);
        if (it != scratch_at.end()) {
            scratch_ptr = it->second;
        } else {
            if (hipGetSymbolAddress((void **)&scratch_ptr, HIP_SYMBOL(scratch)) != hipSuccess || !scratch_ptr) {
                (void)hipGetLastError();
                return fail("the device address of the relaxation scratch could not be read");
            }
            scratch_at[device] = scratch_ptr;
        }
    }
    cg_ctx *ctx = new cg_ctx();
    ctx->device = device;
    ctx->scratch = scratch_ptr;
    ctx->name = prop.name;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail("hipStreamCreate failed");
    }
    if (hipHostMalloc((void **)&ctx->h_pinned, 4 * kMaxSlots * sizeof(unsigned long long), hipHostMallocDefault) !=
        hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail("hipHostMalloc failed");
    }
    *out = ctx;
    if (reason && reason_len) reason[0] = 0;
    return CG_OK;
}

void cg_ocean_close(cg_ctx *ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &kv : ctx->stencils) free_stencil(kv.second);
        ctx->stencils.clear();
        for (Stencil *st : ctx->orphans) free_stencil(st);
        ctx->orphans.clear();
        for (auto &t : ctx->timed) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        (void)hipHostFree(ctx->h_pinned);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int cg_ocean_name(cg_ctx *ctx, char *buf, size_t len) {
    if (!ctx || !buf || !len) return CG_ERR_ARG;
    snprintf(buf, len, "%s (HIP, gfx950)", ctx->name.c_str());
    return CG_OK;
}

const char *cg_ocean_last_error(cg_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int cg_ocean_set_option(cg_ctx *ctx, int option, long value) {
    if (!ctx) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    switch (option) {
        case CG_OPT_MIN_CELLS: ctx->min_cells = value < 0 ? 0 : value; return CG_OK;
        case CG_OPT_TILE_ROWS:
            if (value != 8 && value != 16 && value != 24 && value != 32) {
                set_err(ctx, "tile rows must be 8, 16, 24 or 32");
                return CG_ERR_ARG;
            }
            ctx->tile_rows = (int)value;
            return CG_OK;
        case CG_OPT_USE_GRAPH: ctx->use_graph = value ? 1 : 0; return CG_OK;
        case CG_OPT_FGMRES_RESTART:
            if (value < 1 || value > 64) {
                set_err(ctx, "FGMRES restart must be 1..64");
                return CG_ERR_ARG;
            }
            ctx->fgmres_m = (int)value;
            return CG_OK;
        case CG_OPT_TIME_KERNELS: ctx->time_kernels = value ? 1 : 0; return CG_OK;
        case CG_OPT_FUSED: ctx->fused = value ? 1 : 0; return CG_OK;
        case CG_OPT_FUSE_OUTER: ctx->fuse_outer = value ? 1 : 0; return CG_OK;
        case CG_OPT_DIRECT_BOTTOM: ctx->direct_bottom = value ? 1 : 0; return CG_OK;
        case CG_OPT_NET_FALLBACK: ctx->net_fallback = value ? 1 : 0; return CG_OK;
        case CG_OPT_GROUP_MERGED: ctx->group_merged = value ? 1 : 0; return CG_OK;
        case CG_OPT_ERROR_PICO:
            if (value < 0) {
                set_err(ctx, "the error bound must be >= 0 (in units of 1e-12; 0: the residual alone)");
                return CG_ERR_ARG;
            }
            ctx->error_tol = (double)value * 1e-12;
            return CG_OK;
        case CG_OPT_STOP_PERMILLE:
            if (value < 1 || value > 1000) {
                set_err(ctx, "the refined methods' stopping share must be 1..1000 per mille of tol");
                return CG_ERR_ARG;
            }
            ctx->stop_permille = (int)value;
            return CG_OK;
        case CG_OPT_LEVEL_PASSES:
            for (int l = 0; l < 8; ++l) {
                const long v = (value >> (4 * l)) & 0xf;
                if (v != 2 && v != 4 && v != 6) {
                    set_err(ctx, "level passes must be 2, 4 or 6 (level %d has %ld)", l, v);
                    return CG_ERR_ARG;
                }
            }
            ctx->level_passes = value;
            return CG_OK;
        default: set_err(ctx, "unknown option %d", option); return CG_ERR_ARG;
    }
}

long cg_ocean_get_option(cg_ctx *ctx, int option) {
    if (!ctx) return -1;
    switch (option) {
        case CG_OPT_MIN_CELLS: return ctx->min_cells;
        case CG_OPT_TILE_ROWS: return ctx->tile_rows;
        case CG_OPT_USE_GRAPH: return ctx->use_graph;
        case CG_OPT_FGMRES_RESTART: return ctx->fgmres_m;
        case CG_OPT_TIME_KERNELS: return ctx->time_kernels;
        case CG_OPT_FUSED: return ctx->fused;
        case CG_OPT_FUSE_OUTER: return ctx->fuse_outer;
        case CG_OPT_DIRECT_BOTTOM: return ctx->direct_bottom;
        case CG_OPT_STOP_PERMILLE: return ctx->stop_permille;
        case CG_OPT_NET_FALLBACK: return ctx->net_fallback;
        case CG_OPT_GROUP_MERGED: return ctx->group_merged;
        case CG_OPT_ERROR_PICO: return (long)std::llround(ctx->error_tol * 1e12);
        case CG_OPT_LEVEL_PASSES: return ctx->level_passes;
        default: return -1;
    }
}

int cg_ocean_stencil_upload(cg_ctx *ctx, int across, int down, const uint8_t *water, const float *east,
                            const float *west, const float *north, const float *south, const double *centre,
                            cg_stencil *out) {
    if (!ctx || !out) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    *out = 0;
    if (across <= 0 || down <= 0 || !water || !east || !west || !north || !south) {
        set_err(ctx, "stencil upload: bad sizes or null arrays");
        return CG_ERR_ARG;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    Stencil *s = new Stencil();
    s->across = across;
    s->down = down;
    s->cells = (long)across * down;
    s->has_centre = centre != nullptr;
    // Row-constant when every cell of a row carries the row's first weights, bit for bit.
    bool rc_w = true, rc_c = true;
    for (int r = 0; r < down && rc_w; ++r) {
        const long b = (long)r * across;
        for (int c = 1; c < across; ++c) {
            const long i = b + c;
            if (std::memcmp(&east[i], &east[b], 4) || std::memcmp(&west[i], &west[b], 4) ||
                std::memcmp(&north[i], &north[b], 4) || std::memcmp(&south[i], &south[b], 4)) {
                rc_w = false;
                break;
            }
        }
    }
    if (centre)
        for (int r = 0; r < down && rc_c; ++r) {
            const long b = (long)r * across;
            for (int c = 1; c < across; ++c)
                if (std::memcmp(&centre[b + c], &centre[b], 8)) {
                    rc_c = false;
                    break;
                }
        }
    // The tiled smoother also takes the weights to be finite and non-negative (an M-matrix, as the
    // circulation's are): its first half-sweep from zero then reduces to +0 - f exactly.
    for (int r = 0; r < down && rc_w; ++r) {
        const long b = (long)r * across;
        for (const float w : {east[b], west[b], north[b], south[b]})
            if (!(w >= 0.0f && w <= 3.4028235e38f)) rc_w = false;
    }
    s->rowconst = rc_w && rc_c;
    s->h_water.assign(water, water + s->cells);
    int rc = 0;
    auto bail = [&](int code) {
        free_stencil(s);
        return code;
    };
    // The device mask holds exactly 0 or 1 per cell (the smoother extracts bits from it).
    for (long i = 0; i < s->cells; ++i) s->h_water[i] = water[i] ? 1 : 0;
    if ((rc = dalloc_grid(ctx, &s->water, across, down))) return bail(rc);
    HIPCHK(ctx, hipMemcpy(s->water, s->h_water.data(), s->cells, hipMemcpyHostToDevice));
    if (s->rowconst) {
        s->h_w.resize(down);
        s->h_c.assign(down, 1.0);
        for (int r = 0; r < down; ++r) {
            const long b = (long)r * across;
            s->h_w[r] = make_float4(east[b], west[b], north[b], south[b]);
            if (centre) s->h_c[r] = centre[b];
        }
        if ((rc = dalloc(ctx, &s->roww, down))) return bail(rc);
        if ((rc = dalloc(ctx, &s->rowc, down))) return bail(rc);
        HIPCHK(ctx, hipMemcpy(s->roww, s->h_w.data(), sizeof(float4) * down, hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(s->rowc, s->h_c.data(), sizeof(double) * down, hipMemcpyHostToDevice));
    } else {
        s->h_w.resize(s->cells);
        s->h_c.assign(s->cells, 1.0);
        for (long i = 0; i < s->cells; ++i) {
            s->h_w[i] = make_float4(east[i], west[i], north[i], south[i]);
            if (centre) s->h_c[i] = centre[i];
        }
        if ((rc = dalloc(ctx, &s->cellw, s->cells))) return bail(rc);
        if ((rc = dalloc(ctx, &s->cellc, s->cells))) return bail(rc);
        HIPCHK(ctx, hipMemcpy(s->cellw, s->h_w.data(), sizeof(float4) * s->cells, hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(s->cellc, s->h_c.data(), sizeof(double) * s->cells, hipMemcpyHostToDevice));
    }
    const uint64_t h = ctx->next_handle++;
    ctx->stencils[h] = s;
    *out = h;
    return CG_OK;
}

int cg_ocean_stencil_release(cg_ctx *ctx, cg_stencil h) {
    if (!ctx) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    auto it = ctx->stencils.find(h);
    if (it == ctx->stencils.end()) {
        set_err(ctx, "no stencil %llu", (unsigned long long)h);
        return CG_ERR_ARG;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    Stencil *st = it->second;
    ctx->stencils.erase(it);
    if (st->plan_refs > 0) {
        st->released = true;
        ctx->orphans.insert(st);
    } else {
        free_stencil(st);
    }
    return CG_OK;
}

int cg_ocean_relax(cg_ctx *ctx, cg_stencil h, const float *forcing, const float *start, float *out, int passes,
                   const volatile int *cancel) {
    if (!ctx) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    Stencil *s = find_stencil(ctx, h);
    if (!s) {
        set_err(ctx, "no stencil %llu", (unsigned long long)h);
        return CG_ERR_ARG;
    }
    // The declines of GpuOcean.solve (GO:52-58).
    if (s->across <= 0 || s->across % 2 != 0 || s->down <= 0 || passes < 0) return CG_DECLINED;
    if (s->cells < ctx->min_cells) return CG_DECLINED;
    if (!forcing || !start || !out) {
        set_err(ctx, "relax: null array");
        return CG_ERR_ARG;
    }
    if (passes == 0) {
        if (out != start) std::memmove(out, start, sizeof(float) * s->cells);
        return CG_OK;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc;
    if (!s->fbuf) {
        if ((rc = dalloc_grid(ctx, &s->fbuf, s->across, s->down))) return rc;
        if ((rc = dalloc_grid(ctx, &s->xbuf[0], s->across, s->down))) return rc;
        if ((rc = dalloc_grid(ctx, &s->xbuf[1], s->across, s->down))) return rc;
    }
    HIPCHK(ctx, hipMemcpyAsync(s->fbuf, forcing, sizeof(float) * s->cells, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(s->xbuf[0], start, sizeof(float) * s->cells, hipMemcpyHostToDevice, ctx->stream));
    int cur = 0;
    // Cancellation is polled between launches of two passes, as GpuOcean polls per pass (GO:85-86).
    int left = passes;
    while (left > 0) {
        if (cancel && *cancel) {
            (void)hipStreamSynchronize(ctx->stream);
            return CG_CANCELLED;
        }
        const int now = std::min(left, 2);
        if ((rc = relax_pp(ctx, s, s->fbuf, s->xbuf, &cur, now, false, false))) return rc;
        left -= now;
    }
    HIPCHK(ctx, hipMemcpyAsync(out, s->xbuf[cur], sizeof(float) * s->cells, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CG_OK;
}

int cg_ocean_relax_noalloc(cg_ctx *ctx, cg_stencil h, const float *forcing, const float *start, float *out,
                           int passes, void *stream) {
    if (!ctx) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    Stencil *st = find_stencil(ctx, h);
    if (!st) {
        set_err(ctx, "no stencil %llu", (unsigned long long)h);
        return CG_ERR_ARG;
    }
    if (st->across <= 0 || st->across % 2 != 0 || st->down <= 0 || passes < 0) return CG_DECLINED;
    if (st->cells < ctx->min_cells) return CG_DECLINED;
    if (!forcing || !start || !out) {
        set_err(ctx, "relax_noalloc: null array");
        return CG_ERR_ARG;
    }
    if (st->cells > kScratchCells || st->across > kScratchAcross) {
        set_err(ctx, "relax_noalloc: %d x %d exceeds the scratch's %ld cells of at most %ld columns", st->across,
                st->down, kScratchCells, kScratchAcross);
        return CG_ERR_ARG;
    }
    if (!ctx->scratch) {
        set_err(ctx, "relax_noalloc: the context has no scratch address");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const long n = st->cells, total = n + st->across;
    if (passes == 0) {
        hipLaunchKernelGGL(k_copy_f32, dim3(blocks_for(n, 256)), dim3(256), 0, s, start, out, n);
        HIPCHK(ctx, hipGetLastError());
        return CG_OK;
    }
    // The scratch as three grids with their zero rows: the forcing, then the ping-pong pair. Every
    // call writes all the rows the kernels read (the zero rows included), since the array keeps
    // whatever the previous call, of any size, left in it.
    float *f = ctx->scratch;
    float *x[2] = {ctx->scratch + kScratchGrid, ctx->scratch + 2 * kScratchGrid};
    hipLaunchKernelGGL(k_stage_grid, dim3(blocks_for(total, 256)), dim3(256), 0, s, forcing, f, n, total);
    hipLaunchKernelGGL(k_stage_grid, dim3(blocks_for(total, 256)), dim3(256), 0, s, start, x[0], n, total);
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(st->across, 256)), dim3(256), 0, s, x[1] + n, (long)st->across,
                       0.0f);
    HIPCHK(ctx, hipGetLastError());
    int cur = 0;
    int rc = relax_pp(ctx, st, f, x, &cur, passes, false, false, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_copy_f32, dim3(blocks_for(n, 256)), dim3(256), 0, s, x[cur], out, n);
    HIPCHK(ctx, hipGetLastError());
    return CG_OK;
}

int cg_ocean_plan_create(cg_ctx *ctx, const cg_stencil *levels, int nlevels, int pole_is_wall, cg_plan **out) {
    return cg_ocean_plan_create_batch(ctx, levels, nlevels, pole_is_wall, 1, out);
}

int cg_ocean_plan_create_batch(cg_ctx *ctx, const cg_stencil *levels, int nlevels, int pole_is_wall, int instances,
                               cg_plan **out) {
    if (!ctx || !out) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    *out = nullptr;
    if (!levels || nlevels < 1 || nlevels > kMaxLevels) {
        set_err(ctx, "plan: 1..%d levels required", kMaxLevels);
        return CG_ERR_ARG;
    }
    if (instances < 1 || instances > kMaxInstances) {
        set_err(ctx, "plan: 1..%d instances", kMaxInstances);
        return CG_ERR_ARG;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    cg_plan *p = new cg_plan();
    p->ctx = ctx;
    p->n = nlevels;
    p->pole_is_wall = pole_is_wall ? 1 : 0;
    const int K = instances;
    p->K = K;
    p->red_base = K == 1 ? 0 : 4;
    p->largest_k.assign(K, -1.0);
    p->ans_par.assign(K, 0);
    p->last_steps.assign(K, 0);
    int rc = 0;
    auto bail = [&](int code) {
        free_plan(p);
        return code;
    };
    for (int l = 0; l < nlevels; ++l) {
        Stencil *s = find_stencil(ctx, levels[l]);
        if (!s) {
            set_err(ctx, "plan: no stencil %llu at level %d", (unsigned long long)levels[l], l);
            return bail(CG_ERR_ARG);
        }
        if (!s->has_centre) {
            set_err(ctx, "plan: level %d was uploaded without centre weights", l);
            return bail(CG_ERR_ARG);
        }
        if (s->across % 2 != 0) {
            set_err(ctx, "plan: level %d has an odd width; red-black colouring needs an even one", l);
            return bail(CG_ERR_ARG);
        }
        if (l > 0) {
            const Stencil *f = p->lv[l - 1].st;
            if (f->across % s->across || f->down % s->down) {
                set_err(ctx, "plan: level %d does not divide level %d", l, l - 1);
                return bail(CG_ERR_ARG);
            }
        }
        p->lv[l].st = s;
        ++s->plan_refs;
        // K copies of the grid, each with its zero row, all zero.
        if ((rc = dalloc_grid(ctx, &p->lv[l].f, s->across, (s->down + 1) * K - 1))) return bail(rc);
        if ((rc = dalloc_grid(ctx, &p->lv[l].x[0], s->across, (s->down + 1) * K - 1))) return bail(rc);
        if ((rc = dalloc_grid(ctx, &p->lv[l].x[1], s->across, (s->down + 1) * K - 1))) return bail(rc);
        if (l > 0 && (rc = dalloc(ctx, &p->lv[l].bal, s->cells))) return bail(rc);
    }
    const long n0 = p->lv[0].st->cells;
    if ((rc = dalloc_grid(ctx, &p->f0, p->lv[0].st->across, (p->lv[0].st->down + 1) * K - 1))) return bail(rc);
    if ((rc = dalloc_grid(ctx, &p->f_alt, p->lv[0].st->across, (p->lv[0].st->down + 1) * K - 1))) return bail(rc);
    if ((rc = dalloc(ctx, &p->r64, n0))) return bail(rc);
    if ((rc = dalloc(ctx, &p->x64[0], n0 * K))) return bail(rc);
    if ((rc = dalloc(ctx, &p->x64[1], n0 * K))) return bail(rc);
    p->first_base = std::max(p->red_base + 2 * K, 4);
    const int alloc_groups = p->first_base + K;
    if ((rc = dalloc(ctx, &p->d_u64, (size_t)alloc_groups * kMaxSlots))) return bail(rc);
    HIPCHK(ctx, hipMemset(p->d_u64, 0, (size_t)alloc_groups * kMaxSlots * sizeof(unsigned long long)));
    if (hipHostMalloc((void **)&p->h_red, (size_t)alloc_groups * kMaxSlots * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
        p->h_red = nullptr;
        set_err(ctx, "pinned host memory for the plan's maxima was refused");
        return bail(CG_ERR_NOMEM);
    }
    if ((rc = dalloc(ctx, &p->d_active, K))) return bail(rc);
    {
        const size_t gate_bytes = (sizeof(StepGate) * 2 * K + 511) & ~(size_t)511;
        const size_t res_bytes = sizeof(unsigned long long) * 2 * K * kMaxSlots;
        const size_t sums_bytes = sizeof(double) * (2 * K + 2) * 2 * kMaxSlots;
        p->check_bytes = gate_bytes + res_bytes + sizeof(double) * 4 * K * kMaxSlots;
        if ((rc = dalloc(ctx, &p->d_check, gate_bytes + res_bytes + sums_bytes))) return bail(rc);
        HIPCHK(ctx, hipMemset(p->d_check, 0, gate_bytes + res_bytes + sums_bytes));
        if (hipHostMalloc((void **)&p->h_check, p->check_bytes, hipHostMallocDefault) != hipSuccess) {
            p->h_check = nullptr;
            set_err(ctx, "pinned host memory for the plan's checks was refused");
            return bail(CG_ERR_NOMEM);
        }
        p->d_gate = reinterpret_cast<StepGate *>(p->d_check);
        p->d_res = reinterpret_cast<unsigned long long *>(p->d_check + gate_bytes);
        p->d_sums = reinterpret_cast<double *>(p->d_check + gate_bytes + res_bytes);
        p->h_gate = reinterpret_cast<StepGate *>(p->h_check);
        p->h_res = reinterpret_cast<unsigned long long *>(p->h_check + gate_bytes);
        p->h_sums = reinterpret_cast<double *>(p->h_check + gate_bytes + res_bytes);
    }
    if ((rc = dalloc(ctx, &p->d_thr, 3 * K))) return bail(rc);
    if ((rc = dalloc(ctx, &p->d_want, K))) return bail(rc);
    HIPCHK(ctx, hipMemset(p->d_want, 0, sizeof(int) * K));
    HIPCHK(ctx, hipMemset(p->d_active, 0, sizeof(int) * K));
    if (hipHostMalloc((void **)&p->h_active, sizeof(int) * K, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&p->h_thr, sizeof(double) * 3 * K, hipHostMallocDefault) != hipSuccess) {
        set_err(ctx, "pinned host memory for the plan's flags was refused");
        return bail(CG_ERR_NOMEM);
    }
    if ((rc = dalloc(ctx, &p->d_passes, 4))) return bail(rc);
    HIPCHK(ctx, hipMemset(p->x64[0], 0, sizeof(double) * n0 * K));
    HIPCHK(ctx, hipMemset(p->x64[1], 0, sizeof(double) * n0 * K));
    if (nlevels > 1 && (rc = build_coarsest_inverse(ctx, p))) return bail(rc);
    if ((rc = plan_fusion_setup(ctx, p))) return bail(rc);
    *out = p;
    return CG_OK;
}

int cg_ocean_plan_destroy(cg_plan *p) {
    if (!p) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->group_refs > 0) {
        set_err(ctx, "the plan belongs to %d group(s): destroy them first", p->group_refs);
        return CG_ERR_STATE;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_plan(p);
    return CG_OK;
}

int cg_ocean_plan_set_forcing(cg_plan *p, const float *forcing) { return cg_ocean_plan_set_forcing_at(p, 0, forcing); }

int cg_ocean_plan_set_forcing_at(cg_plan *p, int instance, const float *forcing) {
    if (!p || !forcing || instance < 0 || instance >= p->K) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Stencil *st = p->lv[0].st;
    HIPCHK(ctx, hipMemcpyAsync(p->f0 + instance * vstride(st), forcing, sizeof(float) * st->cells, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (instance == 0) p->largest = -1.0;
    p->largest_k[instance] = -1.0;
    p->x64_valid = false;
    return CG_OK;
}

// Whether the plan can run CG_METHOD_LEARNED: a coarse network set, and the fused cycle.
static bool learned_ready(cg_plan *p) {
    if (p->net.level <= 0) {
        set_err(p->ctx, "CG_METHOD_LEARNED needs cg_ocean_plan_set_coarse_net first");
        return false;
    }
    if (!(p->ctx->fused && p->fusable)) {
        set_err(p->ctx, "CG_METHOD_LEARNED needs the fused cycle (row-constant weights, halving grids)");
        return false;
    }
    return true;
}

int cg_ocean_plan_solve_all(cg_plan *p, double tol, int method, cg_solution *infos) {
    if (!p || !infos) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (method != CG_METHOD_REFINED && method != CG_METHOD_LEARNED) {
        set_err(ctx, "a plan's instances solve together by CG_METHOD_REFINED or CG_METHOD_LEARNED only");
        return CG_ERR_ARG;
    }
    if (p->K > 1 && !(ctx->fused && p->fusable)) {
        set_err(ctx, "several instances need the fused cycle (row-constant weights, halving grids)");
        return CG_ERR_STATE;
    }
    if (method == CG_METHOD_LEARNED && !learned_ready(p)) return CG_ERR_STATE;
    return refined_solve(p, tol, false, infos, nullptr, method == CG_METHOD_LEARNED);
}

int cg_ocean_plan_read_at(cg_plan *p, int instance, double *out64, float *out32) {
    if (!p || instance < 0 || instance >= p->K) return CG_ERR_ARG;
    if (p->K == 1) return cg_ocean_plan_read(p, out64, out32);
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (p->last_answer != 1) {
        set_err(ctx, "the instances have no answer yet");
        return CG_ERR_STATE;
    }
    const long n = p->lv[0].st->cells;
    const double *x = p->x64[p->ans_par[instance]] + instance * n;
    if (out64) HIPCHK(ctx, hipMemcpyAsync(out64, x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (out32) {
        float *tmp = p->lv[0].x[1 - p->lv[0].cur];
        hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, x, tmp, n);
        HIPCHK(ctx, hipMemcpyAsync(out32, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CG_OK;
}

static int plan_solve_locked(cg_plan *p, double tol, int method, int warm, cg_solution *info,
                             const volatile int *cancel) {
    cg_ctx *ctx = p->ctx;
    cg_solution local{};
    if (!info) info = &local;
    *info = cg_solution{};
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (method == CG_METHOD_REFERENCE || method == CG_METHOD_RESCUE || method == CG_METHOD_KRYLOV) {
        PlanLevel &l0 = p->lv[0];
        if (!warm) HIPCHK(ctx, hipMemsetAsync(l0.x[l0.cur], 0, sizeof(float) * l0.st->cells, ctx->stream));
        if (method == CG_METHOD_RESCUE) return rescue_solve(p, tol, info, cancel);
        if (method == CG_METHOD_KRYLOV) return krylov_restarted(p, tol, info, cancel);
        return ref_solve(p, tol, info, cancel);
    }
    if (method == CG_METHOD_REFINED) return refined_solve(p, tol, warm != 0, info, cancel);
    if (method == CG_METHOD_FGMRES) return fgmres_solve(p, tol, warm != 0, info, cancel);
    if (method == CG_METHOD_LEARNED) {
        if (!learned_ready(p)) return CG_ERR_STATE;
        return refined_solve(p, tol, warm != 0, info, cancel, true);
    }
    if (method == CG_METHOD_LEARNED_FGMRES) {
        if (!p->pre_fn) {
            set_err(ctx, "CG_METHOD_LEARNED_FGMRES needs cg_ocean_plan_set_preconditioner first");
            return CG_ERR_STATE;
        }
        return fgmres_solve(p, tol, warm != 0, info, cancel, true);
    }
    set_err(ctx, "unknown or unbuilt method %d", method);
    return CG_ERR_ARG;
}

int cg_unet_weight_count(int base) { return base >= 16 && base % 16 == 0 ? (int)net_weight_count(base) : -1; }

int cg_ocean_plan_set_coarse_net(cg_plan *p, int level, int base, const uint16_t *weights, size_t count,
                                 const float *depth, double out_scale) {
    if (!p) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(p);
    coarse_net_free(p->net);
    if (!weights) return CG_OK;  // removed
    if (!p->fusable || level < 1 || level > p->bottom || level >= p->n) {
        set_err(ctx, "the coarse network serves a grid 1..%d of the fused cycle (asked %d)", p->bottom, level);
        return CG_ERR_ARG;
    }
    const Stencil *st = p->lv[level].st;
    if (!coarse_net_shape_ok(st->down, st->across)) {
        set_err(ctx, "grid %d is %d x %d: the network needs rows %% 8 == 0 and columns %% 128 == 0", level, st->down,
                st->across);
        return CG_ERR_ARG;
    }
    if (!(out_scale > 0.0) || !std::isfinite(out_scale)) {
        set_err(ctx, "the network's output scale must be positive and finite");
        return CG_ERR_ARG;
    }
    const int rc = coarse_net_setup(p->net, level, base, p->K, st->down, st->across, vstride(st), weights, count, depth,
                                    st->water, (float)out_scale);
    if (rc) {
        set_err(ctx, "coarse network setup failed (%d): base %d needs %zu weights, %zu given", rc, base,
                net_weight_count(base), count);
        return rc;
    }
    return CG_OK;
}

int cg_ocean_plan_apply_coarse_net(cg_plan *p, const float *forcing, float *out) {
    if (!p || !forcing || !out) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->net.level <= 0) {
        set_err(ctx, "the plan has no coarse network");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    PlanLevel &N = p->lv[p->net.level];
    const long cells = N.st->cells;
    // Every instance gets the given forcing (the network runs them all); instance 0's answer returns.
    for (int i = 0; i < p->K; ++i)
        HIPCHK(ctx, hipMemcpyAsync(N.f + i * vstride(N.st), forcing, sizeof(float) * cells, hipMemcpyHostToDevice,
                                   ctx->stream));
    const int rc = coarse_net_forward(p->net, N.f, N.x[N.cur], ctx->stream);
    if (rc) {
        set_err(ctx, "the coarse network's launches failed (%d)", rc);
        return rc;
    }
    HIPCHK(ctx, hipMemcpyAsync(out, N.x[N.cur], sizeof(float) * cells, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CG_OK;
}

int cg_ocean_plan_capture_coarse(cg_plan *p, int level, double tol, int most_steps, float *out) {
    if (!p || !out || most_steps < 1) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (p->K != 1 || !(ctx->fused && p->fusable) || level < 1 || level > p->bottom) {
        set_err(ctx, "capture needs a plan of one instance on the fused cycle and a grid 1..%d", p->bottom);
        return CG_ERR_ARG;
    }
    double largest;
    int rc = plan_largest(p, &largest);
    if (rc) return rc;
    const Stencil *st0 = p->lv[0].st;
    const long cells = p->lv[level].st->cells;
    // The plain refinement from zero, step by step, the grid's forcing copied out after each step.
    HIPCHK(ctx, hipMemsetAsync(p->x64[p->x64cur], 0, sizeof(double) * st0->cells, ctx->stream));
    hipLaunchKernelGGL(k_fill_u64, dim3(2), dim3(kMaxSlots), 0, ctx->stream, p->d_res, 2 * kMaxSlots, 0ull);
    if ((rc = outer_step(p, nullptr, outer_fused(p) ? level0_forcing(p, 1 - p->x64cur) : nullptr))) return rc;
    int steps = 0;
    while (steps < most_steps) {
        const int slot = p->x64cur;
        if ((rc = refine_step(p, false))) return rc;
        const PlanLevel &N = p->lv[level];
        HIPCHK(ctx, hipMemcpyAsync(out + (size_t)steps * cells, N.f, sizeof(float) * cells, hipMemcpyDeviceToHost,
                                   ctx->stream));
        ++steps;
        unsigned long long bits;
        if ((rc = read_res(p, slot, &bits))) return rc;
        const double r = largest > 0.0 ? bits_to_double(bits) / largest : 0.0;
        if (!(r >= tol)) break;
    }
    p->x64_valid = false;
    return steps;
}

int cg_ocean_plan_net_fallbacks(cg_plan *p) { return p ? p->net_fallbacks : CG_ERR_ARG; }

int cg_ocean_plan_set_preconditioner(cg_plan *p, cg_precond_fn fn, void *user) {
    if (!p) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(p->ctx->mu);
    if (p->K > 1) {
        set_err(p->ctx, "a learned preconditioner serves a plan of one instance");
        return CG_ERR_STATE;
    }
    p->pre_fn = fn;
    p->pre_user = fn ? user : nullptr;
    return CG_OK;
}

int cg_ocean_plan_solve(cg_plan *p, double tol, int method, int warm, cg_solution *info) {
    if (!p) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(p->ctx->mu);
    if (p->K > 1) {
        set_err(p->ctx, "a plan of several instances solves with cg_ocean_plan_solve_all");
        return CG_ERR_STATE;
    }
    return plan_solve_locked(p, tol, method, warm, info, nullptr);
}

int cg_ocean_plan_solve_batch(cg_plan *const *plans, int count, double tol, int method, cg_solution *infos) {
    if (!plans || count < 0) return CG_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        int rc = cg_ocean_plan_solve(plans[i], tol, method, 0, infos ? &infos[i] : nullptr);
        if (rc) return rc;
    }
    return CG_OK;
}

int cg_ocean_group_create(cg_plan *const *plans, int count, cg_group **out) {
    if (!out) return CG_ERR_ARG;
    *out = nullptr;
    if (!plans || count < 1 || count > kMaxGroupPlans) return CG_ERR_ARG;
    cg_ctx *ctx = plans[0] ? plans[0]->ctx : nullptr;
    if (!ctx) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    for (int i = 0; i < count; ++i) {
        if (!plans[i] || plans[i]->ctx != ctx) {
            set_err(ctx, "every plan of a group must belong to one context");
            return CG_ERR_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i]) {
                set_err(ctx, "a plan appears twice in the group");
                return CG_ERR_ARG;
            }
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    cg_group *gr = new cg_group();
    gr->ctx = ctx;
    gr->plans.assign(plans, plans + count);
    bool ok = hipEventCreateWithFlags(&gr->start, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < count; ++i) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
        if (ok) gr->streams.push_back(s);
        ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (ok) gr->done.push_back(e);
    }
    if (!ok) {
        set_err(ctx, "the group's streams or events could not be created");
        for (hipStream_t s : gr->streams) (void)hipStreamDestroy(s);
        for (hipEvent_t e : gr->done) (void)hipEventDestroy(e);
        if (gr->start) (void)hipEventDestroy(gr->start);
        delete gr;
        return CG_ERR_DEVICE;
    }
    for (cg_plan *p : gr->plans) p->group_refs++;
    *out = gr;
    return CG_OK;
}

int cg_ocean_group_destroy(cg_group *gr) {
    if (!gr) return CG_ERR_ARG;
    cg_ctx *ctx = gr->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t s : gr->streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (hipEvent_t e : gr->done) (void)hipEventDestroy(e);
    (void)hipEventDestroy(gr->start);
    drop_merged(gr);
    for (cg_plan *p : gr->plans) p->group_refs--;
    delete gr;
    return CG_OK;
}

int cg_ocean_group_solve(cg_group *gr, double tol, int method, cg_solution *infos) {
    if (!gr || !infos || !(tol > 0.0)) return CG_ERR_ARG;
    cg_ctx *ctx = gr->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (method != CG_METHOD_REFINED && method != CG_METHOD_LEARNED) {
        set_err(ctx, "a group solves by CG_METHOD_REFINED or CG_METHOD_LEARNED");
        return CG_ERR_ARG;
    }
    const bool learned = method == CG_METHOD_LEARNED;
    for (cg_plan *p : gr->plans) {
        if (!(ctx->fused && p->fusable)) {
            set_err(ctx, "a group's plans need the fused cycle (row-constant weights, halving grids)");
            return CG_ERR_STATE;
        }
        if (learned && !learned_ready(p)) return CG_ERR_STATE;
    }
    const bool graph = ctx->use_graph && !ctx->time_kernels;
    struct Ungate {
        std::vector<cg_plan *> &ps;
        ~Ungate() {
            for (cg_plan *p : ps) {
                p->gated = false;
                p->net_on = false;
            }
        }
    } ungate{gr->plans};
    // Every plan starts from zero.
    std::vector<RefineRun> runs(gr->plans.size());
    int rc;
    for (size_t k = 0; k < gr->plans.size(); ++k) {
        cg_plan *p = gr->plans[k];
        p->x64cur = 0;
        if ((rc = run_begin(runs[k], p, tol, false, learned, nullptr, graph, learned ? 0.0 : 1.0))) return rc;
    }
    // The merged step: one graph a step for every plan, on the context's stream.
    const bool merged = graph && !learned && group_mergeable(gr);
    if (merged) {
        const std::vector<long> key = merged_key(gr);
        if (key != gr->merged_key) {
            drop_merged(gr);
            gr->merged_key = key;
        }
    }
    // Otherwise the plans' streams start after the first outer steps on the context's stream.
    if (graph && !merged) {
        HIPCHK(ctx, hipEventRecord(gr->start, ctx->stream));
        for (hipStream_t s : gr->streams) HIPCHK(ctx, hipStreamWaitEvent(s, gr->start, 0));
    }
    auto step = [&](const std::vector<char> &on) -> int {
        if (merged) {
            // The plans still running step together; they share a binary64 parity (every plan starts
            // at the same, and one that drops out never returns within the solve).
            unsigned mask = 0;
            int q = -1;
            for (size_t k = 0; k < gr->plans.size(); ++k) {
                if (!on[k]) continue;
                mask |= 1u << k;
                if (q >= 0 && gr->plans[k]->x64cur != q) {
                    set_err(ctx, "the group's running plans left their common binary64 parity");
                    return CG_ERR_STATE;
                }
                q = gr->plans[k]->x64cur;
            }
            if (!mask) return CG_OK;
            cg_group::Merged *ms = nullptr;
            int r = merged_for(gr, mask, &ms);
            if (r) return r;
            HIPCHK(ctx, hipGraphLaunch(ms->exec[q], ctx->stream));
            for (size_t k = 0; k < gr->plans.size(); ++k)
                if (on[k]) gr->plans[k]->x64cur = 1 - q;
            return CG_OK;
        }
        for (size_t k = 0; k < gr->plans.size(); ++k) {
            if (!on[k]) continue;
            cg_plan *p = gr->plans[k];
            if (graph) {
                hipGraphExec_t *const exec = p->net_on ? p->net_exec : p->step_exec;
                HIPCHK(ctx, hipGraphLaunch(exec[p->x64cur], gr->streams[k]));
                p->x64cur = 1 - p->x64cur;
            } else {
                int r = refine_step(p, true);
                if (r) return r;
            }
        }
        return CG_OK;
    };
    auto join = [&]() -> int {
        if (!graph || merged) return CG_OK;
        for (size_t k = 0; k < gr->plans.size(); ++k) {
            HIPCHK(ctx, hipEventRecord(gr->done[k], gr->streams[k]));
            HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, gr->done[k], 0));
        }
        return CG_OK;
    };
    int status = CG_OK;
    if ((rc = run_steps(runs, nullptr, step, join, &status))) return rc;
    size_t at = 0;
    for (RefineRun &r : runs) {
        run_finish(r, r.launched, status, infos + at);
        r.p->gated = false;
        r.p->net_on = false;
        if (learned) {
            if ((rc = learned_fallback(r.p, tol, infos + at, nullptr))) return rc;
        } else {
            for (int i = 0; i < r.K; ++i) infos[at + i].refinements = infos[at + i].cycles + 1;
        }
        at += (size_t)r.K;
    }
    return CG_OK;
}

int cg_ocean_plan_set_bodies(cg_plan *p, const int32_t *body_of, int count) {
    if (!p || count < 0 || (count > 0 && !body_of)) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->K > 1) {
        set_err(ctx, "this call takes a plan of one instance");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Stencil *st = p->lv[0].st;
    const long n = st->cells;
    for (void *v : {(void *)p->d_body_of, (void *)p->d_order, (void *)p->d_seg_start, (void *)p->d_seg_len,
                    (void *)p->d_body_seg0, (void *)p->d_body_nseg, (void *)p->d_margin, (void *)p->d_seg_sum,
                    (void *)p->d_shift})
        dfree(v);
    p->d_body_of = p->d_order = p->d_seg_start = p->d_seg_len = p->d_body_seg0 = p->d_body_nseg = nullptr;
    p->d_margin = p->d_seg_sum = nullptr;
    p->d_shift = nullptr;
    p->nbodies = p->nseg = 0;
    if (count == 0) return CG_OK;
    if (!st->has_centre) {
        set_err(ctx, "bodies need the finest stencil's centre weights");
        return CG_ERR_STATE;
    }
    std::vector<int> per(count, 0);
    for (long c = 0; c < n; ++c) {
        const int b = body_of[c];
        if (b < -1 || b >= count) {
            set_err(ctx, "body %d of cell %ld is outside -1..%d", b, c, count - 1);
            return CG_ERR_ARG;
        }
        if (b >= 0) per[b]++;
    }
    // Cells grouped by body, each body's in cell order; the margins summed in cell order.
    std::vector<int> first(count + 1, 0), order, fill;
    for (int b = 0; b < count; ++b) first[b + 1] = first[b] + per[b];
    order.resize(first[count]);
    fill.assign(first.begin(), first.end() - 1);
    std::vector<double> margin(count, 0.0);
    for (long c = 0; c < n; ++c) {
        const int b = body_of[c];
        if (b < 0) continue;
        order[fill[b]++] = (int)c;
        const int row = (int)(c / st->across);
        const float4 w = st->rowconst ? st->h_w[row] : st->h_w[c];
        const double cc = st->rowconst ? st->h_c[row] : st->h_c[c];
        const double weights = (double)w.x + (double)w.y + (double)w.z + (double)w.w;
        margin[b] += cc * (1.0 - weights);
    }
    std::vector<int> seg_start, seg_len, body_seg0(count), body_nseg(count);
    for (int b = 0; b < count; ++b) {
        body_seg0[b] = (int)seg_start.size();
        for (int at = first[b]; at < first[b + 1]; at += kBodySegment) {
            seg_start.push_back(at);
            seg_len.push_back(std::min(kBodySegment, first[b + 1] - at));
        }
        body_nseg[b] = (int)seg_start.size() - body_seg0[b];
    }
    const int nseg = (int)seg_start.size();
    int rc;
    if ((rc = dalloc(ctx, &p->d_body_of, n)) || (rc = dalloc(ctx, &p->d_order, order.size())) ||
        (rc = dalloc(ctx, &p->d_seg_start, nseg)) || (rc = dalloc(ctx, &p->d_seg_len, nseg)) ||
        (rc = dalloc(ctx, &p->d_body_seg0, count)) || (rc = dalloc(ctx, &p->d_body_nseg, count)) ||
        (rc = dalloc(ctx, &p->d_margin, count)) || (rc = dalloc(ctx, &p->d_seg_sum, nseg)) ||
        (rc = dalloc(ctx, &p->d_shift, count)))
        return rc;
    HIPCHK(ctx, hipMemcpy(p->d_body_of, body_of, sizeof(int) * n, hipMemcpyHostToDevice));
    if (!order.empty()) HIPCHK(ctx, hipMemcpy(p->d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    if (nseg) {
        HIPCHK(ctx, hipMemcpy(p->d_seg_start, seg_start.data(), sizeof(int) * nseg, hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(p->d_seg_len, seg_len.data(), sizeof(int) * nseg, hipMemcpyHostToDevice));
    }
    HIPCHK(ctx, hipMemcpy(p->d_body_seg0, body_seg0.data(), sizeof(int) * count, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(p->d_body_nseg, body_nseg.data(), sizeof(int) * count, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(p->d_margin, margin.data(), sizeof(double) * count, hipMemcpyHostToDevice));
    p->nbodies = count;
    p->nseg = nseg;
    return CG_OK;
}

int cg_ocean_plan_correct_bodies(cg_plan *p, const float *values, float *out) {
    if (!p || !values || !out) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->K > 1) {
        set_err(ctx, "this call takes a plan of one instance");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    PlanLevel &l0 = p->lv[0];
    const long n = l0.st->cells;
    float *x = l0.x[l0.cur];
    HIPCHK(ctx, hipMemcpyAsync(x, values, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    int rc = correct_bodies(p, p->f0, x);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(out, x, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    p->last_answer = 0;
    return CG_OK;
}

int cg_ocean_plan_set_start(cg_plan *p, const float *start) {
    if (!p || !start) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->K > 1) {
        set_err(ctx, "this call takes a plan of one instance");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    PlanLevel &l0 = p->lv[0];
    const long n = l0.st->cells;
    HIPCHK(ctx, hipMemcpyAsync(l0.x[l0.cur], start, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, l0.x[l0.cur],
                       p->x64[p->x64cur], n);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    p->x64_valid = true;
    p->last_answer = 0;
    return CG_OK;
}

// ---- the round's input and output steps ---------------------------------------------------------

int cg_ocean_water_on(cg_ctx *ctx, const uint8_t *map_land, int map_across, int map_down, int across, int down,
                      uint8_t *water) {
    if (!ctx || !map_land || !water || map_across <= 0 || map_down <= 0 || across <= 0 || down <= 0)
        return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    uint8_t *dland = nullptr, *dwater = nullptr;
    const size_t nmap = (size_t)map_across * map_down, n = (size_t)across * down;
    int rc = dalloc(ctx, &dland, nmap);
    if (!rc) rc = dalloc(ctx, &dwater, n);
    if (!rc && hipMemcpyAsync(dland, map_land, nmap, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = CG_ERR_DEVICE;
    if (!rc) {
        hipLaunchKernelGGL(k_water_on, dim3(blocks_for(across, 256), down), dim3(256), 0, ctx->stream, dland, map_across,
                           map_down, across, down, dwater);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(water, dwater, n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_err(ctx, "water_on on the device failed");
            rc = CG_ERR_DEVICE;
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    dfree(dland);
    dfree(dwater);
    return rc;
}

int cg_ocean_curl_forcing(cg_ctx *ctx, int across, int down, double width_m, double height_m, const double *eastward,
                          const double *northward, const double *depth, double *forcing) {
    if (!ctx || !eastward || !northward || !forcing || across <= 0 || down <= 0 || !(width_m > 0.0) ||
        !(height_m > 0.0))
        return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)across * down;
    double *de = nullptr, *dn = nullptr, *dd = nullptr, *dF = nullptr;
    int rc = dalloc(ctx, &de, n);
    if (!rc) rc = dalloc(ctx, &dn, n);
    if (!rc && depth) rc = dalloc(ctx, &dd, n);
    if (!rc) rc = dalloc(ctx, &dF, n);
    if (!rc) {
        bool ok = hipMemcpyAsync(de, eastward, n * 8, hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
                  hipMemcpyAsync(dn, northward, n * 8, hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
                  (!depth || hipMemcpyAsync(dd, depth, n * 8, hipMemcpyHostToDevice, ctx->stream) == hipSuccess);
        if (ok) {
            hipLaunchKernelGGL(k_curl_forcing, dim3(blocks_for(across, 256), down), dim3(256), 0, ctx->stream, across,
                               down, width_m, height_m, de, dn, dd, dF, LevelArgs{}, (float *)nullptr);
            ok = hipGetLastError() == hipSuccess &&
                 hipMemcpyAsync(forcing, dF, n * 8, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
                 hipStreamSynchronize(ctx->stream) == hipSuccess;
        }
        if (!ok) {
            set_err(ctx, "curl_forcing on the device failed");
            rc = CG_ERR_DEVICE;
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    for (double *v : {de, dn, dd, dF}) dfree(v);
    return rc;
}

int cg_ocean_plan_set_stress(cg_plan *p, double width_m, double height_m, const double *eastward,
                             const double *northward, const double *depth) {
    if (!p || !eastward || !northward || !(width_m > 0.0) || !(height_m > 0.0)) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->K > 1) {
        set_err(ctx, "this call takes a plan of one instance");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Stencil *st = p->lv[0].st;
    if (!st->has_centre) {
        set_err(ctx, "the finest stencil was uploaded without centre weights");
        return CG_ERR_STATE;
    }
    const size_t n = st->cells;
    int rc;
    if (!p->tau_e && ((rc = dalloc(ctx, &p->tau_e, n)) || (rc = dalloc(ctx, &p->tau_n, n)))) return rc;
    if (depth && !p->depth && (rc = dalloc(ctx, &p->depth, n))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(p->tau_e, eastward, n * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(p->tau_n, northward, n * 8, hipMemcpyHostToDevice, ctx->stream));
    if (depth) HIPCHK(ctx, hipMemcpyAsync(p->depth, depth, n * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_curl_forcing, dim3(blocks_for(st->across, 256), st->down), dim3(256), 0, ctx->stream,
                       st->across, st->down, width_m, height_m, p->tau_e, p->tau_n, depth ? p->depth : nullptr,
                       (double *)nullptr, st->args(), p->f0);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    p->largest = -1.0;
    p->x64_valid = false;
    return CG_OK;
}

int cg_ocean_plan_velocities(cg_plan *p, double width_m, double height_m, float *eastward, float *northward) {
    if (!p || !(width_m > 0.0) || !(height_m > 0.0)) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (p->K > 1) {
        set_err(ctx, "this call takes a plan of one instance");
        return CG_ERR_STATE;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Stencil *st = p->lv[0].st;
    const size_t n = st->cells;
    int rc;
    if (!p->vel && (rc = dalloc(ctx, &p->vel, 2 * n))) return rc;
    const float *s32 = p->last_answer == 0 ? p->lv[0].x[p->lv[0].cur] : nullptr;
    const double *s64 = p->last_answer == 0 ? nullptr : p->x64[p->x64cur];
    hipLaunchKernelGGL(k_velocities, dim3(blocks_for(st->across, 256), st->down), dim3(256), 0, ctx->stream,
                       st->across, st->down, st->water, s32, s64, width_m, height_m, p->vel, p->vel + n);
    HIPCHK(ctx, hipGetLastError());
    if (eastward) HIPCHK(ctx, hipMemcpyAsync(eastward, p->vel, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (northward) HIPCHK(ctx, hipMemcpyAsync(northward, p->vel + n, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CG_OK;
}

int cg_ocean_plan_read(cg_plan *p, double *out64, float *out32) {
    if (!p) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Stencil *st = p->lv[0].st;
    const long n = st->cells;
    if (p->last_answer == 0) {
        const float *x = p->lv[0].x[p->lv[0].cur];
        if (out32) HIPCHK(ctx, hipMemcpyAsync(out32, x, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
        if (out64) {
            hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, x, p->r64, n);
            HIPCHK(ctx, hipMemcpyAsync(out64, p->r64, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
        }
    } else {
        const double *x = p->x64[p->x64cur];
        if (out64) HIPCHK(ctx, hipMemcpyAsync(out64, x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
        if (out32) {
            float *tmp = p->lv[0].x[1 - p->lv[0].cur];
            hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, x, tmp, n);
            HIPCHK(ctx, hipMemcpyAsync(out32, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CG_OK;
}

int cg_ocean_solve(cg_ctx *ctx, const cg_stencil *levels, int nlevels, const float *forcing, const float *start,
                   double tol, int pole_is_wall, int method, float *out, double *out64, cg_solution *info,
                   const volatile int *cancel) {
    if (!ctx || !forcing) return CG_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    cg_plan *p = nullptr;
    int rc = cg_ocean_plan_create(ctx, levels, nlevels, pole_is_wall, &p);
    if (rc) return rc;
    rc = cg_ocean_plan_set_forcing(p, forcing);
    const long n = p->lv[0].st->cells;
    int warm = 0;
    if (!rc && start) {
        warm = 1;
        if (method == CG_METHOD_REFERENCE || method == CG_METHOD_RESCUE) {
            PlanLevel &l0 = p->lv[0];
            hipError_t e = hipMemcpy(l0.x[l0.cur], start, sizeof(float) * n, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = CG_ERR_DEVICE;
        } else {
            std::vector<double> s64(start, start + n);
            hipError_t e = hipMemcpy(p->x64[p->x64cur], s64.data(), sizeof(double) * n, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = CG_ERR_DEVICE;
            p->x64_valid = true;
        }
    }
    if (!rc) rc = plan_solve_locked(p, tol, method, warm, info, cancel);
    if (!rc) rc = cg_ocean_plan_read(p, out64, out);
    cg_ocean_plan_destroy(p);
    return rc;
}

double cg_ocean_plan_time_kernel(cg_plan *p, int kind, int reps) {
    if (!p || reps < 1 || kind < 0 || (kind > 6 && kind < 16) || kind >= 16 + 2 * kMaxLevels) return -1.0;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return -1.0;
    if (kind >= 16) {
        // Grid l's down leg (even kinds) or up leg (odd) as the refined cycle launches it.
        const int l = (kind - 16) / 2;
        const bool up = (kind - 16) % 2 == 1;
        if (!(ctx->fused && p->fusable) || l >= p->bottom) {
            set_err(ctx, "the plan has no fused leg on grid %d", l);
            return -1.0;
        }
        auto leg = [&]() -> hipError_t {
            return up ? up_leg(p, l, nullptr, nullptr, false, plan_tile_rows(p, l))
                      : down_leg(p, l, nullptr, plan_tile_rows(p, l));
        };
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0;
        if (leg() != hipSuccess) return -1.0;  // warm
        (void)hipEventRecord(a, ctx->stream);
        for (int i = 0; i < reps; ++i)
            if (leg() != hipSuccess) return -1.0;
        (void)hipEventRecord(b, ctx->stream);
        float ms = 0.0f;
        const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ok ? 1000.0 * ms / reps : -1.0;
    }
    const bool with_sums = kind == 5;
    // Kind 6: kind 3 over every instance of the plan, as a batched step launches it.
    const int instances = kind == 6 ? p->K : 1;
    if (with_sums || kind == 6) kind = 3;
    if (kind == 3 && !outer_fused(p)) {
        set_err(ctx, "the plan's steps do not fuse the outer step");
        return -1.0;
    }
    if (kind != 2 && (!p->fusable || (kind != 4 && p->bottom == 0))) {
        set_err(ctx, "plan has no fused finest tile level to time");
        return -1.0;
    }
    PlanLevel &L = p->lv[0];
    const int nh = level_half_sweeps(ctx, 0), th = plan_tile_rows(p, 0);
    // The fused kinds' binary64 output goes to scratch: a batched plan's instances keep their answers
    // in the parity where each stopped, so neither parity is free of answers.
    double *x64_out = p->r64;
    struct Tmp {
        double *q = nullptr;
        ~Tmp() { dfree(q); }
    } tmp;
    if (kind == 3 && instances > 1) {
        int rc = dalloc(ctx, &tmp.q, (size_t)L.st->cells * instances);
        if (rc) return -1.0;
        x64_out = tmp.q;
    }
    auto launch = [&]() -> hipError_t {
        if (kind == 4) {
            bottom_launch(p, nullptr, nullptr);
            return hipGetLastError();
        }
        if (kind == 2) {
            MaxReduce red;
            red.out = p->d_u64 + 3 * kMaxSlots;
            red.clear = nullptr;
            const Stencil *st = L.st;
            const dim3 grid(blocks_for(st->across / 2, 256), blocks_for(st->down, kRowsPerBlock));
            hipLaunchKernelGGL(k_outer_residual, grid, dim3(256), 0, ctx->stream, st->args(), p->f0, p->x64[p->x64cur],
                               L.x[L.cur], p->x64[1 - p->x64cur], L.f, red, OuterBatch{});
            return hipGetLastError();
        }
        TileExtra ex{};
        PlanLevel &C = p->lv[1];
        if (kind == 0) {
            ex.cwater = C.st->water;
            ex.g = L.g;
            ex.fc = C.f;
            return tile_dispatch(ctx, L.st->args(), L.f, L.x[L.cur], L.x[1 - L.cur], nh, true, TILE_DOWN, ex, th);
        }
        ex.xc = C.x[C.cur];
        ex.ca = C.st->across;
        ex.cd = C.st->down;
        ex.beyond = p->pole_is_wall ? -1.0f : 1.0f;
        if (kind == 3) {
            // The fused finest up leg and outer step: reads parity q, writes the binary64 answer to
            // scratch, its maxima into the scratch slot groups 3 (out) and 2 (cleared).
            const int q = p->x64cur;
            ex.x64 = p->x64[q];
            ex.x64n = x64_out;
            ex.f0 = p->f0;
            ex.fcn = level0_forcing(p, 1 - q);
            ex.red.out = p->d_u64 + 3 * kMaxSlots;
            ex.red.clear = p->d_u64 + 2 * kMaxSlots;
            ex.s64 = L.st->cells;
            ex.red_stride = 2 * kMaxSlots;
            if (with_sums) {  // the error estimate's sums, which only a solve's last steps take
                ex.sums.out = sums_group(p, 2 * p->K);
                ex.sums.clear = sums_group(p, 2 * p->K + 1);
            }
            if (instances > 1) {
                // Every instance's maxima into its own pair of the plan's residual slot groups
                // (parity q out, 1 - q cleared: 2 K kMaxSlots in all); no active flags, so every
                // instance runs.
                ex.fs = vstride(L.st);
                ex.cs = vstride(C.st);
                ex.red.out = p->d_res + q * kMaxSlots;
                ex.red.clear = p->d_res + (1 - q) * kMaxSlots;
            }
            return tile_dispatch(ctx, L.st->args(), level0_forcing(p, q), L.x[L.cur], L.x[1 - L.cur], nh, false,
                                 TILE_UP_OUTER, ex, th, instances);
        }
        return tile_dispatch(ctx, L.st->args(), L.f, L.x[L.cur], L.x[1 - L.cur], nh, false, TILE_UP, ex, th);
    };
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0;
    if (launch() != hipSuccess) return -1.0;  // warm
    (void)hipEventRecord(a, ctx->stream);
    for (int i = 0; i < reps; ++i)
        if (launch() != hipSuccess) return -1.0;
    (void)hipEventRecord(b, ctx->stream);
    float ms = 0.0f;
    const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ok ? 1000.0 * ms / reps : -1.0;
}

int cg_ocean_plan_set_tile_rows(cg_plan *p, const int *rows) {
    if (!p || !rows) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    for (int l = 0; l < p->n; ++l)
        if (rows[l] != 0 && rows[l] != 2 && rows[l] != 4 && rows[l] != 6 && rows[l] != 8 && rows[l] != 16 &&
            rows[l] != 24 && rows[l] != 32) {
            set_err(ctx, "tile rows of grid %d must be 0, 2, 4, 6, 8, 16, 24 or 32", l);
            return CG_ERR_ARG;
        }
    for (int l = 0; l < p->n; ++l) p->tuned_rows[l] = rows[l];
    drop_graphs(p);
    return CG_OK;
}

int cg_ocean_plan_autotune(cg_plan *p, int reps, int *rows_out) {
    if (!p || reps < 1) return CG_ERR_ARG;
    cg_ctx *ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (!(ctx->fused && p->fusable)) {
        set_err(ctx, "only the fused cycle's tiles are tuned");
        return CG_ERR_STATE;
    }
    static const int kCandidates[] = {2, 4, 6, 8, 16, 24, 32};
    const bool graph = ctx->use_graph && !ctx->time_kernels;

    // A refinement in progress from zero, as a solve starts one, so the steps timed are the solve's.
    const int K = p->K;
    const Stencil *st = p->lv[0].st;
    HIPCHK(ctx, hipMemsetAsync(p->x64[p->x64cur], 0, sizeof(double) * st->cells * K, ctx->stream));
    // The step gate with thresholds of zero never stops an instance, so every timed step runs.
    int rc = gate_reset(p, 1 - p->x64cur, std::vector<double>(K, 0.0), ctx->error_tol, 0);
    if (rc) return rc;
    rc = outer_step(p, nullptr, outer_fused(p) ? level0_forcing(p, 1 - p->x64cur) : nullptr);
    if (rc) return rc;
    p->gated = true;
    struct Ungate {
        cg_plan *p;
        ~Ungate() { p->gated = false; }
    } ungate{p};
    hipEvent_t a, b;
    HIPCHK(ctx, hipEventCreate(&a));
    HIPCHK(ctx, hipEventCreate(&b));
    // Coordinate descent over the grids, finest first: each candidate height timed over `reps`
    // whole refinement steps (graph replays where graphs are on), the other grids as chosen so far.
    auto time_steps = [&](double *ms_out) -> int {
        drop_graphs(p);
        int r;
        if (graph && (r = build_step_graphs(p))) return r;
        auto step = [&]() -> int {
            if (graph) {
                HIPCHK(ctx, hipGraphLaunch(p->step_exec[p->x64cur], ctx->stream));
                p->x64cur = 1 - p->x64cur;
                return CG_OK;
            }
            return refine_step(p, false);
        };
        for (int i = 0; i < 2; ++i)
            if ((r = step())) return r;
        HIPCHK(ctx, hipEventRecord(a, ctx->stream));
        for (int i = 0; i < reps; ++i)
            if ((r = step())) return r;
        HIPCHK(ctx, hipEventRecord(b, ctx->stream));
        HIPCHK(ctx, hipEventSynchronize(b));
        float ms = 0.0f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, a, b));
        *ms_out = ms;
        return CG_OK;
    };
    // Steps before any timing, so the clocks have left their idle state; then each candidate is
    // timed twice, in forward and in reverse order, and its faster time counts, so a clock still
    // ramping favours no end of the list.
    {
        double warm = 0.0;
        for (int i = 0; i < 4 && rc == CG_OK; ++i) rc = time_steps(&warm);
    }
    constexpr int kNumCandidates = (int)(sizeof(kCandidates) / sizeof(kCandidates[0]));
    for (int l = 0; l < p->bottom && rc == CG_OK; ++l) {
        const int keep = plan_tile_rows(p, l);
        double best_ms[kNumCandidates];
        for (int c = 0; c < kNumCandidates; ++c) best_ms[c] = -1.0;
        for (int round = 0; round < 2 && rc == CG_OK; ++round)
            for (int k = 0; k < kNumCandidates; ++k) {
                const int c = round == 0 ? k : kNumCandidates - 1 - k;
                if (kCandidates[c] > p->lv[l].st->down) continue;
                p->tuned_rows[l] = kCandidates[c];
                double ms = 0.0;
                if ((rc = time_steps(&ms))) break;
                if (best_ms[c] < 0.0 || ms < best_ms[c]) best_ms[c] = ms;
            }
        int best_rows = keep;
        double best = -1.0;
        for (int c = 0; c < kNumCandidates; ++c)
            if (best_ms[c] >= 0.0 && (best < 0.0 || best_ms[c] < best)) {
                best = best_ms[c];
                best_rows = kCandidates[c];
            }
        p->tuned_rows[l] = rc ? keep : best_rows;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    drop_graphs(p);
    p->x64_valid = false;
    if (rc) return rc;
    if (rows_out)
        for (int l = 0; l < p->n; ++l) rows_out[l] = l < p->bottom ? p->tuned_rows[l] : 0;
    return CG_OK;
}

double cg_ocean_kernel_time_us(cg_ctx *ctx, int kind, long *launches) {
    if (!ctx) return -1.0;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (kind < 0) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->timed_used = 0;
        ctx->timed_dropped = 0;
        if (launches) *launches = 0;
        return 0.0;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1.0;
    double total = 0.0;
    long counted = 0;
    for (size_t i = 0; i < ctx->timed_used; ++i) {
        if (ctx->timed[i].kind != kind) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ctx->timed[i].a, ctx->timed[i].b) == hipSuccess) {
            total += ms;
            counted++;
        }
    }
    if (launches) *launches = counted;
    return counted ? total * 1000.0 / counted : 0.0;
}

}  // extern "C"
