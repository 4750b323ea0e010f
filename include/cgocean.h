/*
 * cgocean.h: the C ABI of the MI355X (gfx950) ocean-circulation solver.
 *
 * Plain pointers and sizes only. No C++ exception crosses this boundary; every entry point returns
 * an int status (CG_OK, CG_DECLINED, CG_CANCELLED or a negative error) and cg_ocean_last_error()
 * gives the reason of the last failure on a context.
 *
 * Reference interfaces are cited as paths relative to the reference repository:
 *   OA = worldgen/src/commonMain/kotlin/com/cartogenesis/worldgen/pipeline/OceanAccelerator.kt
 *   OC = worldgen/src/commonMain/kotlin/com/cartogenesis/worldgen/pipeline/OceanCirculation.kt
 *   OS = worldgen/src/commonMain/kotlin/com/cartogenesis/worldgen/pipeline/OceanStage.kt
 *   GO = desktop/src/main/kotlin/com/cartogenesis/desktop/GpuOcean.kt
 *   WGC = worldgen/src/commonMain/kotlin/com/cartogenesis/worldgen/model/WorldGenConfig.kt
 *
 * Grids are row-major, cell = row * across + column; columns wrap east to west; row 0 is the
 * northern edge. Masks are one byte per cell, non-zero for water.
 *
 * Threading: a context owns one HIP stream and one mutex; calls on one context are serialised.
 * Use one context per thread for concurrency.
 */
#ifndef CGOCEAN_H
#define CGOCEAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------------- */
#define CG_OK 0
#define CG_DECLINED 1   /* the device cannot take this job; the caller relaxes on the CPU (OA:21-25) */
#define CG_CANCELLED 2  /* the cancel flag was seen set; nothing more was run */
#define CG_ERR_ARG (-1)
#define CG_ERR_DEVICE (-2)
#define CG_ERR_NOMEM (-3)
#define CG_ERR_STATE (-4)

/* ---- solve methods -------------------------------------------------------------------------- */
/* The reference algorithm on the device: binary32 values, binary64 residual, V(2,2) red-black
 * Gauss-Seidel, coarsest grid relaxed to 0.1 x tol (OC:473-493, 740-794). Bit-identical to the
 * processor's solve on the same ladder. */
#define CG_METHOD_REFERENCE 0
/* solveWithKrylovRescue (OC:507-528): the above, then BiCGSTAB preconditioned by the cycle. */
#define CG_METHOD_RESCUE 1
/* Binary32 V(2,2) cycles on the correction, wrapped in binary64 iterative refinement; the
 * outer residual and the accumulated answer are binary64; the coarsest grid is solved directly. */
#define CG_METHOD_REFINED 2
/* Binary64 iterative refinement whose inner solve is flexible GMRES(m) over a binary32 basis,
 * right-preconditioned by one binary32 V(2,2) cycle per iteration. */
#define CG_METHOD_FGMRES 3
/* solveByKrylov (OC:548-607) from the plan's start, with the plan's water bodies, started afresh up
 * to 3 times while it stops short of tol: the heat's solve (OS:486-497). cg_solution.cycles counts
 * the iterations of every start, refinements the starts. */
#define CG_METHOD_KRYLOV 4
/* CG_METHOD_REFINED whose binary32 cycle takes the plan's learned coarse-grid correction
 * (cg_ocean_plan_set_coarse_net): down to the network's grid as usual, then the network's correction
 * for that grid's forcing in place of the grids below it, then up as usual; every step one captured
 * hipGraph, batched plans included. The learned steps stop once the residual is under tol, or as a
 * fallback once it falls by less than a factor 0.6 a step (cg_ocean_plan_net_fallbacks counts those).
 * Plain multigrid then judges every answer: each instance goes on from it with the plain cycle, at
 * least three steps and until CG_METHOD_REFINED's residual and error rule holds with half the error
 * bound (an answer that is not
 * a number starts the solve again, plainly). With CG_OPT_NET_FALLBACK 0 neither happens (the network
 * measured alone). The binary64 fixed point is CG_METHOD_REFINED's. Not a reference method. */
#define CG_METHOD_LEARNED 5
/* CG_METHOD_FGMRES with a caller's preconditioner (cg_ocean_plan_set_preconditioner) on the first
 * inner iteration of every outer step (the outer residual itself) and the binary32 V(2,2) cycle on
 * the others; once an outer step fails to halve the residual, the rest of the solve uses plain
 * cycles only. cg_solution.cycles counts inner iterations of both kinds. Not a reference method. */
#define CG_METHOD_LEARNED_FGMRES 6

/* ---- options (cg_ocean_set_option) ------------------------------------------------------------ */
#define CG_OPT_MIN_CELLS 1       /* relax declines grids with fewer cells (GO:167-172); default 65536 */
#define CG_OPT_TILE_ROWS 2       /* rows per wavefront tile of the fused smoother: 8, 16, 24 or 32 */
#define CG_OPT_USE_GRAPH 3       /* 1: replay each refinement step as a captured hipGraph */
#define CG_OPT_FGMRES_RESTART 4  /* restart length m of CG_METHOD_FGMRES (1..64), default 20 */
#define CG_OPT_FUSED 7           /* 1 (default): fused cycle (one launch per leg per grid, small grids in
                                    one workgroup); 0: separate smoothing/residual/transfer kernels */
#define CG_OPT_LEVEL_PASSES 8    /* passes per grid of the fused binary32 cycle (the unfused one runs two
                                    on every grid), one hex digit per grid from the
                                    finest (lowest digit) up, each 2, 4 or 6; default 0x22224442 (two on
                                    the finest grid, four on the three below it, two below those) */
#define CG_OPT_FUSE_OUTER 9      /* 1 (default): CG_METHOD_REFINED runs the binary64 outer step inside the
                                    finest grid's up-leg launch (fused cycle, two passes there); 0: as a
                                    launch of its own. Both give the same bits. */
#define CG_OPT_DIRECT_BOTTOM 10  /* 1 (default): the binary32 cycle of CG_METHOD_REFINED and
                                    CG_METHOD_FGMRES ends at the largest grid of at most 32 x 64 cells,
                                    solved by its dense inverse (M = -(I - W)^-1 over its water cells,
                                    built when the plan is created, held and applied in binary64) in one
                                    chip-wide launch; 0: that
                                    grid and the coarser ones are cycled in one workgroup down to the
                                    coarsest. The reference method and the Krylov preconditioner
                                    always cycle the whole ladder (OC:740-781). */
#define CG_OPT_STOP_PERMILLE 11  /* CG_METHOD_REFINED, FGMRES and LEARNED take tol x value / 1000 as their
                                    residual target (1..1000, default 1000) */
#define CG_OPT_NET_FALLBACK 13   /* 1 (default): CG_METHOD_LEARNED falls back to the plain cycle and has its
                                    answers judged by plain steps as described there; 0: neither (to
                                    measure the network alone) */
#define CG_OPT_GROUP_MERGED 14   /* 1 (default): cg_ocean_group_solve by CG_METHOD_REFINED steps every plan of
                                    the group together, one captured graph a step whose tile launches each
                                    cover the grids of one depth of every plan's ladder (the leg of every
                                    world and every month in one launch); 0: each plan's own step graph on
                                    a stream of the group's. Both give the same bits. */
#define CG_OPT_ERROR_PICO 12     /* ... and stop only once, besides, the estimated relative L2 error of the
                                    binary64 answer against the exact solution of the operator is at most
                                    value x 1e-12 (default 100: 1e-10; 0: the residual alone decides). The
                                    estimate is rho / (1 - rho) |delta| / |x| for the last correction delta
                                    and its contraction rho against the one before; a residual under tol
                                    alone does not bound the error (a forcing of fine-scale noise leaves a
                                    smooth error whose residual is small). Each instance stops at the first
                                    step where both hold, tested on the device before every step.
                                    CG_METHOD_LEARNED's steps stop on the residual alone; plain steps
                                    then judge its answers (see CG_METHOD_LEARNED). */

typedef struct cg_ctx cg_ctx;
typedef struct cg_plan cg_plan;
typedef uint64_t cg_stencil; /* a device-resident stencil; 0 is never a valid handle */

/* What a solve returned: OceanCirculation.Solution (OC:125) plus the refinement count. */
typedef struct {
    int cycles;               /* V-cycles run (for Krylov methods: preconditioner applications) */
    double relative_residual; /* max|F - Lx| / max|F| of the returned values, binary64 */
    int refinements;          /* outer binary64 residual evaluations (CG_METHOD_REFINED/FGMRES) */
    int status;               /* CG_OK, or CG_CANCELLED */
} cg_solution;

/* ---- device and context -------------------------------------------------------------------- */

/* Probes HIP device `device` and prepares a context: GpuOcean.createOrNull (GO:189-202). On
 * failure returns a negative status, *out = NULL and writes the reason into reason[reason_len]. */
int cg_ocean_open(int device, cg_ctx **out, char *reason, size_t reason_len);
/* Frees the context and every stencil it still holds; destroy its plans first. */
void cg_ocean_close(cg_ctx *ctx);
/* OceanAccelerator.name (OA:18): the device's marketing name. */
int cg_ocean_name(cg_ctx *ctx, char *buf, size_t len);
const char *cg_ocean_last_error(cg_ctx *ctx);
int cg_ocean_set_option(cg_ctx *ctx, int option, long value);
long cg_ocean_get_option(cg_ctx *ctx, int option);

/* ---- stencils ------------------------------------------------------------------------------- */

/* Makes an OceanStencil (OC:24-42) resident on the device: mask, the four neighbour weights and
 * the centre weight. Replaces GpuOcean.holding's identity-keyed LRU (GO:104-137) with an explicit
 * handle. Weights that are equal along every row, finite and non-negative (the circulation's,
 * OC:207-213) are stored once per row and relaxed by the tiled smoother; others once per cell. centre may be NULL when only relaxation is asked for. */
int cg_ocean_stencil_upload(cg_ctx *ctx, int across, int down, const uint8_t *water,
                            const float *east, const float *west, const float *north,
                            const float *south, const double *centre, cg_stencil *out);
/* The handle is invalid once released. The stencil's device memory goes at once unless plans
 * built on it are alive (an LRU like GpuOcean's, GO:109-115, may evict a plan's level); then
 * it goes when the last of those plans is destroyed. */
int cg_ocean_stencil_release(cg_ctx *ctx, cg_stencil stencil);

/* OceanAccelerator.solve (OA:37-41) as GpuOcean.solve runs it (GO:45-102): `passes` red-black
 * passes, colour 0 then colour 1, of x = ((e xE + w xW) + n xN) + s xS - f in binary32 without
 * fused multiply-add, land set to zero, columns wrapping, zero beyond either pole. Bit-identical to
 * OceanCirculation.relax (OC:241-274). `start` and `forcing` are read only; `out` receives the
 * result (it may alias start). Returns CG_DECLINED for an odd or non-positive width, a
 * non-positive height, negative passes or fewer cells than CG_OPT_MIN_CELLS (GO:52-58), and
 * CG_CANCELLED if *cancel is seen non-zero between passes (GO:85-86). */
int cg_ocean_relax(cg_ctx *ctx, cg_stencil stencil, const float *forcing, const float *start,
                   float *out, int passes, const volatile int *cancel);

/* No counterpart in the reference: cg_ocean_relax for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). forcing,
 * start and out are DEVICE pointers of the stencil's cells (out may alias start); the work is only
 * enqueued on `stream` (a hipStream_t; NULL: the context's stream), and the call allocates nothing:
 * the forcing and the two value buffers of the ping-pong live in a static device array of the
 * library (25.2 MB per GPU, held until the process ends), so grids of at most 2^21 cells and 4096
 * columns are accepted (CG_ERR_ARG beyond). Declines as cg_ocean_relax does; no cancellation. Every
 * call on one GPU shares that array: calls that overlap on one GPU (on two streams or two contexts)
 * race, and the caller must order them. */
int cg_ocean_relax_noalloc(cg_ctx *ctx, cg_stencil stencil, const float *forcing, const float *start,
                           float *out, int passes, void *stream);

/* ---- whole solves --------------------------------------------------------------------------- */

/* OceanCirculation.solve / solveWithKrylovRescue (OC:473-528) on the device, over a ladder of
 * resident stencils, finest first, as OceanCirculation.levels builds it (OC:434-462). forcing is
 * the finest grid's f (OceanStencil.forcing); start the initial values (NULL for zeros). tol is
 * the relative max-norm residual to reach; pole_is_wall is prolongAdd's (OC:354-392). out (binary32)
 * and/or out64 (binary64) receive the answer; either may be NULL. requireSolved stays with the
 * caller: an unconverged solve returns CG_OK with info->relative_residual >= tol. */
int cg_ocean_solve(cg_ctx *ctx, const cg_stencil *levels, int nlevels, const float *forcing,
                   const float *start, double tol, int pole_is_wall, int method, float *out,
                   double *out64, cg_solution *info, const volatile int *cancel);

/* A device-resident solve: the ladder, its buffers, the finest forcing and the answer stay in
 * HBM between calls, so repeated solves pay no host transfer. */
int cg_ocean_plan_create(cg_ctx *ctx, const cg_stencil *levels, int nlevels, int pole_is_wall,
                         cg_plan **out);
int cg_ocean_plan_destroy(cg_plan *plan);
/* A plan of `instances` right-hand sides (1..64) on one ladder, solved together: every kernel of the
 * binary32 cycle and the binary64 outer step covers all instances in one launch, so the grids whose
 * launches are latency-bound are paid once for all of them (the months of one world, for instance).
 * Forcings are set per instance; cg_ocean_plan_solve_all solves them all from zero, each instance
 * stopping at its own tolerance check. Calls that name no instance act on instance 0; the
 * single-instance calls (solve, set_stress, set_bodies, set_start, velocities, correct_bodies)
 * return CG_ERR_STATE on a plan of several. */
int cg_ocean_plan_create_batch(cg_ctx *ctx, const cg_stencil *levels, int nlevels, int pole_is_wall,
                               int instances, cg_plan **out);
int cg_ocean_plan_set_forcing_at(cg_plan *plan, int instance, const float *forcing);
/* CG_METHOD_REFINED over every instance; infos[i] receives instance i's result. */
int cg_ocean_plan_solve_all(cg_plan *plan, double tol, int method, cg_solution *infos);
int cg_ocean_plan_read_at(cg_plan *plan, int instance, double *out64, float *out32);
/* Uploads the finest forcing (host pointer). */
int cg_ocean_plan_set_forcing(cg_plan *plan, const float *forcing);
/* Solves from zero (or from the plan's current answer if warm != 0) to tol. */
int cg_ocean_plan_solve(cg_plan *plan, double tol, int method, int warm, cg_solution *info);
/* Solves several plans of one context one after another, each from zero, to one tol (plans of one
 * instance; right-hand sides on one ladder go faster as one batched plan, below). */
int cg_ocean_plan_solve_batch(cg_plan *const *plans, int count, double tol, int method,
                              cg_solution *infos);
/* A group of plans of one context solved together by CG_METHOD_REFINED, from zero: the worlds of
 * uneven size of one batch, each a plan of its own ladder (batched: its months). No reference
 * counterpart (the reference solves one grid per call, OceanStage.kt:436-443). One host loop batches
 * the steps of every instance of every plan between reads of their step gates. With
 * CG_OPT_GROUP_MERGED (the default) the plans still running step together: a refinement step of
 * them is one captured hipGraph launch, each of whose tile launches covers the grids of one depth of
 * every such plan's ladder (every world, every month), then one launch for every plan's direct
 * bottom, then the up legs the same way; otherwise each plan's own step graph is launched on a
 * stream of the group's. Either way a plan whose instances have all stopped gets no more launches. Each instance
 * stops at the step it would stop at alone, with the same answer to the bit. A plan may belong to
 * several groups (at most 16 plans a group) and cannot be destroyed while it does. */
typedef struct cg_group cg_group;
int cg_ocean_group_create(cg_plan *const *plans, int count, cg_group **out);
/* method: CG_METHOD_REFINED, or CG_METHOD_LEARNED (every plan with its coarse network set; the
 * fallback and the plain judge of the answers per plan, as for cg_ocean_plan_solve_all). infos: one per instance, the
 * plans' instances in the order of the plans given to create. */
int cg_ocean_group_solve(cg_group *group, double tol, int method, cg_solution *infos);
int cg_ocean_group_destroy(cg_group *group);
/* The finest grid's water bodies (OceanCirculation.WaterBodies, OC:670-697: body per cell, -1 on
 * land, `count` bodies) for correctBodyMeans (OC:709-731), which then follows every cycle of
 * CG_METHOD_REFERENCE/RESCUE and every preconditioning cycle of RESCUE/KRYLOV. count 0 removes them.
 * The per-body residual sums are deterministic tree sums, not the reference's sequential sums. */
int cg_ocean_plan_set_bodies(cg_plan *plan, const int32_t *body_of, int count);
/* correctBodyMeans on the device alone: values (host) against the plan's finest forcing, into out. */
int cg_ocean_plan_correct_bodies(cg_plan *plan, const float *values, float *out);
/* Sets the answer a warm solve starts from (binary32; the binary64 answer becomes its widening). */
int cg_ocean_plan_set_start(cg_plan *plan, const float *start);
/* Copies the answer back; either pointer may be NULL. */
int cg_ocean_plan_read(cg_plan *plan, double *out64, float *out32);
/* Synchronises the context's stream and reports the average duration, in microseconds, of the
 * finest grid's two-pass fused smoother launches of `kind` (0: from zero, the cycle's down leg;
 * 1: from given values, its up leg and the relaxation seam) recorded with HIP events on the stream
 * they were launched on while option CG_OPT_TIME_KERNELS was 1 (graph replay is then off); kind -1
 * resets the record. *launches receives the count averaged. */
#define CG_OPT_TIME_KERNELS 6
double cg_ocean_kernel_time_us(cg_ctx *ctx, int kind, long *launches);

/* Times `reps` back-to-back launches of one of the plan's kernels between two HIP events on the
 * context's stream and returns the average in microseconds (-1 on failure): kind 0 the finest
 * grid's fused down leg, 1 its fused up leg, 2 the binary64 outer residual, 3 the finest up leg
 * with the outer step fused in (as most steps run it: without the error estimate's sums), 4 the
 * bottom of the cycle (the dense-inverse launch with CG_OPT_DIRECT_BOTTOM, else the one-workgroup
 * cycle of the coarsest grids), 5 kind 3 with the sums (as a solve's last steps run it), 6 kind 3
 * over every instance of a batched plan in one launch (each instance active), 16 + 2 l grid l's
 * down leg and 17 + 2 l its up leg (without the outer step), for grids above the bottom. Each
 * launch reads the plan's current buffers and writes the same outputs (never the answer), so
 * repeats leave the plan as one launch does; call it between solves, not during one. */
double cg_ocean_plan_time_kernel(cg_plan *plan, int kind, int reps);

/* ---- the round's input and output steps, on the device ------------------------------------- */

/* OceanStage.waterOn (OS:551-573): the solve grid's water mask (across x down) from a map's land mask
 * (map_land, map_across x map_down, non-zero for land): a solve cell is water when every map cell it
 * overlaps by more than 1e-6 of a map cell is water. Host buffers in and out. */
int cg_ocean_water_on(cg_ctx *ctx, const uint8_t *map_land, int map_across, int map_down, int across,
                      int down, uint8_t *water);
/* OceanStage.curlForcing (OS:689-708): forcing = curl_z(tau) / (rho H) per cell in binary64, tau
 * given eastward and northward in N/m^2, central differences over cells of width_m x height_m,
 * columns wrapping, edge rows one-sided. depth NULL takes the reference's 500 m wind-driven layer
 * (OS:74); otherwise H = depth[cell]. Host buffers in and out. */
int cg_ocean_curl_forcing(cg_ctx *ctx, int across, int down, double width_m, double height_m,
                          const double *eastward, const double *northward, const double *depth,
                          double *forcing);
/* The same curl made the plan's finest forcing on the device, f = (float)(F / c) on water and 0 on
 * land (withBalance, OC:222-231): the circulation half of OceanStage.circulateUnder (OS:427-432)
 * with only the stress crossing the bus. */
int cg_ocean_plan_set_stress(cg_plan *plan, double width_m, double height_m, const double *eastward,
                             const double *northward, const double *depth);
/* OceanCirculation.velocities (OC:801-828) of the plan's last answer rounded to binary32, computed
 * on the device; either output may be NULL. */
int cg_ocean_plan_velocities(cg_plan *plan, double width_m, double height_m, float *eastward,
                             float *northward);

/* The plan's tile autotuner: for every grid of the fused cycle above the bottom, finest first, times
 * `reps` whole refinement steps (graph replays, HIP events on the context's stream) with that grid's
 * tiles 2, 4, 6, 8, 16, 24 and 32 rows high and the other grids as chosen so far, and keeps the fastest;
 * the plan's later solves use the choice (results do not depend on the tiles). rows_out (NULL or
 * nlevels ints) receives the height per grid, 0 where a grid is not tiled. Call it between solves;
 * it leaves the plan without an answer to start a warm solve from. */
int cg_ocean_plan_autotune(cg_plan *plan, int reps, int *rows_out);
/* Sets the tile height of every grid of the plan (nlevels ints, each 0 for the context's rule, or 2,
 * 4, 6, 8, 16, 24 or 32), as the autotuner would: to repeat a tuned configuration. */
int cg_ocean_plan_set_tile_rows(cg_plan *plan, const int *rows);

/* ---- host-side problem builders (no device needed) --------------------------------------------- */

/* BOTTOM_DRAG_PER_S (OS:102-103) with Earth's rotation rate as the reference's Float (WGC:302). */
double cg_bottom_drag(void);
/* beta per row, 2 Omega cos(phi) / a, with phi the row's centre latitude computed in binary32
 * (ClimateStage.latitudeOf, ClimateStage.kt:1313-1314; WGC:246, 272-273, 302). */
int cg_beta_rows(int down, double world_width_km, double *beta);

/* OceanCirculation.stencil then withBalance (OC:161-231). balance may be NULL (zero forcing). */
int cg_circulation_stencil(int across, int down, double dx, double dy, const uint8_t *water,
                           const double *beta, double drag, const double *balance, float *east,
                           float *west, float *north, float *south, double *centre,
                           float *forcing);
/* OceanCirculation.coarseWater (OC:404-423). */
int cg_coarse_water(const uint8_t *fine, int fine_across, int fine_down, int coarse_across,
                    int coarse_down, int every_cell_water, uint8_t *coarse);
/* The direct bottom's inverse, M = -(I - W)^-1 over a grid's water cells (no counterpart in the
 * reference, whose cycle relaxes its coarsest grid, OC:784-794): per-cell weights as an
 * OceanStencil holds them; writes the water cells (row-major) to idx[n] and M row-major (binary64,
 * by banded LU of the M-matrix I - W) to m[n * n]. Returns n; 0 when there is no water, more than
 * max_n water cells, or I - W has a pivot that is not positive. */
int cg_bottom_inverse(int across, int down, const uint8_t *water, const float *east, const float *west,
                      const float *north, const float *south, int max_n, int threads, int32_t *idx,
                      double *m);
/* The level shapes OceanCirculation.levels walks (OC:434-462): returns the count (<= max_levels)
 * and fills across/down/cell width/cell height per level, finest first. */
int cg_level_shapes(int across, int down, double dx, double dy, int max_levels, int *level_across,
                    int *level_down, double *cell_width, double *cell_height);

/* OceanHeat.diffusivity (OceanHeat.kt:128-129): K = V R_d in m^2/s at a latitude on a planet. */
double cg_heat_diffusivity(double latitude_deg, double planet_radius_m);
/* OceanHeat.stencil then withBalance (OceanHeat.kt:151-228): per-cell exponentially fitted
 * (Scharfetter-Gummel) weights of the heat's advection-diffusion-relaxation balance on one grid.
 * stream is psi (binary32); target the temperature each cell relaxes to over relaxation_s; with_target
 * 0 gives a zero forcing (a coarse grid's). The diffusivity is constant_k where that is positive,
 * else cg_heat_diffusivity at each row's latitude on a planet of planet_radius_m. entrainment (w/h per
 * second) and subsurface (its temperature) may be NULL. */
int cg_heat_stencil(int across, int down, double dx, double dy, const uint8_t *water, const float *stream,
                    const float *target, double relaxation_s, int with_target, double planet_radius_m,
                    double constant_k, const float *entrainment, const float *subsurface, float *east,
                    float *west, float *north, float *south, double *centre, float *forcing);
/* OceanCirculation.waterBodies (OC:670-697): returns the count, fills body_of (-1 on land). */
int cg_water_bodies(const uint8_t *water, int across, int down, int32_t *body_of);
/* OceanStage.resample (OS:1199-1223): a field read bilinearly at every cell centre of another grid
 * covering the same map (the heat's coarse streams). */
int cg_resample(const float *field, int from_across, int from_down, int to_across, int to_down,
                float *out);
/* OceanStage.coarseMean (OS:857-860): block means of a rate per unit area (the heat's coarse
 * entrainment). */
int cg_coarse_mean(const float *fine, int fine_across, int fine_down, int coarse_across,
                   int coarse_down, float *coarse);

/* ---- learned preconditioner ------------------------------------------------------------------ */
/* No counterpart in the reference (its solve has no learned part). A U-Net of 3x3 convolutions on
 * bfloat16 NHWC activations with binary32 accumulation on the matrix cores. Activations:
 * act[((b * down + row) * across + col) * channels + c]; weights [9][cout][cin] bfloat16, tap =
 * 3 * (dr + 1) + (dc + 1); rows are zero-padded and columns wrap. Channel counts are multiples of
 * 16, across a multiple of 16. Pointers are device pointers; `stream` is a hipStream_t (NULL: the
 * default stream); every call only enqueues work. */

/* The learned coarse-grid correction of CG_METHOD_LEARNED: a 4-level U-Net of 3x3 convolutions with
 * b = base channels (a multiple of 16), in this order: enc0a 16->b, enc0b b->b, (2x2 mean pool)
 * enc1a b->2b, enc1b, (pool) enc2a 2b->4b, enc2b, (pool) mid_a 4b->8b, mid_b, dec2a (mid_b upsampled,
 * 8b, and enc2b, 4b) -> 4b, dec2b, dec1a (dec2b upsampled, 4b, and enc1b, 2b) -> 2b, dec1b,
 * dec0a (dec1b upsampled, 2b, and enc0b, b) -> b, dec0b, out b->16 (every layer ReLU but out; nearest
 * upsampling; the layout of cg_conv3x3), on grid `level` of the plan
 * (1 .. the fused cycle's bottom grid; its rows a multiple of 8 and columns of 128). Input per cell:
 * the grid's mask, depth[cell] / 6000 m (depth: host, the grid's cells in metres, NULL for zeros), and
 * the grid's binary32 forcing over its root mean square; the correction is out_scale x that rms x the
 * output's channel 0 on water, zero on land. weights: host bfloat16, every layer's [9][cout][cin] in
 * table order, count = cg_unet_weight_count(base). Copied into the plan, which owns the network's
 * buffers for all its instances; weights NULL removes the network. */
int cg_ocean_plan_set_coarse_net(cg_plan *plan, int level, int base, const uint16_t *weights, size_t count,
                                 const float *depth, double out_scale);
/* The bfloat16 weight count of the U-Net of `base` channels (-1 for a base it does not take). */
int cg_unet_weight_count(int base);
/* The plan's coarse network on one forcing of its grid (host, the grid's cells), its correction into
 * out (host): what the network contributes to a learned step. Between solves only. */
int cg_ocean_plan_apply_coarse_net(cg_plan *plan, const float *forcing, float *out);
/* Training data for the coarse network: the plain refined method from zero on a plan of one instance,
 * step by step, with grid `level`'s forcing after every step copied into out (most_steps x the grid's
 * cells, host) until the residual is under tol; returns the steps captured. Leaves the plan without an
 * answer to start a warm solve from. */
int cg_ocean_plan_capture_coarse(cg_plan *plan, int level, double tol, int most_steps, float *out);
/* How many times the last CG_METHOD_LEARNED solve of the plan fell back to the plain cycle. */
int cg_ocean_plan_net_fallbacks(cg_plan *plan);

/* The preconditioner callback of CG_METHOD_LEARNED_FGMRES: z = M(v) for one binary32 vector of the
 * plan's finest grid (cells values), enqueued on `stream`; returns CG_OK or an error. */
typedef int (*cg_precond_fn)(void *user, const float *v, float *z, void *stream);
/* Sets (fn non-NULL) or clears the plan's learned preconditioner. Plans of one instance only. */
int cg_ocean_plan_set_preconditioner(cg_plan *plan, cg_precond_fn fn, void *user);

/* y = relu?(conv(concat(x1[c1], x2[c2]), w)) with cout output channels: channels [0, o1) into y1
 * (o1 per pixel) and [o1, cout) into y2; or, with y32 non-NULL, all of them into y32 as binary32. */
int cg_conv3x3(const uint16_t *x1, int c1, const uint16_t *x2, int c2, int n, int down, int across,
               const uint16_t *w, int cout, int relu, uint16_t *y1, int o1, uint16_t *y2, float *y32,
               void *stream);
/* cg_conv3x3 of a 16-channel input whose channels 3..15 are zero (the network's input: mask, depth,
 * vector), reading channels 0..2 only: values in channels 3..15 are ignored (never multiplied, so not
 * even a NaN there reaches the output), and the result equals cg_conv3x3's only when they are zero.
 * Sums in binary32: on the matrix cores (K = 9 taps x channels 0..3, channel 3 zeroed) when cout is a
 * multiple of 16 and w 8-byte aligned, else on the vector units. w is [9][cout][16], cout a multiple
 * of 8 up to 32; bfloat16 output. x must be 8-byte and y 16-byte aligned (CG_ERR_ARG otherwise). */
int cg_conv3x3_in3(const uint16_t *x, int n, int down, int across, const uint16_t *w, int cout, int relu, uint16_t *y,
                   void *stream);
/* cg_conv3x3 whose first input x1 is down/2 x across/2 and read through nearest upsampling (the
 * U-Net decoder's input without the upsampled copy). Needs c1 and c2 multiples of 32, across % 64
 * == 0 and down % 4 == 0. */
int cg_conv3x3_up(const uint16_t *x1, int c1, const uint16_t *x2, int c2, int n, int down, int across,
                  const uint16_t *w, int cout, int relu, uint16_t *y1, int o1, uint16_t *y2, float *y32,
                  void *stream);
/* The data gradient of cg_conv3x3 with weights w[9][cout][cin]: dx = conv^T(dy), split as the
 * forward's inputs were (channels [0, c1) into dx1, the rest into dx2), or binary32 into dx32. */
int cg_conv3x3_adjoint(const uint16_t *dy, int cout, int n, int down, int across, const uint16_t *w, int cin,
                       uint16_t *dx1, int c1, uint16_t *dx2, float *dx32, void *stream);
/* wt[9][cin][cout] = w[9][cout][cin] with the taps mirrored: cg_conv3x3 of dy with wt (cin output
 * channels) is cg_conv3x3_adjoint of dy with w, on the forward kernels. */
int cg_conv3x3_flip_weights(const uint16_t *w, int cout, int cin, uint16_t *wt, void *stream);
/* The weight gradient dw[9][cout][c1 + c2] (binary32) of cg_conv3x3 for input x and output gradient
 * dy, summed over all pixels in a fixed order. work holds slices x 9 x cout x (c1 + c2) floats,
 * slices = cg_conv3x3_wgrad_slices(n, down, across). */
int cg_conv3x3_wgrad_slices(int n, int down, int across);
int cg_conv3x3_wgrad(const uint16_t *x1, int c1, const uint16_t *x2, int c2, const uint16_t *dy, int cout, int n,
                     int down, int across, float *work, size_t work_floats, float *dw, void *stream);
/* y = scale x the mean of each 2x2 block (y is down/2 x across/2). */
int cg_pool2(const uint16_t *x, int channels, int n, int down, int across, float scale, uint16_t *y, void *stream);
/* y (2 down x 2 across) = scale x the nearest coarse value. */
int cg_upsample2(const uint16_t *x, int channels, int n, int down, int across, float scale, uint16_t *y,
                 void *stream);
/* dy = 0 wherever y <= 0 (the gradient through a ReLU with output y), in place. */
int cg_relu_mask(uint16_t *dy, const uint16_t *y, long count, void *stream);
/* out = a + b elementwise (out may alias a or b). */
int cg_add_bf16(const uint16_t *a, const uint16_t *b, long count, uint16_t *out, void *stream);
/* The network input: 16 channels per cell; 0..1 from fixed[cell][2] (mask, scaled depth), channel
 * 2 = scale x v[cell], the rest zero. */
int cg_unet_pack_input(const uint16_t *fixed, const float *v, float scale, long cells, uint16_t *out,
                       void *stream);
/* z[cell] = water[cell] ? scale x y32[cell * 16] : 0 (channel 0 of the network's output). */
int cg_unet_unpack_output(const float *y32, const uint8_t *water, float scale, long cells, float *z,
                          void *stream);

/* Synthetic circulation inputs: a seeded 4-octave value-noise land mask thresholded so that
 * `land_fraction` of cells are land, depth H ~ U(200, 6000) m on water, and the balance
 * curl_z(tau) / (rho H) of a zonal stress tau_x = amplitude * cos(3 (lat - shift)) plus Gaussian
 * noise, by OceanStage.curlForcing's differences (OS:689-708). Not a reference function: the
 * reference's stress comes from its climate stage. month in 0..11 sets
 * amplitude = 0.1 cos(2 pi month / 12) and shift = 10 deg sin(2 pi month / 12); month < 0 gives
 * the annual case, amplitude 0.1 and no shift. Any of the outputs may be NULL. */
int cg_problem_generate(int across, int down, uint64_t seed, double land_fraction, int month,
                        uint8_t *water, double *depth, double *balance);
/* The synthetic workload's wind stress that cg_problem_generate's balance is the curl of: eastward
 * tau_x per cell as described there, northward zero. Either output may be NULL. */
int cg_problem_stress(int across, int down, uint64_t seed, int month, double *eastward,
                      double *northward);

#ifdef __cplusplus
}
#endif

#endif /* CGOCEAN_H */
