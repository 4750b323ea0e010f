/*
 * ORACLE: test infrastructure only.
 *
 * A plain-C restatement of the reference's ocean-circulation multigrid, used as the checker for the
 * HIP library and as the CPU baseline's timed code. It is never linked into, loaded by or called
 * from the product library; only tests, the smoke check and the benchmark's CPU leg use it.
 *
 * Every function cites the reference file and lines it restates. Paths are relative to the
 * reference repository; OC = worldgen/src/commonMain/kotlin/com/cartogenesis/worldgen/pipeline/
 * OceanCirculation.kt, OS = .../pipeline/OceanStage.kt.
 *
 * Arithmetic model. The reference runs on the JVM, whose float and double arithmetic is strict IEEE
 * binary32/binary64 with round-to-nearest and no fused multiply-add. This file is compiled with
 * -O2 -ffp-contract=off on x86-64 (SSE arithmetic, no x87), which gives the same roundings for
 * every +, -, *, / below, in the same order as the Kotlin source evaluates them (left to right).
 * The only library calls are cos and expm1 (glibc), which may differ from the JVM's Math by at most
 * one ulp; that moves a weight at a rounding boundary only.
 *
 * Threading mirrors the reference's parallelChunks (Parallel.jvm.kt:27-45): row bands over disjoint
 * writes, so results are bit-identical at any thread count. The sequential reductions (restrict,
 * dot) stay sequential, as in the reference.
 *
 * Parity pinning of each function is recorded in the oracle's test file, tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_SERIES_BELOW_CELL_PECLET 1e-4   /* OC:88 */
#define OR_PASSES_EACH_WAY 2               /* OC:101 */
#define OR_COARSEST_ROWS 8                 /* OC:104 */
#define OR_COARSEST_TOLERANCE_SHARE 0.1    /* OC:110 */
#define OR_COARSEST_PASSES_PER_CHECK 20    /* OC:113 */
#define OR_MOST_CYCLES 200                 /* OC:116 */
#define OR_SQUARE_ENOUGH 1.25              /* OC:122 */
#define OR_KRYLOV_RESTARTS 3               /* OC:535 */
#define OR_PRECONDITIONER_COARSEST_PASSES 200 /* OC:629 */
#define OR_MAX_LEVELS 24

static int or_threads = 1;

void or_set_threads(int n) { or_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return or_threads; }

/* ------------------------------------------------------------------------------------------ */
/* Stencil construction                                                                        */
/* ------------------------------------------------------------------------------------------ */

/*
 * OceanCirculation.stencil (OC:161-219) followed by withBalance (OC:222-231).
 * Weights per cell (row-constant), the centre weight in double, and the forcing f = balance/centre
 * rounded to float on water, zero on land.
 */
void or_stencil(int across, int down, double dx, double dy, const uint8_t *water,
                const double *beta, double drag, const double *balance,
                float *east, float *west, float *north, float *south, double *centre, float *forcing)
{
    double along_y = drag / (dy * dy);
    double width_sq = dx * dx;
    for (int row = 0; row < down; row++) {
        double b = beta[row];
        double peclet = fabs(b) * dx / (2.0 * drag);
        double east_x, west_x;
        if (peclet < OR_SERIES_BELOW_CELL_PECLET) {
            double fitted = drag * (1.0 + peclet * peclet / 3.0) / width_sq;
            double upstream = fabs(b) / (2.0 * dx);
            east_x = fitted + upstream;
            west_x = fitted - upstream;
        } else {
            double scale = drag * peclet / width_sq;
            double coth_less_one = 2.0 / expm1(2.0 * peclet);
            east_x = scale * (2.0 + coth_less_one);
            west_x = scale * coth_less_one;
        }
        double north_y = row == 0 ? 0.0 : along_y;
        double south_y = row == down - 1 ? 0.0 : along_y;
        double c = east_x + west_x + 2.0 * along_y + (along_y - north_y) + (along_y - south_y);
        for (long cell = (long)row * across; cell < (long)(row + 1) * across; cell++) {
            east[cell] = (float)(east_x / c);
            west[cell] = (float)(west_x / c);
            north[cell] = (float)(north_y / c);
            south[cell] = (float)(south_y / c);
            centre[cell] = c;
        }
    }
    long cells = (long)across * down;
    for (long cell = 0; cell < cells; cell++)
        forcing[cell] = water[cell] ? (float)(balance[cell] / centre[cell]) : 0.0f;
}

/* withBalance (OC:222-231): the same operator with another right-hand side. */
void or_with_balance(long cells, const uint8_t *water, const double *centre, const double *balance,
                     float *forcing)
{
    for (long cell = 0; cell < cells; cell++)
        forcing[cell] = water[cell] ? (float)(balance[cell] / centre[cell]) : 0.0f;
}

/* balanceOf (OC:325-327). */
void or_balance_of(long cells, const float *forcing, const double *centre, double *out)
{
    for (long cell = 0; cell < cells; cell++) out[cell] = (double)forcing[cell] * centre[cell];
}

/* ------------------------------------------------------------------------------------------ */
/* Relaxation, residual and transfers                                                          */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int across, down;
    const uint8_t *water;
    const float *e, *w, *n, *s;
    const float *f;
    const double *c;
} or_st;

/* updated (OC:263-274): ((e xE + w xW) + n xN) + s xS - f, in binary32. */
static inline float or_updated(const or_st *st, const float *x, int row, int col)
{
    int across = st->across;
    long cell = (long)row * across + col;
    int ce = col + 1 == across ? 0 : col + 1;
    int cw = col == 0 ? across - 1 : col - 1;
    float vn = row > 0 ? x[cell - across] : 0.0f;
    float vs = row + 1 < st->down ? x[cell + across] : 0.0f;
    return st->e[cell] * x[(long)row * across + ce] + st->w[cell] * x[(long)row * across + cw] +
           st->n[cell] * vn + st->s[cell] * vs - st->f[cell];
}

/* relax (OC:241-257): passes x {colour 0, colour 1}, in place, land set to zero. */
static void or_relax_st(const or_st *st, float *x, int passes)
{
    for (int p = 0; p < passes; p++) {
        for (int colour = 0; colour <= 1; colour++) {
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
            for (int row = 0; row < st->down; row++) {
                for (int col = (colour - row) & 1; col < st->across; col += 2) {
                    long cell = (long)row * st->across + col;
                    x[cell] = st->water[cell] ? or_updated(st, x, row, col) : 0.0f;
                }
            }
        }
    }
}

void or_relax(int across, int down, const uint8_t *water, const float *e, const float *w,
              const float *n, const float *s, const float *f, float *x, int passes)
{
    or_st st = {across, down, water, e, w, n, s, f, NULL};
    or_relax_st(&st, x, passes);
}

/* residual (OC:292-315): (x - (sum w x - f)) * centre in binary64, zero on land. */
static void or_residual_st(const or_st *st, const float *x, double *r)
{
    int across = st->across, down = st->down;
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
    for (int row = 0; row < down; row++) {
        for (int col = 0; col < across; col++) {
            long cell = (long)row * across + col;
            if (!st->water[cell]) { r[cell] = 0.0; continue; }
            int ce = col + 1 == across ? 0 : col + 1;
            int cw = col == 0 ? across - 1 : col - 1;
            double vn = row > 0 ? (double)x[cell - across] : 0.0;
            double vs = row + 1 < down ? (double)x[cell + across] : 0.0;
            double bal = (double)st->e[cell] * (double)x[(long)row * across + ce] +
                         (double)st->w[cell] * (double)x[(long)row * across + cw] +
                         (double)st->n[cell] * vn + (double)st->s[cell] * vs - (double)st->f[cell];
            r[cell] = ((double)x[cell] - bal) * st->c[cell];
        }
    }
}

void or_residual(int across, int down, const uint8_t *water, const float *e, const float *w,
                 const float *n, const float *s, const float *f, const double *c, const float *x,
                 double *r)
{
    or_st st = {across, down, water, e, w, n, s, f, c};
    or_residual_st(&st, x, r);
}

/* largest (OC:318-322). NaN never wins a max, as in kotlin.math.max ... except it does:
 * Kotlin's max(a, NaN) is NaN. That is restated: a NaN anywhere makes the result NaN. */
double or_largest(long n, const double *v)
{
    double m = 0.0;
    for (long i = 0; i < n; i++) {
        double a = fabs(v[i]);
        if (isnan(a) || isnan(m)) m = NAN;
        else if (a > m) m = a;
    }
    return m;
}

/* restrict (OC:334-346): sequential row-major accumulation of fine * share. */
void or_restrict(const double *fine, int fa, int fd, int ca, int cd, double *coarse)
{
    int cpc = fa / ca, rpc = fd / cd;
    double share = 1.0 / (double)(cpc * rpc);
    memset(coarse, 0, sizeof(double) * (size_t)ca * cd);
    for (int row = 0; row < fd; row++) {
        int crow = row / rpc;
        for (int col = 0; col < fa; col++)
            coarse[(long)crow * ca + col / cpc] += fine[(long)row * fa + col] * share;
    }
}

/* prolongAdd (OC:354-392): bilinear between coarse centres, binary32, water cells only. */
void or_prolong_add(const float *coarse, int ca, int cd, float *fine, int fa, int fd,
                    const uint8_t *water, int pole_is_wall)
{
    int cpc = fa / ca, rpc = fd / cd;
    float beyond = pole_is_wall ? -1.0f : 1.0f;
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
    for (int row = 0; row < fd; row++) {
        float crow = ((float)row + 0.5f) / (float)rpc - 0.5f;
        int above = (int)floorf(crow);
        float db = crow - (float)above;
        for (int col = 0; col < fa; col++) {
            long cell = (long)row * fa + col;
            if (!water[cell]) continue;
            float ccol = ((float)col + 0.5f) / (float)cpc - 0.5f;
            int left = (int)floorf(ccol);
            float ab = ccol - (float)left;
            int wc = (left + ca) % ca;
            int ec = (left + 1) % ca;
            float tw, te, bw, be;
#define OR_AT(rr, cc) ((rr) < 0 ? beyond * coarse[(cc)] : (rr) >= cd ? beyond * coarse[(long)(cd - 1) * ca + (cc)] : coarse[(long)(rr) * ca + (cc)])
            tw = OR_AT(above, wc);
            te = OR_AT(above, ec);
            bw = OR_AT(above + 1, wc);
            be = OR_AT(above + 1, ec);
#undef OR_AT
            float top = tw * (1.0f - ab) + te * ab;
            float bottom = bw * (1.0f - ab) + be * ab;
            fine[cell] += top * (1.0f - db) + bottom * db;
        }
    }
}

/* coarseWater (OC:404-423). */
void or_coarse_water(const uint8_t *fine, int fa, int fd, int ca, int cd, int every_cell_water,
                     uint8_t *coarse)
{
    int cpc = fa / ca, rpc = fd / cd;
    uint8_t ev = every_cell_water ? 1 : 0;
    for (long i = 0; i < (long)ca * cd; i++) coarse[i] = ev;
    for (int row = 0; row < fd; row++)
        for (int col = 0; col < fa; col++)
            if ((fine[(long)row * fa + col] ? 1 : 0) != ev) coarse[(long)(row / rpc) * ca + col / cpc] = !ev;
}

/*
 * The shapes of levels (OC:434-462): writes up to max_levels (across, down) pairs, finest first,
 * and the cell width and height of each, and returns the count.
 */
int or_level_shapes(int across, int down, double dx, double dy, int max_levels, int *acr, int *dwn,
                    double *wm, double *hm)
{
    int count = 0;
    acr[0] = across; dwn[0] = down; wm[0] = dx; hm[0] = dy; count = 1;
    double width = dx, height = dy;
    while (count < max_levels) {
        int wide = width > height * OR_SQUARE_ENOUGH;
        int tall = height > width * OR_SQUARE_ENOUGH;
        int na = wide ? across : across / 2;
        int nd = tall ? down : down / 2;
        int even = (na == across || across % 2 == 0) && (nd == down || down % 2 == 0);
        if (!even || nd < OR_COARSEST_ROWS || na % 2 != 0) break;
        width *= (double)across / na;
        height *= (double)down / nd;
        across = na; down = nd;
        acr[count] = across; dwn[count] = down; wm[count] = width; hm[count] = height;
        count++;
    }
    return count;
}

/* ------------------------------------------------------------------------------------------ */
/* The multigrid solve, binary32 values (the reference path)                                   */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int count;
    or_st st[OR_MAX_LEVELS];
    const int32_t *body_of; /* waterBodies of the finest grid, or NULL (OC:670-697) */
    int body_count;
} or_ladder;

static void or_ladder_from(or_ladder *L, int count, const int *acr, const int *dwn,
                           const uint8_t *const *water, const float *const *e, const float *const *w,
                           const float *const *n, const float *const *s, const float *const *f,
                           const double *const *c)
{
    L->count = count;
    L->body_of = NULL;
    L->body_count = 0;
    for (int i = 0; i < count; i++) {
        or_st st = {acr[i], dwn[i], water[i], e[i], w[i], n[i], s[i], f[i], c[i]};
        L->st[i] = st;
    }
}

/* relaxToConvergence (OC:784-794). */
static void or_relax_to_convergence_st(const or_st *st, float *x, double tol, double *scratch)
{
    long cells = (long)st->across * st->down;
    or_balance_of(cells, st->f, st->c, scratch);
    double largest = or_largest(cells, scratch);
    if (!(largest > 0.0)) return;
    int passes = 0;
    while (passes < OR_MOST_CYCLES * OR_COARSEST_PASSES_PER_CHECK) {
        or_residual_st(st, x, scratch);
        if (!(or_largest(cells, scratch) > largest * tol)) break;
        or_relax_st(st, x, OR_COARSEST_PASSES_PER_CHECK);
        passes += OR_COARSEST_PASSES_PER_CHECK;
    }
}

void or_relax_to_convergence(int across, int down, const uint8_t *water, const float *e,
                             const float *w, const float *n, const float *s, const float *f,
                             const double *c, float *x, double tol)
{
    or_st st = {across, down, water, e, w, n, s, f, c};
    double *scratch = malloc(sizeof(double) * (size_t)across * down);
    or_relax_to_convergence_st(&st, x, tol, scratch);
    free(scratch);
}

/* correctBodyMeans (OC:709-731): shift each body by minus its summed residual over its summed margin
 * c (1 - (e + w + n + s)), both summed sequentially in cell order in binary64; the shift is rounded to
 * binary32 and subtracted in binary32. */
static void or_correct_body_means(const or_st *st, float *x, const int32_t *body_of, int count)
{
    long cells = (long)st->across * st->down;
    double *r = malloc(sizeof(double) * (size_t)cells);
    double *rs = calloc((size_t)(count > 0 ? count : 1), sizeof(double));
    double *ms = calloc((size_t)(count > 0 ? count : 1), sizeof(double));
    or_residual_st(st, x, r);
    for (long cell = 0; cell < cells; cell++) {
        int b = body_of[cell];
        if (b < 0) continue;
        double weights = (double)st->e[cell] + (double)st->w[cell] + (double)st->n[cell] + (double)st->s[cell];
        rs[b] += r[cell];
        ms[b] += st->c[cell] * (1.0 - weights);
    }
    for (long cell = 0; cell < cells; cell++) {
        int b = body_of[cell];
        if (b < 0 || ms[b] <= 0.0) continue;
        x[cell] -= (float)(rs[b] / ms[b]);
    }
    free(r);
    free(rs);
    free(ms);
}

void or_correct_body_means_flat(int across, int down, const uint8_t *water, const float *e, const float *w,
                                const float *n, const float *s, const float *f, const double *c, float *x,
                                const int32_t *body_of, int count)
{
    or_st st = {across, down, water, e, w, n, s, f, c};
    or_correct_body_means(&st, x, body_of, count);
}

/* waterBodies (OC:670-697): four-connected labels, columns wrapping, numbered in the order of each
 * body's first cell; -1 on land. Returns the count. */
int or_water_bodies(const uint8_t *water, int across, int down, int32_t *body_of)
{
    long cells = (long)across * down;
    int32_t *pending = malloc(sizeof(int32_t) * (size_t)(cells > 0 ? cells : 1));
    int count = 0;
    for (long i = 0; i < cells; i++) body_of[i] = -1;
    for (long seed = 0; seed < cells; seed++) {
        if (!water[seed] || body_of[seed] >= 0) continue;
        long top = 0;
        pending[top++] = (int32_t)seed;
        body_of[seed] = count;
        while (top > 0) {
            long cell = pending[--top];
            int row = (int)(cell / across);
            int col = (int)(cell - (long)row * across);
            long nb[4] = {(long)row * across + (col + 1 == across ? 0 : col + 1),
                          (long)row * across + (col == 0 ? across - 1 : col - 1),
                          row > 0 ? cell - across : -1, row + 1 < down ? cell + across : -1};
            for (int k = 0; k < 4; k++) {
                long next = nb[k];
                if (next >= 0 && water[next] && body_of[next] < 0) {
                    body_of[next] = count;
                    pending[top++] = (int32_t)next;
                }
            }
        }
        count++;
    }
    free(pending);
    return count;
}

/* vCycle (OC:740-781). levels[0]'s forcing is the cycle's right-hand side; x is level 0's values,
 * updated in place. */
static void or_vcycle_l(const or_ladder *L, float *x0, double tol, int pole_is_wall, int fixed_coarsest)
{
    int count = L->count;
    or_st st[OR_MAX_LEVELS];
    float *vals[OR_MAX_LEVELS];
    float *forc[OR_MAX_LEVELS];
    st[0] = L->st[0];
    vals[0] = x0;
    forc[0] = NULL;
    double *res = malloc(sizeof(double) * (size_t)L->st[0].across * L->st[0].down);
    double *cbal = malloc(sizeof(double) * (size_t)L->st[0].across * L->st[0].down);
    for (int l = 0; l < count - 1; l++) {
        or_relax_st(&st[l], vals[l], OR_PASSES_EACH_WAY);
        const or_st *coarse = &L->st[l + 1];
        or_residual_st(&st[l], vals[l], res);
        or_restrict(res, st[l].across, st[l].down, coarse->across, coarse->down, cbal);
        long cc = (long)coarse->across * coarse->down;
        forc[l + 1] = malloc(sizeof(float) * (size_t)cc);
        or_with_balance(cc, coarse->water, coarse->c, cbal, forc[l + 1]);
        st[l + 1] = *coarse;
        st[l + 1].f = forc[l + 1];
        vals[l + 1] = calloc((size_t)cc, sizeof(float));
    }
    if (count > 1) {
        const or_st *cst = &st[count - 1];
        if (fixed_coarsest > 0) or_relax_st(cst, vals[count - 1], fixed_coarsest);
        else or_relax_to_convergence_st(cst, vals[count - 1], tol * OR_COARSEST_TOLERANCE_SHARE, res);
    }
    for (int l = count - 2; l >= 0; l--) {
        or_prolong_add(vals[l + 1], st[l + 1].across, st[l + 1].down, vals[l], st[l].across, st[l].down,
                       st[l].water, pole_is_wall);
        or_relax_st(&st[l], vals[l], OR_PASSES_EACH_WAY);
    }
    if (count == 1) or_relax_st(&st[0], vals[0], OR_PASSES_EACH_WAY);
    for (int l = 1; l < count; l++) { free(vals[l]); free(forc[l]); }
    free(res);
    free(cbal);
}

typedef struct {
    int cycles;
    double relative_residual;
} or_solution;

/* solve (OC:473-493) with the processor's relax; with bodies, correctBodyMeans after each cycle. x is start in, answer out. */
static or_solution or_solve_l(const or_ladder *L, float *x, double tol, int pole_is_wall)
{
    const or_st *fin = &L->st[0];
    long cells = (long)fin->across * fin->down;
    double *tmp = malloc(sizeof(double) * (size_t)cells);
    or_balance_of(cells, fin->f, fin->c, tmp);
    double largest = or_largest(cells, tmp);
    int cycles = 0;
    double residual;
    if (largest > 0.0) { or_residual_st(fin, x, tmp); residual = or_largest(cells, tmp) / largest; }
    else residual = 0.0;
    while (residual >= tol && cycles < OR_MOST_CYCLES) {
        or_vcycle_l(L, x, tol, pole_is_wall, 0);
        if (L->body_of) or_correct_body_means(fin, x, L->body_of, L->body_count);
        cycles++;
        or_residual_st(fin, x, tmp);
        residual = or_largest(cells, tmp) / largest;
    }
    free(tmp);
    or_solution s = {cycles, residual};
    return s;
}

/* applyOperator (OC:637-659). */
static void or_apply_operator(const or_st *st, const double *z, double *into)
{
    int across = st->across, down = st->down;
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
    for (int row = 0; row < down; row++) {
        for (int col = 0; col < across; col++) {
            long cell = (long)row * across + col;
            if (!st->water[cell]) { into[cell] = 0.0; continue; }
            int ce = col + 1 == across ? 0 : col + 1;
            int cw = col == 0 ? across - 1 : col - 1;
            double zn = row > 0 ? z[cell - across] : 0.0;
            double zs = row + 1 < down ? z[cell + across] : 0.0;
            double nb = (double)st->e[cell] * z[(long)row * across + ce] + (double)st->w[cell] * z[(long)row * across + cw] +
                        (double)st->n[cell] * zn + (double)st->s[cell] * zs;
            into[cell] = (nb - z[cell]) * st->c[cell];
        }
    }
}

/* dot (OC:661-665): sequential. */
static double or_dot(long n, const double *a, const double *b)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* precondition (OC:610-622): one V-cycle on L z = balance from zero, coarsest 200 passes, then the
 * bodies' means where there are bodies. */
static void or_precondition(const or_ladder *L, const double *balance, double tol, int pole_is_wall,
                            float *zf, double *out)
{
    const or_st *fin = &L->st[0];
    long cells = (long)fin->across * fin->down;
    float *forcing = malloc(sizeof(float) * (size_t)cells);
    or_with_balance(cells, fin->water, fin->c, balance, forcing);
    or_ladder corrected = *L;
    corrected.st[0].f = forcing;
    memset(zf, 0, sizeof(float) * (size_t)cells);
    or_vcycle_l(&corrected, zf, tol, pole_is_wall, OR_PRECONDITIONER_COARSEST_PASSES);
    if (L->body_of) or_correct_body_means(&corrected.st[0], zf, L->body_of, L->body_count);
    for (long i = 0; i < cells; i++) out[i] = (double)zf[i];
    free(forcing);
}

/* solveByKrylov (OC:548-607): right-preconditioned BiCGSTAB in binary64, the ladder's bodies in
 * the preconditioner. */
static or_solution or_solve_by_krylov_l(const or_ladder *L, float *x_io, double tol, int pole_is_wall,
                                        int most_iterations)
{
    const or_st *fin = &L->st[0];
    long cells = (long)fin->across * fin->down;
    double *bal = malloc(sizeof(double) * (size_t)cells);
    or_balance_of(cells, fin->f, fin->c, bal);
    double largest = or_largest(cells, bal);
    double *x = malloc(sizeof(double) * (size_t)cells);
    double *r = malloc(sizeof(double) * (size_t)cells);
    double *shadow = malloc(sizeof(double) * (size_t)cells);
    double *p = calloc((size_t)cells, sizeof(double));
    double *v = calloc((size_t)cells, sizeof(double));
    double *phat = malloc(sizeof(double) * (size_t)cells);
    double *shat = malloc(sizeof(double) * (size_t)cells);
    double *t = malloc(sizeof(double) * (size_t)cells);
    float *zf = malloc(sizeof(float) * (size_t)cells);
    for (long i = 0; i < cells; i++) x[i] = (double)x_io[i];
    or_apply_operator(fin, x, r);
    for (long i = 0; i < cells; i++) r[i] = fin->water[i] ? bal[i] - r[i] : 0.0;
    double residual_now = largest > 0.0 ? or_largest(cells, r) / largest : 0.0;
    memcpy(shadow, r, sizeof(double) * (size_t)cells);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    while (residual_now >= tol && it < most_iterations) {
        double rho_next = or_dot(cells, shadow, r);
        if (rho_next == 0.0 || omega == 0.0) break;
        double beta = (rho_next / rho) * (alpha / omega);
        for (long i = 0; i < cells; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        or_precondition(L, p, tol, pole_is_wall, zf, phat);
        or_apply_operator(fin, phat, v);
        double sv = or_dot(cells, shadow, v);
        if (sv == 0.0) break;
        alpha = rho_next / sv;
        for (long i = 0; i < cells; i++) r[i] -= alpha * v[i];
        for (long i = 0; i < cells; i++) x[i] += alpha * phat[i];
        rho = rho_next;
        it++;
        residual_now = or_largest(cells, r) / largest;
        if (residual_now < tol) break;
        or_precondition(L, r, tol, pole_is_wall, zf, shat);
        or_apply_operator(fin, shat, t);
        double tt = or_dot(cells, t, t);
        omega = tt == 0.0 ? 0.0 : or_dot(cells, t, r) / tt;
        for (long i = 0; i < cells; i++) {
            x[i] += omega * shat[i];
            r[i] -= omega * t[i];
        }
        residual_now = or_largest(cells, r) / largest;
    }
    for (long i = 0; i < cells; i++) x_io[i] = (float)x[i];
    free(bal); free(x); free(r); free(shadow); free(p); free(v); free(phat); free(shat); free(t); free(zf);
    or_solution s = {it, residual_now};
    return s;
}

static int or_all_finite(long n, const float *x)
{
    for (long i = 0; i < n; i++) if (!isfinite(x[i])) return 0;
    return 1;
}

/* solveWithKrylovRescue (OC:507-528). */
static or_solution or_solve_with_rescue_l(const or_ladder *L, float *x, double tol, int pole_is_wall)
{
    long cells = (long)L->st[0].across * L->st[0].down;
    or_solution cycled = or_solve_l(L, x, tol, pole_is_wall);
    if (cycled.relative_residual < tol || !isfinite(cycled.relative_residual) || !or_all_finite(cells, x))
        return cycled;
    or_solution rescued = or_solve_by_krylov_l(L, x, tol, pole_is_wall, OR_MOST_CYCLES);
    int iterations = rescued.cycles, restarts = 0;
    while (rescued.relative_residual >= tol && restarts < OR_KRYLOV_RESTARTS && or_all_finite(cells, x)) {
        rescued = or_solve_by_krylov_l(L, x, tol, pole_is_wall, OR_MOST_CYCLES);
        iterations += rescued.cycles;
        restarts++;
    }
    or_solution s = {cycled.cycles + iterations, rescued.relative_residual};
    return s;
}

/*
 * Flat entry points over a ladder given as arrays of per-level pointers. mode 0 = solve,
 * 1 = solveWithKrylovRescue, 2 = one vCycle (the result's cycles is 1), 3 = solveByKrylov.
 */
int or_solve(int count, const int *acr, const int *dwn, const uint8_t *const *water,
             const float *const *e, const float *const *w, const float *const *n,
             const float *const *s, const float *const *f, const double *const *c, float *x,
             double tol, int pole_is_wall, int mode, double *relative_residual)
{
    if (count < 1 || count > OR_MAX_LEVELS) return -1;
    or_ladder L;
    or_ladder_from(&L, count, acr, dwn, water, e, w, n, s, f, c);
    or_solution sol = {0, 0.0};
    if (mode == 0) sol = or_solve_l(&L, x, tol, pole_is_wall);
    else if (mode == 1) sol = or_solve_with_rescue_l(&L, x, tol, pole_is_wall);
    else if (mode == 2) { or_vcycle_l(&L, x, tol, pole_is_wall, 0); sol.cycles = 1; }
    else if (mode == 3) sol = or_solve_by_krylov_l(&L, x, tol, pole_is_wall, OR_MOST_CYCLES);
    else return -1;
    if (relative_residual) *relative_residual = sol.relative_residual;
    return sol.cycles;
}

/* or_solve with the finest grid's water bodies (body_of may be NULL) and, for solveByKrylov, its
 * iteration limit (mostIterations, OC:548-556). */
int or_solve_bodies(int count, const int *acr, const int *dwn, const uint8_t *const *water,
                    const float *const *e, const float *const *w, const float *const *n,
                    const float *const *s, const float *const *f, const double *const *c, float *x,
                    double tol, int pole_is_wall, int mode, const int32_t *body_of, int body_count,
                    int most_iterations, double *relative_residual)
{
    if (count < 1 || count > OR_MAX_LEVELS) return -1;
    or_ladder L;
    or_ladder_from(&L, count, acr, dwn, water, e, w, n, s, f, c);
    L.body_of = body_of;
    L.body_count = body_count;
    or_solution sol = {0, 0.0};
    if (mode == 0) sol = or_solve_l(&L, x, tol, pole_is_wall);
    else if (mode == 1) sol = or_solve_with_rescue_l(&L, x, tol, pole_is_wall);
    else if (mode == 2) { or_vcycle_l(&L, x, tol, pole_is_wall, 0); sol.cycles = 1; }
    else if (mode == 3) sol = or_solve_by_krylov_l(&L, x, tol, pole_is_wall, most_iterations);
    else return -1;
    if (relative_residual) *relative_residual = sol.relative_residual;
    return sol.cycles;
}

/* ------------------------------------------------------------------------------------------ */
/* The binary64 twin: the same multigrid with every value and transfer in double precision.    */
/* Its fixed point is the exact solution of the float-weight operator with the float forcing.  */
/* ------------------------------------------------------------------------------------------ */

static void or_relax64(const or_st *st, double *x, int passes)
{
    int across = st->across, down = st->down;
    for (int p = 0; p < passes; p++) {
        for (int colour = 0; colour <= 1; colour++) {
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
            for (int row = 0; row < down; row++) {
                for (int col = (colour - row) & 1; col < across; col += 2) {
                    long cell = (long)row * across + col;
                    if (!st->water[cell]) { x[cell] = 0.0; continue; }
                    int ce = col + 1 == across ? 0 : col + 1;
                    int cw = col == 0 ? across - 1 : col - 1;
                    double vn = row > 0 ? x[cell - across] : 0.0;
                    double vs = row + 1 < down ? x[cell + across] : 0.0;
                    x[cell] = (double)st->e[cell] * x[(long)row * across + ce] + (double)st->w[cell] * x[(long)row * across + cw] +
                              (double)st->n[cell] * vn + (double)st->s[cell] * vs - (double)st->f[cell];
                }
            }
        }
    }
}

/* The twin's right-hand side is carried in double (fd), not rounded to float per level. */
static void or_relax64_fd(const or_st *st, const double *fd, double *x, int passes)
{
    int across = st->across, down = st->down;
    for (int p = 0; p < passes; p++) {
        for (int colour = 0; colour <= 1; colour++) {
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
            for (int row = 0; row < down; row++) {
                for (int col = (colour - row) & 1; col < across; col += 2) {
                    long cell = (long)row * across + col;
                    if (!st->water[cell]) { x[cell] = 0.0; continue; }
                    int ce = col + 1 == across ? 0 : col + 1;
                    int cw = col == 0 ? across - 1 : col - 1;
                    double vn = row > 0 ? x[cell - across] : 0.0;
                    double vs = row + 1 < down ? x[cell + across] : 0.0;
                    x[cell] = (double)st->e[cell] * x[(long)row * across + ce] + (double)st->w[cell] * x[(long)row * across + cw] +
                              (double)st->n[cell] * vn + (double)st->s[cell] * vs - fd[cell];
                }
            }
        }
    }
}

/* Residual in the balance's unit with a double right-hand side fd (in the unknown's unit). */
static void or_residual64_fd(const or_st *st, const double *fd, const double *x, double *r)
{
    int across = st->across, down = st->down;
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
    for (int row = 0; row < down; row++) {
        for (int col = 0; col < across; col++) {
            long cell = (long)row * across + col;
            if (!st->water[cell]) { r[cell] = 0.0; continue; }
            int ce = col + 1 == across ? 0 : col + 1;
            int cw = col == 0 ? across - 1 : col - 1;
            double vn = row > 0 ? x[cell - across] : 0.0;
            double vs = row + 1 < down ? x[cell + across] : 0.0;
            double bal = (double)st->e[cell] * x[(long)row * across + ce] + (double)st->w[cell] * x[(long)row * across + cw] +
                         (double)st->n[cell] * vn + (double)st->s[cell] * vs - fd[cell];
            r[cell] = (x[cell] - bal) * st->c[cell];
        }
    }
}

static void or_prolong_add64(const double *coarse, int ca, int cd, double *fine, int fa, int fd,
                             const uint8_t *water, int pole_is_wall)
{
    int cpc = fa / ca, rpc = fd / cd;
    double beyond = pole_is_wall ? -1.0 : 1.0;
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
    for (int row = 0; row < fd; row++) {
        double crow = ((double)row + 0.5) / (double)rpc - 0.5;
        int above = (int)floor(crow);
        double db = crow - (double)above;
        for (int col = 0; col < fa; col++) {
            long cell = (long)row * fa + col;
            if (!water[cell]) continue;
            double ccol = ((double)col + 0.5) / (double)cpc - 0.5;
            int left = (int)floor(ccol);
            double ab = ccol - (double)left;
            int wc = (left + ca) % ca, ec = (left + 1) % ca;
#define OR_AT(rr, cc) ((rr) < 0 ? beyond * coarse[(cc)] : (rr) >= cd ? beyond * coarse[(long)(cd - 1) * ca + (cc)] : coarse[(long)(rr) * ca + (cc)])
            double top = OR_AT(above, wc) * (1.0 - ab) + OR_AT(above, ec) * ab;
            double bottom = OR_AT(above + 1, wc) * (1.0 - ab) + OR_AT(above + 1, ec) * ab;
#undef OR_AT
            fine[cell] += top * (1.0 - db) + bottom * db;
        }
    }
}

static void or_vcycle64(const or_ladder *L, const double *fd0, double *x0, int pole_is_wall,
                        int coarsest_passes)
{
    int count = L->count;
    double *vals[OR_MAX_LEVELS], *fds[OR_MAX_LEVELS];
    long n0 = (long)L->st[0].across * L->st[0].down;
    double *res = malloc(sizeof(double) * (size_t)n0);
    double *cbal = malloc(sizeof(double) * (size_t)n0);
    vals[0] = x0;
    fds[0] = (double *)fd0;
    for (int l = 0; l < count - 1; l++) {
        const or_st *st = &L->st[l];
        or_relax64_fd(st, fds[l], vals[l], OR_PASSES_EACH_WAY);
        const or_st *co = &L->st[l + 1];
        or_residual64_fd(st, fds[l], vals[l], res);
        or_restrict(res, st->across, st->down, co->across, co->down, cbal);
        long cc = (long)co->across * co->down;
        fds[l + 1] = malloc(sizeof(double) * (size_t)cc);
        for (long i = 0; i < cc; i++) fds[l + 1][i] = co->water[i] ? cbal[i] / co->c[i] : 0.0;
        vals[l + 1] = calloc((size_t)cc, sizeof(double));
    }
    if (count > 1) or_relax64_fd(&L->st[count - 1], fds[count - 1], vals[count - 1], coarsest_passes);
    for (int l = count - 2; l >= 0; l--) {
        const or_st *st = &L->st[l];
        or_prolong_add64(vals[l + 1], L->st[l + 1].across, L->st[l + 1].down, vals[l], st->across, st->down,
                         st->water, pole_is_wall);
        or_relax64_fd(st, fds[l], vals[l], OR_PASSES_EACH_WAY);
    }
    if (count == 1) or_relax64_fd(&L->st[0], fds[0], vals[0], OR_PASSES_EACH_WAY);
    for (int l = 1; l < count; l++) { free(vals[l]); free(fds[l]); }
    free(res);
    free(cbal);
}

/*
 * The binary64 twin of solve: x (double) from its start until max|F - Lx| / max|F| < tol, with
 * F = f * c and the float forcing f of level 0. V(2,2) cycles; coarsest relaxed coarsest_passes
 * each cycle. If tol_mark > 0, *seconds_at_mark receives nothing (timing is the caller's); the
 * residual history (one entry per cycle, up to max_cycles) is written to history if given.
 * Returns the cycles run.
 */
int or_solve64(int count, const int *acr, const int *dwn, const uint8_t *const *water,
               const float *const *e, const float *const *w, const float *const *n,
               const float *const *s, const float *const *f, const double *const *c, double *x,
               double tol, int pole_is_wall, int max_cycles, int coarsest_passes, double *history,
               double *relative_residual)
{
    if (count < 1 || count > OR_MAX_LEVELS) return -1;
    or_ladder L;
    or_ladder_from(&L, count, acr, dwn, water, e, w, n, s, f, c);
    const or_st *fin = &L.st[0];
    long cells = (long)fin->across * fin->down;
    double *fd = malloc(sizeof(double) * (size_t)cells);
    double *tmp = malloc(sizeof(double) * (size_t)cells);
    for (long i = 0; i < cells; i++) fd[i] = fin->water[i] ? (double)fin->f[i] : 0.0;
    or_balance_of(cells, fin->f, fin->c, tmp);
    double largest = or_largest(cells, tmp);
    int cycles = 0;
    double residual = 0.0;
    if (largest > 0.0) { or_residual64_fd(fin, fd, x, tmp); residual = or_largest(cells, tmp) / largest; }
    while (residual >= tol && cycles < max_cycles) {
        or_vcycle64(&L, fd, x, pole_is_wall, coarsest_passes);
        cycles++;
        or_residual64_fd(fin, fd, x, tmp);
        residual = or_largest(cells, tmp) / largest;
        if (history) history[cycles - 1] = residual;
    }
    if (relative_residual) *relative_residual = residual;
    free(fd);
    free(tmp);
    return cycles;
}

/* Binary64 relax and residual of a double vector on a float-forcing stencil, exported for tests. */
void or_relax64_flat(int across, int down, const uint8_t *water, const float *e, const float *w,
                     const float *n, const float *s, const float *f, double *x, int passes)
{
    or_st st = {across, down, water, e, w, n, s, f, NULL};
    or_relax64(&st, x, passes);
}

void or_residual64(int across, int down, const uint8_t *water, const float *e, const float *w,
                   const float *n, const float *s, const float *f, const double *c, const double *x,
                   double *r)
{
    long cells = (long)across * down;
    double *fd = malloc(sizeof(double) * (size_t)cells);
    for (long i = 0; i < cells; i++) fd[i] = water[i] ? (double)f[i] : 0.0;
    or_st st = {across, down, water, e, w, n, s, f, c};
    or_residual64_fd(&st, fd, x, r);
    free(fd);
}

/* velocities (OC:801-828). */
void or_velocities(int across, int down, const uint8_t *water, const float *stream, double dx,
                   double dy, float *eastward, float *northward)
{
    for (int row = 0; row < down; row++) {
        for (int col = 0; col < across; col++) {
            long cell = (long)row * across + col;
            if (!water[cell]) { eastward[cell] = 0.0f; northward[cell] = 0.0f; continue; }
            int ce = col + 1 == across ? 0 : col + 1;
            int cw = col == 0 ? across - 1 : col - 1;
            float sn = row > 0 ? stream[cell - across] : -stream[cell];
            float ss = row + 1 < down ? stream[cell + across] : -stream[cell];
            eastward[cell] = (float)(-(double)(sn - ss) / (2.0 * dy));
            northward[cell] = (float)((double)(stream[(long)row * across + ce] - stream[(long)row * across + cw]) / (2.0 * dx));
        }
    }
}

/* OceanStage.waterOn (OS:551-573): a solve cell is water when every map cell it overlaps is water.
 * The overlap of solve row `row` is map rows floor(row * rpc + 1e-6) .. ceil((row + 1) * rpc - 1e-6) - 1,
 * capped at the last map row (TOUCHING_SHARE, OS:129), likewise for columns; no wrap. */
void or_water_on(const uint8_t *map_land, int map_across, int map_down, int across, int down, uint8_t *water)
{
    const double touching = 1e-6;
    const double cols_per = (double)map_across / across;
    const double rows_per = (double)map_down / down;
    for (int row = 0; row < down; row++) {
        int first_row = (int)floor(row * rows_per + touching);
        int last_row = (int)ceil((row + 1) * rows_per - touching) - 1;
        if (last_row > map_down - 1) last_row = map_down - 1;
        for (int col = 0; col < across; col++) {
            int first_col = (int)floor(col * cols_per + touching);
            int last_col = (int)ceil((col + 1) * cols_per - touching) - 1;
            if (last_col > map_across - 1) last_col = map_across - 1;
            int all_water = 1;
            for (int mr = first_row; mr <= last_row; mr++)
                for (int mc = first_col; mc <= last_col; mc++)
                    if (map_land[(long)mr * map_across + mc]) all_water = 0;
            water[(long)row * across + col] = (uint8_t)all_water;
        }
    }
}

/* OceanStage.curlForcing (OS:689-708): F = (d tau_n/dx - d tau_e/dy) / (rho H) per cell, binary64,
 * central differences in meters, columns wrapping, the edge rows one-sided, y running north (the row
 * above is the northern one). rho = 1025 kg/m^3 (OS:64); H the wind-driven layer's 500 m (OS:74)
 * when depth is NULL, else depth[cell] (the per-cell reading of the synthetic workload). */
void or_curl_forcing(int across, int down, double width_m, double height_m, const double *eastward,
                     const double *northward, const double *depth, double *forcing)
{
    const double rho = 1025.0, layer = 500.0;
    const double per_density_depth = 1.0 / (rho * layer);
    for (int row = 0; row < down; row++) {
        int rn = row - 1 < 0 ? 0 : row - 1;
        int rs = row + 1 > down - 1 ? down - 1 : row + 1;
        double ns_m = (rs - rn) * height_m;
        for (int col = 0; col < across; col++) {
            long cell = (long)row * across + col;
            int ce = col + 1 == across ? 0 : col + 1;
            int cw = col == 0 ? across - 1 : col - 1;
            double dtn_dx = (northward[(long)row * across + ce] - northward[(long)row * across + cw]) / (2.0 * width_m);
            double dte_dy = (eastward[(long)rn * across + col] - eastward[(long)rs * across + col]) / ns_m;
            double scale = depth ? 1.0 / (rho * depth[cell]) : per_density_depth;
            forcing[cell] = (dtn_dx - dte_dy) * scale;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* The heat problem (OceanHeat.kt, "OH")                                                      */
/* ------------------------------------------------------------------------------------------ */

#define OR_ROTATION_RATE ((double)7.2921159e-5f) /* WorldScale.ROTATION_RATE_PER_S, a Float (WGC:302) */
#define OR_EARTH_MEAN_RADIUS_M 6.371e6            /* OS:80 */

/* BAROCLINIC_WAVE_SPEED_M_PER_S (OH:86-88): 2 beta_eq R^2 with Earth's equatorial beta, R = 240 km. */
static double or_wave_speed(void)
{
    return 2.0 * (2.0 * OR_ROTATION_RATE / OR_EARTH_MEAN_RADIUS_M) * 240000.0 * 240000.0;
}

/* deformationRadiusMeters (OH:98-105). */
double or_deformation_radius(double latitude_deg, double radius_m)
{
    double latitude = latitude_deg * M_PI / 180.0;
    double coriolis = fabs(2.0 * OR_ROTATION_RATE * sin(latitude));
    double beta = 2.0 * OR_ROTATION_RATE * cos(latitude) / radius_m;
    double c = or_wave_speed();
    double off_equator = coriolis == 0.0 ? INFINITY : c / coriolis;
    double equatorial = beta <= 0.0 ? INFINITY : sqrt(c / (2.0 * beta));
    return off_equator < equatorial ? off_equator : equatorial;
}

/* diffusivity (OH:128-129) with EDDY_VELOCITY_MPS (OH:112-114): K = V R_d, V fixed so that K is
 * 2,500 m^2/s at 45 degrees. */
double or_heat_diffusivity(double latitude_deg, double radius_m)
{
    double v = 2500.0 * 2.0 * OR_ROTATION_RATE * sin(45.0 * M_PI / 180.0) / or_wave_speed();
    return v * or_deformation_radius(latitude_deg, radius_m);
}

/* bernoulli (OH:240-244): x / (e^x - 1), its series 1 - x/2 below 1e-3, zero above 700. */
static double or_bernoulli(double x)
{
    if (fabs(x) < 1e-3) return 1.0 - x / 2.0;
    if (x > 700.0) return 0.0;
    return x / expm1(x);
}

/* faceWeight (OH:231-234). */
static double or_face_weight(double k, double outward, double spacing)
{
    double peclet = outward * spacing / k;
    return k / (spacing * spacing) * or_bernoulli(peclet);
}

/* corner (OH:262-281): psi at the corner between two columns on the north or south side of row:
 * zero when any of the four cells around it is land (beyond a pole counts as not land), otherwise
 * their mean, a cell beyond a pole taking minus the same column's value in row. */
static double or_corner(const float *stream, const uint8_t *water, int across, int down, int row, int wc, int ec,
                        int north)
{
    int other = north ? row - 1 : row + 1;
#define OR_ON_LAND(r, c) ((r) >= 0 && (r) < down && !water[(long)(r) * across + (c)])
    if (OR_ON_LAND(row, wc) || OR_ON_LAND(row, ec) || OR_ON_LAND(other, wc) || OR_ON_LAND(other, ec)) return 0.0;
#undef OR_ON_LAND
#define OR_AT(r, c) (((r) < 0 || (r) >= down) ? -(double)stream[(long)row * across + (c)] : (double)stream[(long)(r) * across + (c)])
    return (OR_AT(row, wc) + OR_AT(row, ec) + OR_AT(other, wc) + OR_AT(other, ec)) / 4.0;
#undef OR_AT
}

/* OceanHeat.stencil then withBalance (OH:151-228). The diffusivity is `constant_k` where that is
 * positive, else diffusivity(latitude, radius_m). entrainment and subsurface may be NULL. Land keeps
 * zero weights, centre 1 and forcing 0. */
void or_heat_stencil(int across, int down, double dx, double dy, const uint8_t *water, const float *stream,
                     const float *target, double relaxation_s, int with_target, double radius_m, double constant_k,
                     const float *entrainment, const float *subsurface, float *e, float *w, float *n, float *s,
                     double *c, float *f)
{
    long cells = (long)across * down;
    double rate = 1.0 / relaxation_s;
    double *along = malloc(sizeof(double) * (size_t)down);
    double *boundary = malloc(sizeof(double) * (size_t)(down + 1));
    for (int row = 0; row < down; row++) {
        float lat = 90.0f - 180.0f * ((float)row + 0.5f) / (float)down; /* ClimateStage.latitudeOf */
        along[row] = constant_k > 0.0 ? constant_k : or_heat_diffusivity((double)lat, radius_m);
    }
    for (int b = 0; b <= down; b++)
        boundary[b] = constant_k > 0.0 ? constant_k : or_heat_diffusivity(90.0 - 180.0 * b / down, radius_m);
    for (long i = 0; i < cells; i++) { e[i] = w[i] = n[i] = s[i] = 0.0f; c[i] = 1.0; f[i] = 0.0f; }
    for (int row = 0; row < down; row++) {
        for (int col = 0; col < across; col++) {
            long cell = (long)row * across + col;
            if (!water[cell]) continue;
            int ce = col + 1 == across ? 0 : col + 1;
            int cw = col == 0 ? across - 1 : col - 1;
            double ne = or_corner(stream, water, across, down, row, col, ce, 1);
            double se = or_corner(stream, water, across, down, row, col, ce, 0);
            double nw = or_corner(stream, water, across, down, row, cw, col, 1);
            double sw = or_corner(stream, water, across, down, row, cw, col, 0);
            double out_e = -(ne - se) / dy;
            double out_w = (nw - sw) / dy;
            double out_n = (ne - nw) / dx;
            double out_s = -(se - sw) / dx;
            double we = water[(long)row * across + ce] ? or_face_weight(along[row], out_e, dx) : 0.0;
            double ww = water[(long)row * across + cw] ? or_face_weight(along[row], out_w, dx) : 0.0;
            double wn = (row > 0 && water[cell - across]) ? or_face_weight(boundary[row], out_n, dy) : 0.0;
            double ws = (row + 1 < down && water[cell + across]) ? or_face_weight(boundary[row + 1], out_s, dy) : 0.0;
            double ent = entrainment ? (double)entrainment[cell] : 0.0;
            double centre = we + ww + wn + ws + rate + ent;
            e[cell] = (float)(we / centre);
            w[cell] = (float)(ww / centre);
            n[cell] = (float)(wn / centre);
            s[cell] = (float)(ws / centre);
            c[cell] = centre;
            if (with_target) {
                double balance = (double)(-target[cell]) * rate - ent * (subsurface ? (double)subsurface[cell] : 0.0);
                f[cell] = (float)(balance / centre);
            }
        }
    }
    free(along);
    free(boundary);
}

/* resample and sample (OS:1199-1223): bilinear at each target cell centre, columns wrapping, rows
 * clamped, all in binary32. */
void or_resample(const float *field, int fa, int fd, int ta, int td, float *out)
{
    for (int row = 0; row < td; row++) {
        float from_row = ((float)row + 0.5f) * (float)fd / (float)td - 0.5f;
        for (int col = 0; col < ta; col++) {
            float from_col = ((float)col + 0.5f) * (float)fa / (float)ta - 0.5f;
            int left = (int)floorf(from_col);
            float ab = from_col - (float)left;
            int wc = ((left % fa) + fa) % fa;
            int ec = wc + 1 == fa ? 0 : wc + 1;
            float cr = from_row < 0.0f ? 0.0f : (from_row > (float)(fd - 1) ? (float)(fd - 1) : from_row);
            int above = (int)cr < fd - 1 ? (int)cr : fd - 1;
            int below = above + 1 < fd - 1 ? above + 1 : fd - 1;
            float db = cr - (float)above;
            float top = field[(long)above * fa + wc] * (1.0f - ab) + field[(long)above * fa + ec] * ab;
            float bottom = field[(long)below * fa + wc] * (1.0f - ab) + field[(long)below * fa + ec] * ab;
            out[(long)row * ta + col] = top * (1.0f - db) + bottom * db;
        }
    }
}

/* coarseMean (OS:857-860): restrict of the binary64 promotion, rounded to binary32. */
void or_coarse_mean(const float *fine, int fa, int fd, int ca, int cd, float *coarse)
{
    long nf = (long)fa * fd, nc = (long)ca * cd;
    double *f64 = malloc(sizeof(double) * (size_t)nf);
    double *c64 = malloc(sizeof(double) * (size_t)nc);
    for (long i = 0; i < nf; i++) f64[i] = (double)fine[i];
    or_restrict(f64, fa, fd, ca, cd, c64);
    for (long i = 0; i < nc; i++) coarse[i] = (float)c64[i];
    free(f64);
    free(c64);
}
