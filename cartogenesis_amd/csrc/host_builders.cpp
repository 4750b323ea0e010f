// Host-side problem builders of the ocean solver: the circulation and heat stencils, coarse masks,
// the level ladder's shapes, beta per row, the bottom drag, water bodies, the heat's coarse-grid
// resampling, and the synthetic workload generator.
//
// Compiled with -ffp-contract=off so every double and float operation rounds once, in the order
// written, as the reference's JVM arithmetic does (the reference is Kotlin; JVM floating point is
// strict IEEE with no fused multiply-add).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "cgocean.h"

namespace {

// WorldScale.ROTATION_RATE_PER_S (WGC:302): a Float constant, promoted to double where used.
constexpr float kRotationRateF = 7.2921159e-5f;
// OceanStage constants (OS:64, 80, 83, 91).
constexpr double kSeawaterDensity = 1025.0;
constexpr double kEarthMeanRadius = 6.371e6;
constexpr double kGulfStreamLatitudeDeg = 30.0;
constexpr double kGulfStreamWidth = 50000.0;
// OceanCirculation constants (OC:88, 104, 122).
constexpr double kSeriesBelowCellPeclet = 1e-4;
constexpr int kCoarsestRows = 8;
constexpr double kSquareEnough = 1.25;

// splitmix64 (Steele, Lea and Flood 2014): the generator's only source of randomness.
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// A uniform double in [0, 1) from 53 bits of the hash of (seed, stream, index).
inline double uniform01(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + index);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

// One octave of value noise at the centre of cell (row, col): a lattice of nx by ny cells over the
// whole map, periodic in x, with a seeded value at each lattice point, blended by smoothstep.
double octave(uint64_t seed, int o, int nx, int ny, int across, int down, int row, int col) {
    double u = ((double)col + 0.5) / (double)across * nx;
    double v = ((double)row + 0.5) / (double)down * ny;
    int i0 = (int)std::floor(u), j0 = (int)std::floor(v);
    double fu = smooth(u - i0), fv = smooth(v - j0);
    int i1 = (i0 + 1) % nx;
    int j1 = std::min(j0 + 1, ny);
    uint64_t stream = 100 + (uint64_t)o;
    auto at = [&](int i, int j) { return uniform01(seed, stream, (uint64_t)j * (uint64_t)(nx + 1) + (uint64_t)i); };
    double top = at(i0, j0) * (1.0 - fu) + at(i1, j0) * fu;
    double bottom = at(i0, j1) * (1.0 - fu) + at(i1, j1) * fu;
    return top * (1.0 - fv) + bottom * fv;
}

// OceanHeat's deformation radius and eddy diffusivity (OceanHeat.kt:86-129).
struct EddyModel {
    double wave_speed;    // BAROCLINIC_WAVE_SPEED_M_PER_S: 2 beta_eq R^2, beta_eq = 2 Omega / a_Earth, R = 240 km
    double eddy_velocity; // EDDY_VELOCITY_MPS: K = 2,500 m^2/s at 45 degrees on any planet
    EddyModel() {
        const double omega = (double)kRotationRateF;
        wave_speed = 2.0 * (2.0 * omega / kEarthMeanRadius) * 240000.0 * 240000.0;
        eddy_velocity = 2500.0 * 2.0 * omega * std::sin(45.0 * M_PI / 180.0) / wave_speed;
    }
    double radius(double latitude_deg, double planet_radius_m) const {
        const double omega = (double)kRotationRateF;
        const double phi = latitude_deg * M_PI / 180.0;
        const double f = std::fabs(2.0 * omega * std::sin(phi));
        const double beta = 2.0 * omega * std::cos(phi) / planet_radius_m;
        const double away = f == 0.0 ? INFINITY : wave_speed / f;
        const double near = beta <= 0.0 ? INFINITY : std::sqrt(wave_speed / (2.0 * beta));
        return std::min(away, near);
    }
    double diffusivity(double latitude_deg, double planet_radius_m) const {
        return eddy_velocity * radius(latitude_deg, planet_radius_m);
    }
};

// The Scharfetter-Gummel face weight K B(P) / d^2, B(x) = x / (e^x - 1), P = w d / K (OH:231-244).
inline double sg_face(double k, double outward, double spacing) {
    const double p = outward * spacing / k;
    double b;
    if (std::fabs(p) < 1e-3) b = 1.0 - p / 2.0;
    else if (p > 700.0) b = 0.0;
    else b = p / std::expm1(p);
    return k / (spacing * spacing) * b;
}

}  // namespace

extern "C" {

// OceanStage.BOTTOM_DRAG_PER_S (OS:102-103): 2 Omega cos(30 deg) / a_Earth * 50 km, with Omega the
// Float rotation rate promoted to double, evaluated left to right.
double cg_bottom_drag(void) {
    return 2.0 * (double)kRotationRateF * std::cos(kGulfStreamLatitudeDeg * M_PI / 180.0) / kEarthMeanRadius *
           kGulfStreamWidth;
}

// WorldScale.planetaryVorticityGradientPerMeterSecond (WGC:272-273) at ClimateStage.latitudeOf's
// binary32 row latitude (ClimateStage.kt:1313-1314), with radiusMeters = worldWidthKm * 1000f / 2 pi
// (WGC:246).
int cg_beta_rows(int down, double world_width_km, double *beta) {
    if (down <= 0 || !beta) return CG_ERR_ARG;
    double radius = world_width_km * (double)1000.0f / (2.0 * M_PI);
    for (int row = 0; row < down; row++) {
        float lat = 90.0f - 180.0f * ((float)row + 0.5f) / (float)down;
        beta[row] = 2.0 * (double)kRotationRateF * std::cos((double)lat * M_PI / 180.0) / radius;
    }
    return CG_OK;
}

// OceanCirculation.stencil (OC:161-219) and withBalance (OC:222-231).
int cg_circulation_stencil(int across, int down, double dx, double dy, const uint8_t *water,
                           const double *beta, double drag, const double *balance, float *east,
                           float *west, float *north, float *south, double *centre, float *forcing) {
    if (across <= 0 || down <= 0 || !water || !beta || !east || !west || !north || !south || !centre || !forcing)
        return CG_ERR_ARG;
    const double along_y = drag / (dy * dy);
    const double width_sq = dx * dx;
    for (int row = 0; row < down; row++) {
        const double b = beta[row];
        const double peclet = std::fabs(b) * dx / (2.0 * drag);
        double ex, wx;
        if (peclet < kSeriesBelowCellPeclet) {
            const double fitted = drag * (1.0 + peclet * peclet / 3.0) / width_sq;
            const double upstream = std::fabs(b) / (2.0 * dx);
            ex = fitted + upstream;
            wx = fitted - upstream;
        } else {
            // r p (coth p +- 1) / dx^2 with coth p - 1 = 2 / expm1(2p).
            const double scale = drag * peclet / width_sq;
            const double coth_less_one = 2.0 / std::expm1(2.0 * peclet);
            ex = scale * (2.0 + coth_less_one);
            wx = scale * coth_less_one;
        }
        // Pole walls fold into the edge rows' centre weight.
        const double ny = row == 0 ? 0.0 : along_y;
        const double sy = row == down - 1 ? 0.0 : along_y;
        const double c = ex + wx + 2.0 * along_y + (along_y - ny) + (along_y - sy);
        const float fe = (float)(ex / c), fw = (float)(wx / c), fn = (float)(ny / c), fs = (float)(sy / c);
        for (long cell = (long)row * across; cell < (long)(row + 1) * across; cell++) {
            east[cell] = fe;
            west[cell] = fw;
            north[cell] = fn;
            south[cell] = fs;
            centre[cell] = c;
        }
    }
    const long cells = (long)across * down;
    for (long cell = 0; cell < cells; cell++)
        forcing[cell] = (water[cell] && balance) ? (float)(balance[cell] / centre[cell]) : 0.0f;
    return CG_OK;
}

// OceanCirculation.coarseWater (OC:404-423).
int cg_coarse_water(const uint8_t *fine, int fa, int fd, int ca, int cd, int every_cell_water, uint8_t *coarse) {
    if (!fine || !coarse || fa <= 0 || fd <= 0 || ca <= 0 || cd <= 0 || fa % ca || fd % cd) return CG_ERR_ARG;
    const int cpc = fa / ca, rpc = fd / cd;
    const uint8_t ev = every_cell_water ? 1 : 0;
    std::memset(coarse, ev, (size_t)ca * cd);
    for (int row = 0; row < fd; row++)
        for (int col = 0; col < fa; col++)
            if ((fine[(long)row * fa + col] ? 1 : 0) != ev) coarse[(long)(row / rpc) * ca + col / cpc] = (uint8_t)!ev;
    return CG_OK;
}

// The shapes OceanCirculation.levels walks (OC:434-462).
int cg_level_shapes(int across, int down, double dx, double dy, int max_levels, int *la, int *ld, double *cw,
                    double *ch) {
    if (across <= 0 || down <= 0 || max_levels < 1 || !la || !ld) return CG_ERR_ARG;
    int count = 1;
    la[0] = across;
    ld[0] = down;
    if (cw) cw[0] = dx;
    if (ch) ch[0] = dy;
    double w = dx, h = dy;
    while (count < max_levels) {
        const bool wide = w > h * kSquareEnough;
        const bool tall = h > w * kSquareEnough;
        const int na = wide ? across : across / 2;
        const int nd = tall ? down : down / 2;
        const bool even = (na == across || across % 2 == 0) && (nd == down || down % 2 == 0);
        if (!even || nd < kCoarsestRows || na % 2 != 0) break;
        w *= (double)across / na;
        h *= (double)down / nd;
        across = na;
        down = nd;
        la[count] = across;
        ld[count] = down;
        if (cw) cw[count] = w;
        if (ch) ch[count] = h;
        count++;
    }
    return count;
}

// The synthetic workload's zonal stress: tau_x = amplitude cos(3 (lat - shift)) + N(0, 0.005^2) per
// cell (Box-Muller from two uniforms of the cell), lat the row's binary32 centre latitude.
static void zonal_stress(int across, int down, uint64_t seed, int month, double *tau) {
    double amplitude = 0.1, shift_deg = 0.0;
    if (month >= 0) {
        amplitude = 0.1 * std::cos(2.0 * M_PI * month / 12.0);
        shift_deg = 10.0 * std::sin(2.0 * M_PI * month / 12.0);
    }
    for (int row = 0; row < down; row++) {
        const float lat = 90.0f - 180.0f * ((float)row + 0.5f) / (float)down;
        const double band = amplitude * std::cos(3.0 * ((double)lat - shift_deg) * M_PI / 180.0);
        for (int col = 0; col < across; col++) {
            const long i = (long)row * across + col;
            const double u1 = 1.0 - uniform01(seed, 2, (uint64_t)i);
            const double u2 = uniform01(seed, 3, (uint64_t)i);
            const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            tau[i] = band + 0.005 * g;
        }
    }
}

// The synthetic workload. The curl follows OceanStage.curlForcing (OS:689-708): central
// differences, columns wrapping, edge rows one-sided, with per-cell depth H in rho H.
int cg_problem_generate(int across, int down, uint64_t seed, double land_fraction, int month, uint8_t *water,
                        double *depth, double *balance) {
    if (across <= 0 || down <= 0 || !(land_fraction >= 0.0 && land_fraction < 1.0) || month > 11) return CG_ERR_ARG;
    const long cells = (long)across * down;
    // Land mask: 4 octaves of value noise, lattices 8x4, 16x8, 32x16, 64x32, amplitude halving.
    std::vector<double> noise(cells);
    for (int row = 0; row < down; row++)
        for (int col = 0; col < across; col++) {
            double total = 0.0, amp = 1.0;
            for (int o = 0; o < 4; o++) {
                total += amp * octave(seed, o, 8 << o, 4 << o, across, down, row, col);
                amp *= 0.5;
            }
            noise[(long)row * across + col] = total;
        }
    // The lowest land_fraction of the noise is land: exactly round(land_fraction * cells) cells.
    const long land_cells = std::lround(land_fraction * (double)cells);
    std::vector<uint8_t> wet(cells, 1);
    if (land_cells > 0) {
        std::vector<double> sorted(noise);
        std::nth_element(sorted.begin(), sorted.begin() + (land_cells - 1), sorted.end());
        const double threshold = sorted[land_cells - 1];
        for (long i = 0; i < cells; i++) wet[i] = noise[i] > threshold ? 1 : 0;
    }
    // Depth: i.i.d. uniform 200..6000 m on water, 0 on land.
    std::vector<double> h(cells, 0.0);
    for (long i = 0; i < cells; i++)
        if (wet[i]) h[i] = 200.0 + 5800.0 * uniform01(seed, 1, (uint64_t)i);
    std::vector<double> tau(cells);
    zonal_stress(across, down, seed, month, tau.data());
    const double world_width_m = 40075.0 * 1000.0;
    const double height_m = world_width_m * 0.5 / down;
    if (balance) {
        for (int row = 0; row < down; row++) {
            const int rn = std::max(row - 1, 0), rs = std::min(row + 1, down - 1);
            const double ns_m = (double)(rs - rn) * height_m;
            for (int col = 0; col < across; col++) {
                const long i = (long)row * across + col;
                if (!wet[i]) { balance[i] = 0.0; continue; }
                // tau_y = 0, so d(tau_y)/dx is (0 - 0) / 2dx = 0.
                const double dtn_dx = 0.0;
                const double dte_dy = (tau[(long)rn * across + col] - tau[(long)rs * across + col]) / ns_m;
                balance[i] = (dtn_dx - dte_dy) * (1.0 / (kSeawaterDensity * h[i]));
            }
        }
    }
    if (water) std::memcpy(water, wet.data(), (size_t)cells);
    if (depth) std::memcpy(depth, h.data(), sizeof(double) * (size_t)cells);
    return CG_OK;
}

// OceanHeat.diffusivity (OH:128-129).
double cg_heat_diffusivity(double latitude_deg, double planet_radius_m) {
    static const EddyModel model;
    return model.diffusivity(latitude_deg, planet_radius_m);
}

// OceanHeat.stencil then withBalance (OH:151-228). Face velocities come from psi at the cell's
// corners; a corner touching land is the coast's psi, zero, else the mean of its four cells, a cell
// beyond a pole reading minus its own column's value in the cell's row.
int cg_heat_stencil(int across, int down, double dx, double dy, const uint8_t *water, const float *stream,
                    const float *target, double relaxation_s, int with_target, double planet_radius_m,
                    double constant_k, const float *entrainment, const float *subsurface, float *east, float *west,
                    float *north, float *south, double *centre, float *forcing) {
    if (across <= 0 || down <= 0 || !water || !stream || (with_target && !target) || !(relaxation_s > 0.0) ||
        !east || !west || !north || !south || !centre || !forcing || (constant_k <= 0.0 && !(planet_radius_m > 0.0)))
        return CG_ERR_ARG;
    static const EddyModel model;
    auto k_at = [&](double lat) { return constant_k > 0.0 ? constant_k : model.diffusivity(lat, planet_radius_m); };
    std::vector<double> along(down), across_edge(down + 1);
    for (int row = 0; row < down; ++row) {
        const float lat = 90.0f - 180.0f * ((float)row + 0.5f) / (float)down;  // ClimateStage.latitudeOf
        along[row] = k_at((double)lat);
    }
    for (int b = 0; b <= down; ++b) across_edge[b] = k_at(90.0 - 180.0 * b / down);
    const double rate = 1.0 / relaxation_s;
    auto wet = [&](int r, int c) { return water[(long)r * across + c] != 0; };
    auto corner = [&](int row, int wc, int ec, int other) -> double {
        for (int r : {row, other})
            if (r >= 0 && r < down && (!wet(r, wc) || !wet(r, ec))) return 0.0;
        auto psi = [&](int r, int c) -> double {
            return (r < 0 || r >= down) ? -(double)stream[(long)row * across + c] : (double)stream[(long)r * across + c];
        };
        return (psi(row, wc) + psi(row, ec) + psi(other, wc) + psi(other, ec)) / 4.0;
    };
    for (int row = 0; row < down; ++row) {
        for (int col = 0; col < across; ++col) {
            const long cell = (long)row * across + col;
            east[cell] = west[cell] = north[cell] = south[cell] = 0.0f;
            centre[cell] = 1.0;
            forcing[cell] = 0.0f;
            if (!water[cell]) continue;
            const int ce = col + 1 == across ? 0 : col + 1;
            const int cw = col == 0 ? across - 1 : col - 1;
            const double ne = corner(row, col, ce, row - 1), se = corner(row, col, ce, row + 1);
            const double nw = corner(row, cw, col, row - 1), sw = corner(row, cw, col, row + 1);
            const double ae = wet(row, ce) ? sg_face(along[row], -(ne - se) / dy, dx) : 0.0;
            const double aw = wet(row, cw) ? sg_face(along[row], (nw - sw) / dy, dx) : 0.0;
            const double an = (row > 0 && wet(row - 1, col)) ? sg_face(across_edge[row], (ne - nw) / dx, dy) : 0.0;
            const double as = (row + 1 < down && wet(row + 1, col)) ? sg_face(across_edge[row + 1], -(se - sw) / dx, dy)
                                                                    : 0.0;
            const double ent = entrainment ? (double)entrainment[cell] : 0.0;
            const double c = ae + aw + an + as + rate + ent;
            east[cell] = (float)(ae / c);
            west[cell] = (float)(aw / c);
            north[cell] = (float)(an / c);
            south[cell] = (float)(as / c);
            centre[cell] = c;
            if (with_target) {
                const double balance = (double)(-target[cell]) * rate - ent * (subsurface ? (double)subsurface[cell] : 0.0);
                forcing[cell] = (float)(balance / c);
            }
        }
    }
    return CG_OK;
}

// OceanCirculation.waterBodies (OC:670-697): four-connected, columns wrapping, numbered in the order
// of each body's first cell (a flood fill from each unlabelled water cell in cell order).
int cg_water_bodies(const uint8_t *water, int across, int down, int32_t *body_of) {
    if (!water || !body_of || across <= 0 || down <= 0) return CG_ERR_ARG;
    const long cells = (long)across * down;
    std::fill(body_of, body_of + cells, -1);
    std::vector<int32_t> stack;
    int count = 0;
    for (long seed = 0; seed < cells; ++seed) {
        if (!water[seed] || body_of[seed] >= 0) continue;
        body_of[seed] = count;
        stack.assign(1, (int32_t)seed);
        while (!stack.empty()) {
            const long cell = stack.back();
            stack.pop_back();
            const int row = (int)(cell / across), col = (int)(cell % across);
            const long nbs[4] = {(long)row * across + (col + 1) % across, (long)row * across + (col + across - 1) % across,
                                 row > 0 ? cell - across : -1, row + 1 < down ? cell + across : -1};
            for (long nb : nbs)
                if (nb >= 0 && water[nb] && body_of[nb] < 0) {
                    body_of[nb] = count;
                    stack.push_back((int32_t)nb);
                }
        }
        ++count;
    }
    return count;
}

// OceanStage.resample / sample (OS:1199-1223): bilinear in binary32 at each target cell centre,
// columns wrapping, rows clamped.
int cg_resample(const float *field, int fa, int fd, int ta, int td, float *out) {
    if (!field || !out || fa <= 0 || fd <= 0 || ta <= 0 || td <= 0) return CG_ERR_ARG;
    for (int row = 0; row < td; ++row) {
        const float fr = ((float)row + 0.5f) * (float)fd / (float)td - 0.5f;
        const float clamped = std::min(std::max(fr, 0.0f), (float)(fd - 1));
        const int above = std::min((int)clamped, fd - 1);
        const int below = std::min(above + 1, fd - 1);
        const float db = clamped - (float)above;
        for (int col = 0; col < ta; ++col) {
            const float fc = ((float)col + 0.5f) * (float)fa / (float)ta - 0.5f;
            const int left = (int)std::floor(fc);
            const float ab = fc - (float)left;
            const int wc = ((left % fa) + fa) % fa;
            const int ec = wc + 1 == fa ? 0 : wc + 1;
            const float top = field[(long)above * fa + wc] * (1.0f - ab) + field[(long)above * fa + ec] * ab;
            const float bottom = field[(long)below * fa + wc] * (1.0f - ab) + field[(long)below * fa + ec] * ab;
            out[(long)row * ta + col] = top * (1.0f - db) + bottom * db;
        }
    }
    return CG_OK;
}

// OceanStage.coarseMean (OS:857-860): restrict (OC:334-346) of the binary64 promotion, each coarse
// cell accumulating fine * (1 / (cx cy)) in row-major fine order, then rounded to binary32.
int cg_coarse_mean(const float *fine, int fa, int fd, int ca, int cd, float *coarse) {
    if (!fine || !coarse || fa <= 0 || fd <= 0 || ca <= 0 || cd <= 0 || fa % ca || fd % cd) return CG_ERR_ARG;
    const int cpc = fa / ca, rpc = fd / cd;
    const double share = 1.0 / (double)(cpc * rpc);
    std::vector<double> acc((size_t)ca * cd, 0.0);
    for (int row = 0; row < fd; ++row)
        for (int col = 0; col < fa; ++col)
            acc[(long)(row / rpc) * ca + col / cpc] += (double)fine[(long)row * fa + col] * share;
    for (long i = 0; i < (long)ca * cd; ++i) coarse[i] = (float)acc[i];
    return CG_OK;
}

int cg_problem_stress(int across, int down, uint64_t seed, int month, double *eastward, double *northward) {
    if (across <= 0 || down <= 0 || month > 11) return CG_ERR_ARG;
    const long cells = (long)across * down;
    if (eastward) zonal_stress(across, down, seed, month, eastward);
    if (northward) std::fill(northward, northward + cells, 0.0);
    return CG_OK;
}

// M = -(I - W)^-1 over a grid's water cells (row-major order), for a direct solve of that grid:
// x = M f gives, on water, x = W x - f exactly. I - W over the water cells is banded (half-bandwidth
// at most the grid's width: the east-west wrap and the north and south neighbours) and an M-matrix
// with weakly dominant diagonal (weights non-negative, each row's summing to at most 1), so banded
// LU without pivoting, in binary64, keeps its pivots positive; column j of the inverse is one
// forward and one backward banded solve of e_j, the columns shared among `threads` host threads.
// Writes the water cells to idx[n] and M row-major to m[n * n] in binary32; returns n, 0 when there
// is no water, more than max_n water cells, or a pivot that is not positive and finite, and
// CG_ERR_ARG on bad arguments.
int cg_bottom_inverse(int across, int down, const uint8_t *water, const float *east, const float *west,
                      const float *north, const float *south, int max_n, int threads, int32_t *idx,
                      double *m) {
    if (across <= 0 || down <= 0 || !water || !east || !west || !north || !south || max_n < 0 || !idx || !m)
        return CG_ERR_ARG;
    const long cells = (long)across * down;
    std::vector<int> cell_of;
    std::vector<int> pos(cells, -1);
    for (long c = 0; c < cells; ++c)
        if (water[c]) {
            pos[c] = (int)cell_of.size();
            cell_of.push_back((int)c);
        }
    const int n = (int)cell_of.size();
    if (n == 0 || n > max_n) return 0;
    auto neighbours = [&](int i, auto &&visit) {
        const int cell = cell_of[i];
        const int row = cell / across, col = cell % across;
        const int ce = col + 1 == across ? 0 : col + 1;
        const int cw = col == 0 ? across - 1 : col - 1;
        visit(row * across + ce, (double)east[cell]);
        visit(row * across + cw, (double)west[cell]);
        if (row > 0) visit(cell - across, (double)north[cell]);
        if (row + 1 < down) visit(cell + across, (double)south[cell]);
    };
    int bw = 0;
    for (int i = 0; i < n; ++i)
        neighbours(i, [&](int nb, double) {
            if (pos[nb] >= 0) bw = std::max(bw, std::abs(pos[nb] - i));
        });
    const int width = 2 * bw + 1;
    std::vector<double> a((size_t)n * width, 0.0);
    auto at = [&](int i, int j) -> double & { return a[(size_t)i * width + (j - i + bw)]; };
    for (int i = 0; i < n; ++i) {
        at(i, i) = 1.0;
        neighbours(i, [&](int nb, double wt) {
            if (pos[nb] >= 0) at(i, pos[nb]) -= wt;
        });
    }
    for (int k = 0; k < n; ++k) {
        const double piv = at(k, k);
        if (!(piv > 0.0) || !std::isfinite(piv)) return 0;
        const int last = std::min(n - 1, k + bw);
        for (int i = k + 1; i <= last; ++i) {
            double &lik = at(i, k);
            if (lik == 0.0) continue;
            lik /= piv;
            for (int j = k + 1; j <= last; ++j) at(i, j) -= lik * at(k, j);
        }
    }
    const int nthreads = std::max(1, threads);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&, t]() {
            std::vector<double> y(n);
            for (int j = t; j < n; j += nthreads) {
                std::fill(y.begin(), y.end(), 0.0);
                y[j] = 1.0;
                for (int i = j + 1; i < n; ++i) {  // L y = e_j (unit lower, banded)
                    double v = y[i];
                    for (int k = std::max(j, i - bw); k < i; ++k) v -= at(i, k) * y[k];
                    y[i] = v;
                }
                for (int i = n - 1; i >= 0; --i) {  // U x = y
                    double v = y[i];
                    const int last = std::min(n - 1, i + bw);
                    for (int k = i + 1; k <= last; ++k) v -= at(i, k) * y[k];
                    y[i] = v / at(i, i);
                }
                for (int i = 0; i < n; ++i) m[(size_t)i * n + j] = -y[i];
            }
        });
    for (auto &th : pool) th.join();
    for (int i = 0; i < n; ++i) idx[i] = cell_of[i];
    return n;
}

}  // extern "C"
