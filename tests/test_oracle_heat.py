"""The oracle's heat problem, pinned (CPU only).

The reference's OceanHeatTest (worldgen/src/commonTest/kotlin/com/cartogenesis/worldgen/
OceanHeatTest.kt) holds closed forms and bars for this path; each is restated here against the
oracle with the processor's relaxation, as the reference test runs it:
- the eddy diffusivity (:410-451): 2,500 m^2/s at 45 degrees, monotone poleward of the equatorial
  reach, 240 km equatorial deformation radius on Earth, the square-root scaling with the radius,
  non-negative weights summing below one;
- a current carries heat the same distance whichever way it flows (:70-109): mirror images east/west
  and north/south within 1e-3 C, decay length within 5% of the one-dimensional balance's 1/|m|;
- no heat crosses a strip of land one cell wide (:193-219);
- a closed basin neither makes nor loses heat at its coast (:265-336);
- the solve converges on a difficult fixture and a tenth of its iterations does not (:367-404).
Beyond those: the heat stencil's closed form for still water (every weight K/d^2 over the centre)
and for a uniform current (the upstream face weighted by B(-P), the downstream by B(P)); the margin
c (1 - sum of weights) equal to 1/tau plus the entrainment; the solve against a scipy direct solve
of an independently assembled sparse operator; waterBodies against a brute-force union-find;
correctBodyMeans zeroing each body's summed residual; resample and coarseMean on constants and
linear fields.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle

TAU = 50.0 * 4.0e6 / 25.0          # OceanStage.RELAXATION_SECONDS (OS:113-114, EnergyBalance.kt:688-832)
EARTH_RADIUS = 6.371e6             # OceanStage.EARTH_MEAN_RADIUS_M (OS:80)
WORLD_RADIUS = 40075e3 / (2 * math.pi)  # WorldGenConfig().scale.radiusMeters (WGC:246)
TOL = 1e-3                         # OceanCirculation.RESIDUAL_TOLERANCE
SHORE = 2
STEP_C = 10.0


def heat_solve(across, down, dx, dy, water, stream, target, constant_k=0.0, tol=TOL, most=200,
               radius=WORLD_RADIUS):
    """OceanHeatTest.solve (:30-54): finest stencil, coarse grids by the test's point-sampled stream,
    solveByKrylov with the water bodies, pole not a wall."""
    water = np.asarray(water, np.uint8).reshape(-1)
    stream = np.asarray(stream, np.float32).reshape(-1)
    target = np.asarray(target, np.float32).reshape(-1)
    finest = oracle.heat_stencil(across, down, dx, dy, water, stream, target, TAU, True, radius, constant_k)

    def coarser(a, d, w):
        cell = np.arange(a * d)
        cs = stream[((cell // a) * down // d) * across + (cell % a) * across // a]
        return oracle.heat_stencil(a, d, dx * across / a, dy * down / d, w, cs, np.zeros(a * d, np.float32), TAU,
                                   False, radius, constant_k)

    ladder = oracle.levels(finest, dx, dy, False, coarser)
    body_of, count = oracle.water_bodies(water, across, down)
    x, it, rr = oracle.solve_bodies(ladder, target, tol, False, oracle.KRYLOV, body_of, count, most)
    return x, it, rr, finest


def channel_stream(index, cells, per_cell):
    half = cells / 2.0 - SHORE
    offset = index + 0.5 - cells / 2.0
    steady = half * 2.0 / 3.0
    if abs(offset) <= steady:
        value = offset
    else:
        value = math.copysign(1.0, offset) * steady * (half - abs(offset)) / (half - steady)
    return np.float32(per_cell * value)


def zonal_channel(across, down, dx, dy, speed, k):
    cell = np.arange(across * down)
    row, col = cell // across, cell % across
    water = ((row >= SHORE) & (row < down - SHORE)).astype(np.uint8)
    stream = np.array([channel_stream(r, down, speed * dy) if w else 0.0 for r, w in zip(row, water)], np.float32)
    warm = (col < across // 2) if speed > 0 else (col >= across // 2)
    target = np.where((water != 0) & warm, STEP_C, 0.0).astype(np.float32)
    return heat_solve(across, down, dx, dy, water, stream, target, constant_k=k)[0]


def meridional_channel(across, down, dx, dy, speed, k):
    cell = np.arange(across * down)
    row, col = cell // across, cell % across
    water = ((col >= SHORE) & (col < across - SHORE)).astype(np.uint8)
    stream = np.array([channel_stream(c, across, speed * dx) if w else 0.0 for c, w in zip(col, water)], np.float32)
    warm = (row >= down // 2) if speed > 0 else (row < down // 2)
    target = np.where((water != 0) & warm, STEP_C, 0.0).astype(np.float32)
    return heat_solve(across, down, dx, dy, water, stream, target, constant_k=k)[0]


def decay_length(profile, spacing):
    threshold = profile[0] / math.e
    for i in range(1, len(profile)):
        if profile[i] <= threshold:
            share = (profile[i - 1] - threshold) / (profile[i - 1] - profile[i])
            return (i - 1 + share) * spacing
    return float("nan")


# ---- the diffusivity ----------------------------------------------------------------------------

def test_eddy_diffusivity_is_v_times_the_deformation_radius():
    for radius in (EARTH_RADIUS, WORLD_RADIUS):
        assert abs(oracle.heat_diffusivity(45.0, radius) - 2500.0) < 1e-6
        assert abs(oracle.heat_diffusivity(-45.0, radius) - 2500.0) < 1e-6
        at_eq = oracle.heat_diffusivity(0.0, radius)
        prev = math.inf
        for tenth in range(900):
            here = oracle.heat_diffusivity(tenth / 10.0, radius)
            assert math.isfinite(here)
            if tenth < 80:
                assert here <= at_eq * 1.01
            else:
                assert here <= prev
                prev = here
            assert abs(here - oracle.heat_diffusivity(-tenth / 10.0, radius)) < 1e-9
    assert abs(oracle.deformation_radius(0.0, EARTH_RADIUS) - 240000.0) < 1e-6
    k_eq = oracle.heat_diffusivity(0.0, EARTH_RADIUS)
    assert 20000.0 <= k_eq <= 30000.0
    assert abs(math.sqrt(EARTH_RADIUS / WORLD_RADIUS) - k_eq / oracle.heat_diffusivity(0.0, WORLD_RADIUS)) < 1e-9
    st = oracle.heat_stencil(8, 8, 10000.0, 5000.0, np.ones(64, np.uint8), np.arange(64, dtype=np.float32) * 1000,
                             np.full(64, 10, np.float32), TAU, True, WORLD_RADIUS)
    total = st.east + st.west + st.north + st.south
    assert np.all(st.east >= 0) and np.all(st.west >= 0) and np.all(st.north >= 0) and np.all(st.south >= 0)
    assert np.all(total < 1)


# ---- the stencil's closed forms -----------------------------------------------------------------

def bernoulli(x):
    return 1.0 - x / 2.0 if abs(x) < 1e-3 else (0.0 if x > 700 else x / math.expm1(x))


def test_still_water_weights_are_the_diffusion_stencil():
    across, down, dx, dy, k = 16, 8, 30000.0, 20000.0, 4000.0
    water = np.ones(across * down, np.uint8)
    water[3 * across + 5] = 0
    st = oracle.heat_stencil(across, down, dx, dy, water, np.zeros(across * down, np.float32),
                             np.full(across * down, 7, np.float32), TAU, True, constant_k=k)
    ax, ay = k / dx ** 2, k / dy ** 2
    W = water.reshape(down, across) != 0
    for r in range(down):
        for c in range(across):
            i = r * across + c
            if not W[r, c]:
                assert st.east[i] == st.west[i] == st.north[i] == st.south[i] == 0 and st.centre[i] == 1
                assert st.forcing[i] == 0
                continue
            e = ax if W[r, (c + 1) % across] else 0.0
            w = ax if W[r, (c - 1) % across] else 0.0
            n = ay if r > 0 and W[r - 1, c] else 0.0
            s = ay if r + 1 < down and W[r + 1, c] else 0.0
            centre = e + w + n + s + 1.0 / TAU
            assert st.centre[i] == pytest.approx(centre, rel=1e-15)
            for got, want in ((st.east[i], e), (st.west[i], w), (st.north[i], n), (st.south[i], s)):
                assert got == np.float32(want / centre)
            # The balance -T / tau over the centre.
            assert st.forcing[i] == np.float32(-7.0 / TAU / centre)


def test_a_uniform_current_weights_its_upstream_face():
    # psi rising southward by u dy a row: u eastward everywhere away from the poles.
    across, down, dx, dy, k, u = 32, 16, 25000.0, 25000.0, 5000.0, 0.3
    water = np.ones(across * down, np.uint8)
    stream = np.repeat(np.arange(down, dtype=np.float32) * np.float32(u * dy), across)
    st = oracle.heat_stencil(across, down, dx, dy, water, stream, np.zeros(across * down, np.float32), TAU, True,
                             constant_k=k)
    p = u * dx / k
    down_w, up_w = k / dx ** 2 * bernoulli(p), k / dx ** 2 * bernoulli(-p)
    for r in range(2, down - 2):
        i = r * across + 7
        centre = st.centre[i]
        # Water leaves through the east face: the east (downstream) neighbour weighs B(P), the west
        # (upstream) B(-P) = B(P) + P, more; north and south see no flow.
        assert st.east[i] == np.float32(down_w / centre)
        assert st.west[i] == np.float32(up_w / centre)
        assert st.north[i] == st.south[i] == np.float32(k / dy ** 2 / centre)
        assert st.west[i] > st.east[i]


def test_the_margin_is_the_relaxation_and_the_entrainment():
    rng = np.random.default_rng(4)
    across, down = 24, 12
    water = (rng.random(across * down) > 0.2).astype(np.uint8)
    stream = rng.normal(0, 3e4, across * down).astype(np.float32)
    ent = rng.uniform(0, 1e-6, across * down).astype(np.float32)
    sub = rng.uniform(2, 10, across * down).astype(np.float32)
    target = rng.uniform(0, 28, across * down).astype(np.float32)
    st = oracle.heat_stencil(across, down, 50e3, 50e3, water, stream, target, TAU, True, WORLD_RADIUS, 0.0, ent, sub)
    wet = water != 0
    weights = (st.east.astype(np.float64) + st.west + st.north + st.south)[wet]
    margin = st.centre[wet] * (1.0 - weights)
    want = 1.0 / TAU + ent[wet].astype(np.float64)
    assert np.allclose(margin, want, rtol=0, atol=1e-6 * st.centre[wet].max() * 1e-1)
    bal = -target[wet].astype(np.float64) / TAU - ent[wet].astype(np.float64) * sub[wet]
    assert np.allclose(st.forcing[wet].astype(np.float64) * st.centre[wet], bal, rtol=1e-6)


# ---- the solve, against a direct solve ------------------------------------------------------------

def sparse_operator(st):
    """(I - W) c over water, identity on land, assembled independently of the oracle's sweeps."""
    n = st.cells
    a, d = st.across, st.down
    rows, cols, vals = [], [], []
    for i in range(n):
        r, c = divmod(i, a)
        if not st.water[i]:
            rows.append(i); cols.append(i); vals.append(1.0)
            continue
        rows.append(i); cols.append(i); vals.append(st.centre[i])
        for wgt, j in ((st.east[i], r * a + (c + 1) % a), (st.west[i], r * a + (c - 1) % a),
                       (st.north[i], i - a if r > 0 else -1), (st.south[i], i + a if r + 1 < d else -1)):
            if j >= 0 and wgt != 0:
                rows.append(i); cols.append(j); vals.append(-float(wgt) * st.centre[i])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    rhs = np.where(st.water != 0, -st.forcing.astype(np.float64) * st.centre, 0.0)
    return A, rhs


def test_heat_solve_reaches_the_direct_solution():
    rng = np.random.default_rng(8)
    across, down, dx = 64, 32, 60e3
    cell = np.arange(across * down)
    row, col = cell // across, cell % across
    water = ((np.sin(col / 5.0) + np.cos(row / 3.0) + rng.normal(0, 0.3, cell.size)) > -0.6).astype(np.uint8)
    stream = (3e4 * np.sin(2 * np.pi * col / across) * np.sin(np.pi * (row + 0.5) / down)).astype(np.float32)
    stream[water == 0] = 0
    target = np.where(water != 0, 28.0 - 25.0 * np.abs(row + 0.5 - down / 2) / (down / 2), 0.0).astype(np.float32)
    x, it, rr, finest = heat_solve(across, down, dx, dx, water, stream, target, tol=1e-8)
    assert rr < 1e-8 and it > 1
    A, rhs = sparse_operator(finest)
    direct = spla.spsolve(A.tocsc(), rhs)
    wet = water != 0
    assert np.abs(x[wet] - direct[wet]).max() < 1e-5 * np.abs(direct[wet]).max()
    assert np.abs(A @ direct - rhs).max() < 1e-9 * np.abs(rhs).max()


# ---- OceanHeatTest's cases ----------------------------------------------------------------------

@pytest.mark.parametrize("aspect", [1.0, 0.5])
def test_a_current_carries_heat_the_same_distance_whichever_way_it_flows(aspect):
    speed, k = 0.1, 5000.0
    expected = 2 * k / (math.sqrt(speed * speed + 4 * k / TAU) - speed)
    dx = 25000.0
    dy = dx * aspect
    across = 256
    down = int(128 / aspect)
    east = zonal_channel(across, down, dx, dy, speed, k)
    west = zonal_channel(across, down, dx, dy, -speed, k)
    mid = down // 2
    E, Wt = east.reshape(down, across), west.reshape(down, across)
    worst_mirror = np.abs(E[mid] - Wt[mid][::-1]).max()
    north = meridional_channel(across, down, dx, dy, speed, k)
    south = meridional_channel(across, down, dx, dy, -speed, k)
    col = across // 2
    N, S = north.reshape(down, across), south.reshape(down, across)
    worst_ns = np.abs(N[:, col] - S[::-1, col]).max()
    east_len = decay_length(E[mid, across // 2:].astype(np.float64), dx)
    north_len = decay_length(N[down // 2 - 1::-1, col].astype(np.float64), dy)
    print(f"OCEAN heat carry aspect {aspect}: mirror {worst_mirror}, {worst_ns}; decay {east_len / 1e3} km east, "
          f"{north_len / 1e3} km north, against {expected / 1e3}")
    assert worst_mirror < 1e-3 and worst_ns < 1e-3
    assert abs(east_len - expected) < 0.05 * expected
    assert abs(north_len - expected) < 0.05 * expected


@pytest.mark.parametrize("aspect", [1.0, 0.5])
def test_no_heat_crosses_a_strip_of_land_one_cell_wide(aspect):
    across, down, dx = 128, 64, 50000.0
    dy = dx * aspect
    strip = across // 2
    col = np.arange(across * down) % across
    water = ((col != strip) & (col != 0)).astype(np.uint8)
    stirred = np.where((water != 0) & (col > strip), 2000.0 * (col - strip) * dx * 1e-3, 0.0).astype(np.float32)

    def far_target(c):
        return np.where(water == 0, 0.0, np.where(col < strip, 5.0, c)).astype(np.float32)

    cold = heat_solve(across, down, dx, dy, water, np.zeros(across * down, np.float32), far_target(0.0))[0]
    warm = heat_solve(across, down, dx, dy, water, stirred, far_target(10.0))[0]
    near = (water != 0) & (col < strip)
    worst = np.abs(cold[near] - warm[near]).max()
    print(f"OCEAN barrier aspect {aspect}: the near basin moved {worst} C")
    assert worst < 1e-3


@pytest.mark.parametrize("aspect", [1.0, 0.5])
def test_a_closed_basin_neither_makes_nor_loses_heat_at_its_coast(aspect):
    across = 96
    down = int(48 / aspect)
    dx = 50000.0
    dy = dx * aspect
    wm, hm = across * dx, down * dy
    radius = 0.42 * min(wm, hm)
    cell = np.arange(across * down)
    x = (cell % across + 0.5) * dx - wm / 2
    y = (cell // across + 0.5) * dy - hm / 2
    ang = np.arctan2(y, x)
    coast = radius * (1 + 0.18 * np.sin(3 * ang) + 0.09 * np.cos(5 * ang) + 0.05 * np.sin(11 * ang))
    ix, iy = x - 0.35 * radius, y + 0.2 * radius
    water = ((np.sqrt(x * x + y * y) < coast) & (np.sqrt(ix * ix + iy * iy) > 0.15 * radius)).astype(np.uint8)
    xs = (cell % across + 0.5) / across
    ys = (cell // across + 0.5) / down
    stream = np.where(water != 0, 50000.0 * (np.sin(2 * np.pi * xs + 0.3) * np.cos(np.pi * ys) + 0.5 * xs),
                      0.0).astype(np.float32)
    target = np.where(water != 0, 25.0 - 20.0 * (cell // across + 0.5) / down + 3.0 * np.sin(0.2 * (cell % across)),
                      0.0).astype(np.float32)
    solved, _, _, finest = heat_solve(across, down, dx, dy, water, stream, target, tol=1e-9)
    carried = scale = 0.0
    for i in np.nonzero(water)[0]:
        r, c = divmod(int(i), across)
        wgt = finest.centre[i]
        here = float(solved[i])
        for share, nb in ((finest.east[i], r * across + (c + 1) % across),
                          (finest.west[i], r * across + (c + across - 1) % across),
                          (finest.north[i], i - across if r > 0 else i),
                          (finest.south[i], i + across if r + 1 < down else i)):
            term = wgt * float(share) * (here - float(solved[nb]))
            carried += term
            scale += abs(term)
    wet = water != 0
    budget = float(np.sum(solved[wet].astype(np.float64) - target[wet]))
    budget_scale = float(np.sum(np.abs(solved[wet].astype(np.float64) - target[wet])))
    uniform = np.where(wet, 14.0, 0.0).astype(np.float32)
    flat = heat_solve(across, down, dx, dy, water, stream, uniform, tol=1e-9)[0]
    flat_worst = np.abs(flat[wet] - 14.0).max()
    print(f"OCEAN heat budget aspect {aspect}: carried {carried / scale}, budget {budget / budget_scale}, "
          f"uniform {flat_worst}")
    assert abs(carried) < 5e-5 * scale
    assert abs(budget) < 1e-4 * budget_scale
    assert flat_worst < 1e-4


def difficult_fixture():
    across, down, dx = 256, 128, 25000.0
    cell = np.arange(across * down)
    row, col = cell // across, cell % across
    sea = (col >= 20) & (col < 200) & (row >= 10) & (row < 118)
    inlet = (row == 60) & (col >= 200) & (col < 240)
    pond = (row >= 30) & (row <= 33) & (col >= 201) & (col <= 204)
    water = (sea | inlet | pond).astype(np.uint8)
    stream = np.where((water != 0) & (col < 200),
                      20000.0 * np.sin(np.pi * (col - 20) / 180.0) * np.sin(np.pi * (row - 10) / 108.0),
                      0.0).astype(np.float32)
    target = np.where(water != 0, np.float32(30.0) - np.float32(0.4) * row.astype(np.float32), 0.0).astype(np.float32)
    return across, down, dx, water, stream, target


def test_the_heat_solve_converges_on_a_difficult_fixture_and_a_tenth_does_not():
    across, down, dx, water, stream, target = difficult_fixture()
    _, it, rr, _ = heat_solve(across, down, dx, dx, water, stream, target)
    print(f"OCEAN difficult heat fixture: {it} iterations to {rr}")
    assert rr < TOL
    _, it2, rr2, _ = heat_solve(across, down, dx, dx, water, stream, target, most=max(it // 10, 1))
    assert rr2 >= TOL


# ---- water bodies, body means, resampling -------------------------------------------------------

def bodies_brute(water, across, down):
    parent = list(range(across * down))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    W = np.asarray(water).reshape(-1) != 0
    for i in range(across * down):
        if not W[i]:
            continue
        r, c = divmod(i, across)
        for j in (r * across + (c + 1) % across, (r + 1) * across + c if r + 1 < down else -1):
            if j >= 0 and W[j]:
                a, b = find(i), find(j)
                if a != b:
                    parent[max(a, b)] = min(a, b)
    label, out, nxt = {}, np.full(across * down, -1, np.int32), 0
    for i in range(across * down):
        if W[i]:
            root = find(i)
            if root not in label:
                label[root] = nxt
                nxt += 1
            out[i] = label[root]
    return out, nxt


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_water_bodies_match_a_brute_force_union_find(seed):
    rng = np.random.default_rng(seed)
    across, down = 40, 20
    water = (rng.random(across * down) > 0.45).astype(np.uint8)
    got, n = oracle.water_bodies(water, across, down)
    want, m = bodies_brute(water, across, down)
    assert n == m and n > 5
    assert np.array_equal(got, want)


def test_a_body_joined_only_across_the_date_line_is_one_body():
    across, down = 10, 3
    water = np.zeros((down, across), np.uint8)
    water[1, 0] = water[1, across - 1] = 1
    body, n = oracle.water_bodies(water.reshape(-1), across, down)
    assert n == 1 and body[across] == body[2 * across - 1] == 0


def test_correct_body_means_zeroes_each_bodys_summed_residual():
    rng = np.random.default_rng(5)
    across, down = 48, 24
    col = np.arange(across * down) % across
    water = ((col != 10) & (col != 30)).astype(np.uint8)  # two bodies, columns wrapping
    stream = rng.normal(0, 2e4, across * down).astype(np.float32)
    target = rng.uniform(0, 20, across * down).astype(np.float32)
    st = oracle.heat_stencil(across, down, 50e3, 50e3, water, stream, target, TAU, True, WORLD_RADIUS)
    body, n = oracle.water_bodies(water, across, down)
    assert n == 2
    x = rng.uniform(0, 20, across * down).astype(np.float32)
    y = oracle.correct_body_means(st, x, body, n)
    shift = (x - y)[water != 0]
    for b in range(n):
        s = (x.astype(np.float64) - y)[body == b]
        # One binary32 shift per body, subtracted in binary32: uniform to the rounding of the values.
        assert np.allclose(s, s[0], rtol=0, atol=4e-6 * 20) and s[0] != 0
        r_before = oracle.residual64(st, x.astype(np.float64))[body == b]
        r_after = oracle.residual64(st, y.astype(np.float64))[body == b]
        assert abs(r_after.sum()) < 1e-6 * np.abs(r_before).sum()
    assert shift.size > 0


def test_resample_and_coarse_mean_on_simple_fields():
    fa, fd = 64, 32
    const = np.full(fa * fd, 3.5, np.float32)
    assert np.all(oracle.resample(const, fa, fd, 16, 8) == np.float32(3.5))
    assert np.all(oracle.coarse_mean(const, fa, fd, 16, 8) == np.float32(3.5))
    # Linear in the row, at the centres: exact where no clamping is involved.
    lin = np.repeat(np.arange(fd, dtype=np.float32), fa)
    r = oracle.resample(lin, fa, fd, fa, fd // 2).reshape(fd // 2, fa)
    assert np.allclose(r[:, 0], 2 * np.arange(fd // 2) + 0.5, atol=1e-6)
    # The identity when both grids are one.
    rng = np.random.default_rng(2)
    f = rng.normal(size=fa * fd).astype(np.float32)
    assert np.array_equal(oracle.resample(f, fa, fd, fa, fd), f)
    # coarseMean conserves the total of each block.
    m = oracle.coarse_mean(f, fa, fd, fa // 2, fd // 2).reshape(fd // 2, fa // 2)
    blocks = f.reshape(fd // 2, 2, fa // 2, 2).astype(np.float64).mean(axis=(1, 3))
    assert np.allclose(m, blocks, atol=1e-6)
    # Columns wrap: a field that is a sine of the column resampled at half the columns.
    c = np.tile(np.arange(fa), fd)
    sine = np.sin(2 * np.pi * c / fa).astype(np.float32)
    half = oracle.resample(sine, fa, fd, fa // 2, fd).reshape(fd, fa // 2)
    cc = np.arange(fa // 2)
    pos = (cc + 0.5) * 2 - 0.5
    assert np.allclose(half[0], np.sin(2 * np.pi * pos / fa) * np.cos(np.pi / fa), atol=2e-3)
