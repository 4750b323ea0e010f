"""The heat problem on the device, against the oracle.

- GpuOceanTest's heat batches (GpuOceanTest.kt:52-68, 170-196): a heat stencil with and without
  upwelling relaxed through the seam is bit-identical to the processor's relax (and so inside the
  reference's batch bound), at both aspects.
- correctBodyMeans (OceanCirculation.kt:709-731) on the device against the oracle: the per-body sums
  are tree sums on the device and sequential in the oracle, so the binary32 shift may differ by one
  unit in its last place; values agree to within that.
- The heat solve (CG_METHOD_KRYLOV with bodies, pole not a wall: OceanStage.kt:486-497) on
  OceanHeatTest's fixtures: converged where the oracle converges, iteration counts within two (the
  BiCGSTAB scalars come from blocked dot products), the answer within the bound the solve's tolerance
  sets (the operator is diagonally dominant by 1/tau, so |x - x*| <= tau max|r|), and against a
  scipy direct solve at a tight tolerance; OceanHeatTest's bars (difficult fixture, barrier, budget)
  hold for the device's answers.
- A whole round on the device: the circulation solved, its velocities, the heat stencil from its
  stream, the heat solved, at 256 x 512.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from conftest import BATCH_PASSES, ROUNDED_OPERATIONS_PER_UPDATE, UNIT_ROUNDOFF, to_oracle
from cartogenesis_amd import CG_METHOD_KRYLOV, CG_METHOD_REFINED, CG_OPT_MIN_CELLS
from cartogenesis_amd.ocean import OceanCirculation as OC, OceanHeat, Plan, Solution
from cartogenesis_amd.problem import circulation_problem
from test_oracle_heat import difficult_fixture, sparse_operator

pytestmark = pytest.mark.gpu

TAU = OceanHeat.RELAXATION_SECONDS
WORLD_RADIUS = 40075e3 / (2 * math.pi)
TOL = OC.RESIDUAL_TOLERANCE


def gpu_heat_stencil(aspect, with_upwelling=False):
    """GpuOceanTest.heatStencil (GpuOceanTest.kt:170-196)."""
    across = 512
    down = int(256 / aspect)
    dx = 6250.0
    dy = dx * aspect
    cell = np.arange(across * down)
    col, row = cell % across, cell // across
    water = ((col != 0) & (col != across - 1)).astype(np.uint8)
    stream = (2e4 * np.sin(np.pi * col / across) * np.sin(np.pi * (row + 0.5) / down)).astype(np.float32)
    target = np.where(water != 0, np.float32(5) + (np.float32(20) * row.astype(np.float32)) / np.float32(down),
                      0).astype(np.float32)
    ent = sub = None
    if with_upwelling:
        rising = np.where(col == across - 2, 1.2e-4, np.where(row >= down // 2, 1.2e-6, 0.0))
        ent = (rising / 50.0).astype(np.float32)
        sub = (target - np.float32(5)).astype(np.float32)
    return OceanHeat.stencil(across, down, dx, dy, water, stream, target, TAU, True, WORLD_RADIUS, None, ent, sub)


@pytest.mark.parametrize("aspect", [1.0, 0.5])
@pytest.mark.parametrize("upwelling", [False, True])
def test_heat_batches_are_bit_identical_to_the_processor(acc, aspect, upwelling):
    st = gpu_heat_stencil(aspect, upwelling)
    start = np.where(st.is_water != 0, 1.0, 0.0).astype(np.float32)
    print(f"OCEAN heat batch aspect {aspect} upwelling {upwelling}", flush=True)
    got = acc.solve(st, start, BATCH_PASSES)
    want = oracle.relax(to_oracle(st), start, BATCH_PASSES)
    assert got is not None
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    bound = BATCH_PASSES * ROUNDED_OPERATIONS_PER_UPDATE * UNIT_ROUNDOFF * np.abs(want).max()
    assert np.abs(got.astype(np.float64) - want).max() <= bound


def heat_ladder(across, down, dx, dy, water, stream, target, constant_k=None):
    """OceanHeatTest.solve's ladder (:30-54): coarse streams point-sampled."""
    stream = np.asarray(stream, np.float32).reshape(-1)
    finest = OceanHeat.stencil(across, down, dx, dy, water, stream, target, TAU, True, WORLD_RADIUS, constant_k)

    def coarser(a, d, w):
        cell = np.arange(a * d)
        cs = stream[((cell // a) * down // d) * across + (cell % a) * across // a]
        return OceanHeat.stencil(a, d, dx * across / a, dy * down / d, w, cs, None, TAU, False, WORLD_RADIUS, constant_k)

    return OC.levels(finest, dx, dy, False, coarser)


def device_heat(acc, levels, target, tol):
    plan = Plan(acc, levels, pole_is_wall=False)
    try:
        fin = levels[0]
        plan.set_bodies(OC.water_bodies(fin.is_water, fin.cells_across, fin.cells_down))
        plan.set_start(target)
        info = plan.solve(tol, CG_METHOD_KRYLOV, warm=True)
        _, x32 = plan.read()
        return x32, info
    finally:
        plan.close()


def oracle_heat(levels, target, tol):
    """The heat's loop (OceanStage.kt:486-497) on the oracle: solveByKrylov, restarted up to three
    times while short of tol."""
    ladder = [to_oracle(s) for s in levels]
    fin = levels[0]
    body, count = oracle.water_bodies(fin.is_water, fin.cells_across, fin.cells_down)
    x, it, rr = oracle.solve_bodies(ladder, target, tol, False, oracle.KRYLOV, body, count)
    total, starts = it, 1
    while rr >= tol and starts <= 3 and np.all(np.isfinite(x)):
        x, it, rr = oracle.solve_bodies(ladder, x, tol, False, oracle.KRYLOV, body, count)
        total += it
        starts += 1
    return x, total, rr


def test_correct_body_means_matches_the_oracle(acc):
    rng = np.random.default_rng(5)
    across, down = 256, 128
    cell = np.arange(across * down)
    col, row = cell % across, cell // across
    water = ((col % 50 != 7) & ~((row == 40) & (col > 100))).astype(np.uint8)
    water[(row > 90) & (row < 95) & (col > 60) & (col < 64)] = 1
    stream = rng.normal(0, 2e4, across * down).astype(np.float32) * water
    target = rng.uniform(0, 25, across * down).astype(np.float32) * water
    levels = heat_ladder(across, down, 40e3, 40e3, water, stream, target)
    bodies = OC.water_bodies(water, across, down)
    assert bodies.count >= 5
    x = rng.uniform(0, 25, across * down).astype(np.float32)
    plan = Plan(acc, levels, pole_is_wall=False)
    try:
        plan.set_bodies(bodies)
        got = plan.correct_bodies(x)
    finally:
        plan.close()
    want = oracle.correct_body_means(to_oracle(levels[0]), x, bodies.body_of, bodies.count)
    ulp = np.spacing(np.maximum(np.abs(want), 1.0).astype(np.float32))
    assert np.all(np.abs(got - want) <= 2 * ulp)
    assert np.mean(got == want) > 0.9
    assert not np.array_equal(got, x)


def test_heat_solve_on_the_difficult_fixture(acc):
    across, down, dx, water, stream, target = difficult_fixture()
    levels = heat_ladder(across, down, dx, dx, water, stream, target)
    print("OCEAN heat difficult fixture", flush=True)
    x, info = device_heat(acc, levels, target, TOL)
    xo, it_o, rr_o = oracle_heat(levels, target, TOL)
    print(f"OCEAN heat difficult fixture: device {info.cycles} iterations to {info.relative_residual}, "
          f"oracle {it_o} to {rr_o}")
    assert info.relative_residual < TOL and rr_o < TOL
    assert abs(info.cycles - it_o) <= 2
    bal = np.abs(OC.balance_of(levels[0])).max()
    bound = 2 * TAU * TOL * bal
    assert np.abs(x.astype(np.float64) - xo).max() <= bound


def test_heat_solve_reaches_the_direct_solution(acc):
    rng = np.random.default_rng(8)
    across, down, dx = 128, 64, 60e3
    cell = np.arange(across * down)
    row, col = cell // across, cell % across
    water = ((np.sin(col / 5.0) + np.cos(row / 3.0) + rng.normal(0, 0.3, cell.size)) > -0.6).astype(np.uint8)
    stream = (3e4 * np.sin(2 * np.pi * col / across) * np.sin(np.pi * (row + 0.5) / down)).astype(np.float32) * water
    target = np.where(water != 0, 28.0 - 25.0 * np.abs(row + 0.5 - down / 2) / (down / 2), 0.0).astype(np.float32)
    levels = heat_ladder(across, down, dx, dx, water, stream, target)
    print("OCEAN heat direct", flush=True)
    x, info = device_heat(acc, levels, target, 1e-8)
    assert info.relative_residual < 1e-8
    A, rhs = sparse_operator(to_oracle(levels[0]))
    direct = spla.spsolve(A.tocsc(), rhs)
    wet = water != 0
    assert np.abs(x[wet] - direct[wet]).max() < 1e-5 * np.abs(direct[wet]).max()


@pytest.mark.parametrize("aspect", [1.0, 0.5])
def test_no_heat_crosses_a_strip_on_the_device(acc, aspect):
    across, down, dx = 128, 64, 50000.0
    dy = dx * aspect
    strip = across // 2
    col = np.arange(across * down) % across
    water = ((col != strip) & (col != 0)).astype(np.uint8)
    stirred = np.where((water != 0) & (col > strip), 2000.0 * (col - strip) * dx * 1e-3, 0.0).astype(np.float32)

    def far(c):
        return np.where(water == 0, 0.0, np.where(col < strip, 5.0, c)).astype(np.float32)

    cold_t, warm_t = far(0.0), far(10.0)
    cold, _ = device_heat(acc, heat_ladder(across, down, dx, dy, water, np.zeros(across * down), cold_t), cold_t, TOL)
    warm, _ = device_heat(acc, heat_ladder(across, down, dx, dy, water, stirred, warm_t), warm_t, TOL)
    near = (water != 0) & (col < strip)
    assert np.abs(cold[near] - warm[near]).max() < 1e-3


def test_a_closed_basin_keeps_its_heat_budget_on_the_device(acc):
    across, down, dx = 96, 48, 50000.0
    dy = dx
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
    solved, info = device_heat(acc, heat_ladder(across, down, dx, dy, water, stream, target), target, 1e-9)
    wet = water != 0
    budget = float(np.sum(solved[wet].astype(np.float64) - target[wet]))
    scale = float(np.sum(np.abs(solved[wet].astype(np.float64) - target[wet])))
    print(f"OCEAN device heat budget: {budget / scale} after {info.cycles} iterations")
    assert abs(budget) < 1e-4 * scale


def test_a_whole_round_on_the_device(acc):
    """circulateUnder's two solves (OceanStage.kt:417-498) with everything but the stencil building on
    the device: psi by the refined method, velocities, the heat stencil from psi, the heat by
    solveByKrylov with bodies."""
    p = circulation_problem(256, seed=3)
    across, down = p.cols, p.rows
    w, h = 40075e3 / across, 20037.5e3 / down
    plan = Plan(acc, p.levels)
    try:
        info = plan.solve(TOL, CG_METHOD_REFINED)
        assert info.relative_residual < TOL
        _, stream = plan.read()
        east, north = plan.velocities(w, h)
    finally:
        plan.close()
    assert np.abs(east).max() > 0
    row = np.arange(across * down) // across
    lat = 90.0 - 180.0 * (row + 0.5) / down
    target = np.where(p.water != 0, 28.0 * np.cos(np.radians(lat)) - 2.0, 0.0).astype(np.float32)
    levels = OceanHeat.levels(across, down, w, h, p.water, stream, target, WORLD_RADIUS)
    print("OCEAN whole round 512x256", flush=True)
    x, hinfo = device_heat(acc, levels, target, TOL)
    xo, it_o, rr_o = oracle_heat(levels, target, TOL)
    print(f"OCEAN whole round heat: device {hinfo.cycles} iterations to {hinfo.relative_residual}, oracle {it_o} to {rr_o}")
    OC.require_solved("the sea's heat", Solution(x, hinfo.cycles, hinfo.relative_residual), TOL)
    assert rr_o < TOL and abs(hinfo.cycles - it_o) <= 2
    bal = np.abs(OC.balance_of(levels[0])).max()
    assert np.abs(x.astype(np.float64) - xo).max() <= 2 * TAU * TOL * bal
