"""The round's input and output steps on the device, against the oracle, to the bit.

- waterOn (OceanStage.kt:551-573): device mask equal to the oracle's at finer, coarser, equal and
  non-dividing map/solve ratios, including the 1024 x 2048 solve grid from a 1.5x finer map.
- curlForcing (OceanStage.kt:689-708): device binary64 forcing equal to the oracle's, with the
  500 m layer and with per-cell depth; the plan's set_stress (curl then withBalance, on the device)
  gives the same solve, bit for bit, as uploading the oracle's withBalance of the oracle's curl.
- velocities (OceanCirculation.kt:801-828): the plan's device velocities equal the oracle's
  velocities of the plan's binary32 answer, after the reference method and after refinement.
"""
import numpy as np
import pytest

import oracle
from conftest import to_oracle
from cartogenesis_amd import CG_METHOD_REFERENCE, CG_METHOD_REFINED
from cartogenesis_amd.ocean import OceanCirculation as OC, Plan
from cartogenesis_amd.problem import circulation_problem, stress

pytestmark = pytest.mark.gpu


def grid_m(across, down):
    return 40075e3 / across, 20037.5e3 / down


@pytest.mark.parametrize("ma,md,across,down", [
    (128, 64, 64, 32), (64, 32, 128, 64), (200, 100, 200, 100), (300, 150, 200, 100), (370, 190, 160, 80),
    (160, 80, 370, 190), (3072, 1536, 2048, 1024),
])
def test_water_on_matches_the_oracle(acc, ma, md, across, down):
    rng = np.random.default_rng(ma + 7 * across)
    # Blobs rather than salt, so both all-water and mixed solve cells occur.
    coarse = rng.random((md // 8 + 1, ma // 8 + 1)) < 0.35
    land = np.repeat(np.repeat(coarse, 8, 0), 8, 1)[:md, :ma].astype(np.uint8).reshape(-1)
    land |= (rng.random(ma * md) < 0.01).astype(np.uint8)
    print(f"OCEAN water_on map {ma}x{md} -> {across}x{down}", flush=True)
    got = acc.water_on(land, ma, md, across, down)
    want = oracle.water_on(land, ma, md, across, down)
    assert np.array_equal(got, want)
    assert 0 < want.sum() < want.size


@pytest.mark.parametrize("rows", [32, 256, 1024])
@pytest.mark.parametrize("per_cell_depth", [False, True])
def test_curl_forcing_matches_the_oracle(acc, rows, per_cell_depth):
    across, down = 2 * rows, rows
    w, h = grid_m(across, down)
    rng = np.random.default_rng(rows)
    east = rng.normal(0, 0.1, across * down)
    north = rng.normal(0, 0.05, across * down)
    depth = rng.uniform(200, 6000, across * down) if per_cell_depth else None
    print(f"OCEAN curl_forcing {across}x{down} depth={'cell' if per_cell_depth else '500 m'}", flush=True)
    got = acc.curl_forcing(across, down, w, h, east, north, depth)
    want = oracle.curl_forcing(across, down, w, h, east, north, depth)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("rows,month", [(128, -1), (256, 3)])
def test_set_stress_solves_as_the_uploaded_forcing_does(acc, rows, month):
    p = circulation_problem(rows, seed=2, month=month)
    across, down = p.cols, p.rows
    w, h = grid_m(across, down)
    east, north = stress(across, down, 2, month)
    depth = np.where(p.water != 0, p.depth, 1.0)
    F = oracle.curl_forcing(across, down, w, h, east, north, depth)
    f = OC.with_balance(p.levels[0], F).forcing
    print(f"OCEAN set_stress {across}x{down} month {month}", flush=True)
    plan = Plan(acc, p.levels)
    try:
        for method, tol in ((CG_METHOD_REFERENCE, 1e-5), (CG_METHOD_REFINED, 1e-10)):
            plan.set_forcing(f)
            a = plan.solve(tol, method)
            xa64, xa32 = plan.read()
            plan.set_stress(w, h, east, north, depth)
            b = plan.solve(tol, method)
            xb64, xb32 = plan.read()
            assert a.cycles == b.cycles and a.relative_residual == b.relative_residual
            assert np.array_equal(xa64.view(np.uint64), xb64.view(np.uint64))
            assert np.array_equal(xa32.view(np.uint32), xb32.view(np.uint32))
        # The synthetic balance is the curl of the synthetic stress, so the first forcing agrees too.
        assert np.array_equal(f.view(np.uint32), p.levels[0].forcing.view(np.uint32))
    finally:
        plan.close()


@pytest.mark.parametrize("rows", [128, 1024])
@pytest.mark.parametrize("method", [CG_METHOD_REFERENCE, CG_METHOD_REFINED])
def test_plan_velocities_match_the_oracle(acc, rows, method):
    p = circulation_problem(rows, seed=1)
    w, h = grid_m(p.cols, p.rows)
    plan = Plan(acc, p.levels)
    try:
        # The reference method stops at binary32's floor, near 1e-5 at 1024 x 2048.
        tol = 1e-4 if method == CG_METHOD_REFERENCE else 1e-8
        print(f"OCEAN velocities {p.cols}x{p.rows} method {method}", flush=True)
        info = plan.solve(tol, method)
        assert info.relative_residual < tol
        _, x32 = plan.read()
        e, n = plan.velocities(w, h)
        we, wn = oracle.velocities(to_oracle(p.levels[0]), x32, w, h)
        assert np.array_equal(e.view(np.uint32), we.view(np.uint32))
        assert np.array_equal(n.view(np.uint32), wn.view(np.uint32))
        assert np.abs(e).max() > 0 and np.all(e[p.water == 0] == 0)
    finally:
        plan.close()
