"""The workloads at their full sizes, each held to the binary64 oracle.

- 1024 x 2048, seed 0, tolerance 1e-8 (the benchmark's solve): the refined method's answer within
  relative L2 1e-10 of the oracle's binary64 solution (solve64 to 1e-12), its residual recomputed on
  the host under 1e-8.
- The twelve months of that world in one batched plan: three months against the oracle the same way.
- Four worlds of uneven size (512 x 1024, 768 x 1536, 1024 x 2048 twice; seeds 1-4, land 20, 27,
  33 and 40%), twelve months each in one batched plan per world: every instance's residual
  recomputed on the host under 1e-8, its own month and month 3 (noise alone) within 1e-10 of the oracle.
- 2048 x 4096: one instance against the oracle, exact linearity and bit-identical repeats; sixteen
  monthly forcings in one batched plan, each residual recomputed, two instances bit-identical to
  their single-instance solves.
"""
import numpy as np
import pytest

import oracle
from conftest import to_oracle
from cartogenesis_amd import CG_METHOD_REFINED
from cartogenesis_amd.ocean import BatchPlan, Plan
from cartogenesis_amd.problem import circulation_problem, monthly_forcings

pytestmark = pytest.mark.gpu

PARITY_L2 = 1e-10
TOL = 1e-8


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def host_residual(st, x64):
    return float(np.abs(oracle.residual64(st, x64)).max() / np.abs(oracle.balance_of(st)).max())


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    oracle.set_threads(16)
    yield
    oracle.set_threads(1)


def test_headline_solve_matches_the_binary64_oracle(acc):
    p = circulation_problem(1024, seed=0, land_fraction=0.3)
    plan = Plan(acc, p.levels)
    try:
        print("OCEAN headline 2048x1024 seed 0 tol 1e-8", flush=True)
        info = plan.solve(TOL, CG_METHOD_REFINED)
        x64, _ = plan.read()
    finally:
        plan.close()
    ladder = [to_oracle(s) for s in p.levels]
    ref, _, ref_rr = oracle.solve64(ladder, 1e-12)
    err = rel_l2(x64, ref)
    res = host_residual(ladder[0], x64)
    print(f"OCEAN headline: {info.cycles} cycles, residual {info.relative_residual:.3e} (host {res:.3e}), "
          f"rel L2 vs binary64 oracle {err:.3e}")
    assert ref_rr < 1e-12
    assert info.relative_residual < TOL and res < TOL
    assert err <= PARITY_L2


def test_full_size_year_months_match_the_binary64_oracle(acc):
    p = circulation_problem(1024, seed=0)
    forcings = monthly_forcings(p)
    batch = BatchPlan(acc, p.levels, 12)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        print("OCEAN batched year 2048x1024 against the oracle", flush=True)
        infos = batch.solve_all(TOL)
        ladder = [to_oracle(s) for s in p.levels]
        for i in (0, 5, 11):
            y64, _ = batch.read(i)
            ladder[0] = to_oracle(p.levels[0]).with_forcing(forcings[i])
            ref, _, _ = oracle.solve64(ladder, 1e-12)
            err = rel_l2(y64, ref)
            print(f"OCEAN year month {i}: {infos[i].cycles} cycles, residual {infos[i].relative_residual:.3e}, "
                  f"rel L2 {err:.3e}")
            assert infos[i].relative_residual < TOL
            assert host_residual(ladder[0], y64) < TOL
            assert err <= PARITY_L2, f"month {i}"
    finally:
        batch.close()


FOUR_WORLDS = [(512, 1, 0.20), (768, 2, 0.27), (1024, 3, 0.33), (1024, 4, 0.40)]


@pytest.mark.parametrize("rows,seed,land", FOUR_WORLDS)
def test_four_world_year(acc, rows, seed, land):
    p = circulation_problem(rows, seed=seed, land_fraction=land)
    forcings = monthly_forcings(p)
    batch = BatchPlan(acc, p.levels, 12)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        print(f"OCEAN four-world year {2 * rows}x{rows} seed {seed} land {land}: {len(p.levels)} grids, "
              f"coarsest {p.levels[-1].cells_across}x{p.levels[-1].cells_down}", flush=True)
        infos = batch.solve_all(TOL)
        st = to_oracle(p.levels[0])
        answers = []
        for i, f in enumerate(forcings):
            y64, _ = batch.read(i)
            answers.append(y64)
            res = host_residual(st.with_forcing(f), y64)
            assert infos[i].relative_residual < TOL and res < 1.01 * TOL, f"month {i}: {res}"
    finally:
        batch.close()
    # The world's own month and month 3, whose band amplitude cos(pi / 2) vanishes: a forcing of
    # noise alone, whose residual under tol leaves the largest error (the error bound must hold).
    ladder = [to_oracle(s) for s in p.levels]
    for month in sorted({seed % 12, 3}):
        ladder[0] = st.with_forcing(forcings[month])
        ref, _, _ = oracle.solve64(ladder, 1e-12)
        err = rel_l2(answers[month], ref)
        print(f"OCEAN four-world {2 * rows}x{rows}: cycles {min(i.cycles for i in infos)}-"
              f"{max(i.cycles for i in infos)}, month {month}: {infos[month].cycles} cycles, rel L2 {err:.3e}")
        assert err <= PARITY_L2, f"month {month}"


def test_largest_grid_single_instance(acc):
    p = circulation_problem(2048, seed=8, land_fraction=0.3)
    plan = Plan(acc, p.levels)
    try:
        print("OCEAN 4096x2048 single instance", flush=True)
        info = plan.solve(TOL, CG_METHOD_REFINED)
        x1, _ = plan.read()
        info2 = plan.solve(TOL, CG_METHOD_REFINED)
        x2, _ = plan.read()
        plan.set_forcing(p.levels[0].forcing * np.float32(2.0))
        info3 = plan.solve(TOL, CG_METHOD_REFINED)
        x3, _ = plan.read()
    finally:
        plan.close()
    ladder = [to_oracle(s) for s in p.levels]
    assert info.relative_residual < TOL and host_residual(ladder[0], x1) < TOL
    assert info2.cycles == info.cycles and np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
    assert info3.cycles == info.cycles and np.array_equal((2.0 * x1).view(np.uint64), x3.view(np.uint64))
    ref, _, _ = oracle.solve64(ladder, 1e-12)
    err = rel_l2(x1, ref)
    print(f"OCEAN 4096x2048: {info.cycles} cycles, residual {info.relative_residual:.3e}, rel L2 {err:.3e}")
    assert err <= PARITY_L2


def test_largest_grid_sixteen_instances(acc):
    p = circulation_problem(2048, seed=8, land_fraction=0.3)
    months = monthly_forcings(p)
    forcings = [months[i % 12] for i in range(16)]
    singles = {}
    plan = Plan(acc, p.levels)
    try:
        for i in (3, 14):
            plan.set_forcing(forcings[i])
            singles[i] = (plan.solve(TOL, CG_METHOD_REFINED).cycles, plan.read()[0])
    finally:
        plan.close()
    batch = BatchPlan(acc, p.levels, 16)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        print("OCEAN 4096x2048 sixteen instances", flush=True)
        infos = batch.solve_all(TOL)
        st = to_oracle(p.levels[0])
        for i, f in enumerate(forcings):
            y64, _ = batch.read(i)
            assert infos[i].relative_residual < TOL and host_residual(st.with_forcing(f), y64) < 1.01 * TOL, i
            if i in singles:
                assert infos[i].cycles == singles[i][0]
                assert np.array_equal(y64.view(np.uint64), singles[i][1].view(np.uint64)), i
    finally:
        batch.close()
