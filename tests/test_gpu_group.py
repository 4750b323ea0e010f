"""Plans of uneven size solved together (cg_ocean_group_*): one host loop, and either the merged step
(CG_OPT_GROUP_MERGED, the default: one graph a step whose tile launches cover one depth of every
running plan's ladder) or each plan's step graph on a stream of the group's own. Each instance must stop at
the step it stops at alone, with the same answer to the bit, since the device's step gate decides per
instance whatever the host batches.

- The four-world year (twelve months on each of 512x1024, 768x1536, 1024x2048 and 1024x2048, seeds
  1-4, land 20/27/33/40%): every instance of the group's solve against its plan's own solve_all.
- A group mixing a single-instance plan with batched plans, solved twice in a row.
- A plan cannot be destroyed while a group holds it; a plan may not appear twice in a group.
"""
import numpy as np
import pytest

from cartogenesis_amd.ocean import BatchPlan, Plan, PlanGroup
from cartogenesis_amd.problem import circulation_problem, monthly_forcings

pytestmark = pytest.mark.gpu

WORLDS = [(512, 1, 0.2), (768, 2, 0.27), (1024, 3, 0.33), (1024, 4, 0.4)]


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def batched_world(acc, rows, seed, land, months=12):
    p = circulation_problem(rows, seed=seed, land_fraction=land)
    forcings = monthly_forcings(p, range(months))
    plan = BatchPlan(acc, p.levels, months)
    for i, f in enumerate(forcings):
        plan.set_forcing(i, f)
    return plan


@pytest.fixture(params=[1, 0], ids=["merged", "streams"])
def merged(acc, request):
    from cartogenesis_amd import CG_OPT_GROUP_MERGED, lib

    assert lib().cg_ocean_set_option(acc.ctx, CG_OPT_GROUP_MERGED, request.param) == 0
    yield request.param
    assert lib().cg_ocean_set_option(acc.ctx, CG_OPT_GROUP_MERGED, 1) == 0


def test_four_world_year_in_one_group_matches_each_world_alone(acc, merged):
    plans = [batched_world(acc, *w) for w in WORLDS]
    group = None
    try:
        alone = []
        for (rows, _, _), plan in zip(WORLDS, plans):
            print(f"OCEAN group: world {2 * rows}x{rows} alone", flush=True)
            infos = plan.solve_all(1e-8)
            alone.append(([i.cycles for i in infos], [i.relative_residual for i in infos],
                          [plan.read(k)[0].copy() for k in range(plan.instances)]))
        group = PlanGroup(acc, plans)
        print(f"OCEAN group: the four worlds together (merged {merged})", flush=True)
        together = group.solve(1e-8)
        for (rows, _, _), plan, infos, (cycles, res, answers) in zip(WORLDS, plans, together, alone):
            assert [i.cycles for i in infos] == cycles, rows
            assert [i.relative_residual for i in infos] == res, rows
            assert all(i.status == 0 and i.relative_residual < 1e-8 for i in infos)
            for k in range(plan.instances):
                assert same_bits(plan.read(k)[0], answers[k]), (rows, k)
    finally:
        if group:
            group.close()
        for plan in plans:
            plan.close()


def test_group_of_a_single_plan_and_batched_plans(acc, merged):
    single_p = circulation_problem(256, seed=7)
    single = Plan(acc, single_p.levels)
    batched = [batched_world(acc, 128, 5, 0.3, months=3), batched_world(acc, 384, 6, 0.35, months=4)]
    group = None
    try:
        ref = single.solve(1e-10)
        ref_x = single.read()[0].copy()
        refs = []
        for plan in batched:
            infos = plan.solve_all(1e-10)
            refs.append(([i.cycles for i in infos], [plan.read(k)[0].copy() for k in range(plan.instances)]))
        group = PlanGroup(acc, [single] + batched)
        for rep in range(3):
            if rep == 2:
                # New tile heights on the largest plan: the merged step's launches take them, so its
                # graphs are rebuilt; tiles only partition the work.
                batched[1].set_tile_rows([2] * len(batched[1].levels))
            print(f"OCEAN group: mixed, solve {rep} (merged {merged})", flush=True)
            out = group.solve(1e-10)
            assert out[0][0].cycles == ref.cycles
            assert out[0][0].relative_residual == ref.relative_residual
            assert same_bits(single.read()[0], ref_x)
            for plan, infos, (cycles, answers) in zip(batched, out[1:], refs):
                assert [i.cycles for i in infos] == cycles
                for k in range(plan.instances):
                    assert same_bits(plan.read(k)[0], answers[k])
    finally:
        if group:
            group.close()
        single.close()
        for plan in batched:
            plan.close()


def test_group_refusals(acc):
    from cartogenesis_amd import CG_ERR_ARG, CG_ERR_STATE, lib

    p = circulation_problem(128, seed=1)
    a, b = Plan(acc, p.levels), Plan(acc, p.levels)
    try:
        with pytest.raises(Exception):
            PlanGroup(acc, [a, a])
        group = PlanGroup(acc, [a, b])
        assert lib().cg_ocean_plan_destroy(a.handle) == CG_ERR_STATE
        group.close()
        assert lib().cg_ocean_group_destroy(None) == CG_ERR_ARG
    finally:
        a.close()
        b.close()
