"""Stencil residency and plan lifetime (cg_ocean_stencil_release, cg_ocean_plan_destroy).

The accelerator keeps uploaded stencils in an LRU of MOST_RESIDENT_STENCILS, as GpuOcean's holding
does (GpuOcean.kt:109-115), and releases the oldest handle when a new stencil needs the room. A
plan built on a released stencil must keep solving on that stencil's memory until the plan is
destroyed:

- releasing every handle of a live plan's ladder leaves its solves bit-identical to the solves
  before the release, and the released handles are refused afterwards;
- an LRU as long as one ladder, evicting a batched plan's ladder while a second world's plan is
  built, changes none of the first plan's answers or cycle counts.
"""
import ctypes

import numpy as np
import pytest

from cartogenesis_amd import CG_ERR_ARG, CG_METHOD_REFINED, lib
from cartogenesis_amd.ocean import BatchPlan, HipOcean, Plan
from cartogenesis_amd.problem import circulation_problem, monthly_forcings

pytestmark = pytest.mark.gpu


def test_a_plan_outlives_the_release_of_its_stencils(acc):
    prob = circulation_problem(256, seed=11)
    fresh = HipOcean.create_or_null(0).accelerator
    try:
        plan = Plan(fresh, prob.levels)
        try:
            print("OCEAN residency: plan over released stencils 512x256", flush=True)
            before = plan.solve(1e-10, CG_METHOD_REFINED)
            x_before, _ = plan.read()
            handles = [h for (_, h) in fresh._resident.values()]
            assert len(handles) == len(prob.levels)
            for h in handles:
                assert lib().cg_ocean_stencil_release(fresh.ctx, ctypes.c_uint64(h)) == 0
            fresh._resident.clear()
            after = plan.solve(1e-10, CG_METHOD_REFINED)
            x_after, _ = plan.read()
            assert after.cycles == before.cycles
            assert np.array_equal(x_after, x_before)
            assert lib().cg_ocean_stencil_release(fresh.ctx, ctypes.c_uint64(handles[0])) == CG_ERR_ARG
        finally:
            plan.close()
    finally:
        fresh.close()


def test_lru_eviction_leaves_a_batched_plan_solving(acc):
    first = circulation_problem(256, seed=12, land_fraction=0.25)
    second = circulation_problem(256, seed=13, land_fraction=0.35)
    months = monthly_forcings(first, range(3))
    small = HipOcean.create_or_null(0).accelerator
    small.MOST_RESIDENT_STENCILS = len(first.levels)
    assert len(second.levels) == len(first.levels)
    try:
        plan = BatchPlan(small, first.levels, len(months))
        other = None
        try:
            for i, f in enumerate(months):
                plan.set_forcing(i, f)
            print("OCEAN residency: batched 512x256 x3, then a second 512x256 world through an LRU of one ladder", flush=True)
            infos = plan.solve_all(1e-10)
            xs = [plan.read(i)[0] for i in range(len(months))]
            other = Plan(small, second.levels)
            assert all(id(s.east_weight) not in small._resident for s in first.levels)
            other.solve(1e-10, CG_METHOD_REFINED)
            again = plan.solve_all(1e-10)
            for i in range(len(months)):
                assert again[i].cycles == infos[i].cycles > 0
                assert np.array_equal(plan.read(i)[0], xs[i])
        finally:
            plan.close()
            if other is not None:
                other.close()
    finally:
        small.close()
