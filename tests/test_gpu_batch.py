"""Several right-hand sides solved together on one ladder (cg_ocean_plan_create_batch), against the
single-instance solve and the binary64 oracle.

- Copies of one forcing in every instance take exactly the single solve's steps, so each instance's
  answer is bit-identical to the single-instance plan's, with the same cycle count.
- Twelve monthly forcings of one world (the batched workload's months): every instance reaches the
  binary64 oracle's answer of its own forcing to relative L2 1e-10 at tolerance 1e-12.
- An instance with zero forcing stops before its first cycle with a zero answer while the others
  solve (the per-instance active flags).
- The full 1024 x 2048 year: twelve months to 1e-8, each residual recomputed on the host.
"""
import numpy as np
import pytest

import oracle
from conftest import to_oracle
from cartogenesis_amd import CG_METHOD_REFINED
from cartogenesis_amd.ocean import BatchPlan, Plan
from cartogenesis_amd.problem import circulation_problem, monthly_forcings

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("rows,k", [(128, 4), (256, 3)])
def test_copies_of_one_forcing_match_the_single_solve_to_the_bit(acc, rows, k):
    p = circulation_problem(rows, seed=5)
    single = Plan(acc, p.levels)
    batch = BatchPlan(acc, p.levels, k)
    try:
        info = single.solve(1e-10, CG_METHOD_REFINED)
        x64, _ = single.read()
        for i in range(k):
            batch.set_forcing(i, p.levels[0].forcing)
        print(f"OCEAN batch copies {2 * rows}x{rows} x{k}", flush=True)
        infos = batch.solve_all(1e-10)
        for i in range(k):
            y64, _ = batch.read(i)
            assert infos[i].cycles == info.cycles
            assert infos[i].relative_residual == info.relative_residual
            assert np.array_equal(y64.view(np.uint64), x64.view(np.uint64))
    finally:
        batch.close()
        single.close()


@pytest.mark.parametrize("rows", [64, 128])
def test_monthly_forcings_each_reach_the_binary64_oracle(acc, rows):
    p = circulation_problem(rows, seed=2, land_fraction=0.27)
    forcings = monthly_forcings(p)
    batch = BatchPlan(acc, p.levels, 12)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        print(f"OCEAN batch months {2 * rows}x{rows}", flush=True)
        infos = batch.solve_all(1e-12)
        ladder = [to_oracle(s) for s in p.levels]
        for i, f in enumerate(forcings):
            ladder[0] = to_oracle(p.levels[0]).with_forcing(f)
            ref, _, _ = oracle.solve64(ladder, 1e-13)
            y64, _ = batch.read(i)
            assert infos[i].relative_residual < 1e-12
            assert rel_l2(y64, ref) <= 1e-10, f"month {i}"
    finally:
        batch.close()


def test_an_instance_with_no_forcing_stops_at_once(acc):
    p = circulation_problem(128, seed=1)
    forcings = monthly_forcings(p, months=(0, 3))
    batch = BatchPlan(acc, p.levels, 3)
    try:
        batch.set_forcing(0, forcings[0])
        batch.set_forcing(1, np.zeros_like(forcings[0]))
        batch.set_forcing(2, forcings[1])
        infos = batch.solve_all(1e-9)
        assert infos[1].cycles == 0 and infos[1].relative_residual == 0.0
        assert np.all(batch.read(1)[0] == 0)
        for i in (0, 2):
            assert infos[i].cycles > 0 and infos[i].relative_residual < 1e-9
    finally:
        batch.close()


def test_a_year_at_full_size(acc):
    p = circulation_problem(1024, seed=0)
    forcings = monthly_forcings(p)
    batch = BatchPlan(acc, p.levels, 12)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        print("OCEAN batch year 2048x1024", flush=True)
        infos = batch.solve_all(1e-8)
        st = to_oracle(p.levels[0])
        for i in (0, 5, 11):
            y64, _ = batch.read(i)
            sti = st.with_forcing(forcings[i])
            res = np.abs(oracle.residual64(sti, y64)).max() / np.abs(oracle.balance_of(sti)).max()
            assert infos[i].relative_residual < 1e-8 and res < 1.01e-8, f"month {i}: {res}"
    finally:
        batch.close()


def test_autotuned_tiles_change_no_bits(acc):
    """The autotuner picks tile heights per grid by timing; the tiles only partition the work, so
    the answer, steps and residual are those of the default tiles, for one instance and for a batch."""
    p = circulation_problem(512, seed=4)
    base = Plan(acc, p.levels)
    tuned = Plan(acc, p.levels)
    try:
        a = base.solve(1e-10, CG_METHOD_REFINED)
        rows = tuned.autotune(5)
        print(f"OCEAN autotuned tile rows 1024x512: {rows}", flush=True)
        assert rows[0] in (2, 4, 6, 8, 16, 24, 32)
        b = tuned.solve(1e-10, CG_METHOD_REFINED)
        assert a.cycles == b.cycles and a.relative_residual == b.relative_residual
        assert np.array_equal(base.read()[0].view(np.uint64), tuned.read()[0].view(np.uint64))
    finally:
        base.close()
        tuned.close()
    forcings = monthly_forcings(p, months=(1, 7))
    batch = BatchPlan(acc, p.levels, 2)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        first = batch.solve_all(1e-9)
        x0 = [batch.read(i)[0] for i in range(2)]
        batch.autotune(3)
        again = batch.solve_all(1e-9)
        for i in range(2):
            assert first[i].cycles == again[i].cycles
            assert np.array_equal(batch.read(i)[0].view(np.uint64), x0[i].view(np.uint64))
    finally:
        batch.close()


@pytest.mark.parametrize("rows0", [16, 6])
def test_batched_months_repeat_their_single_solves_to_the_bit(acc, rows0):
    """Twelve different months of one 1024 x 2048 world in one batched plan: every instance's answer
    and cycle count equal its own single-instance solve's, on every one of three repeats. With many
    more workgroups than the device holds at once, a kernel that depends on what an earlier
    workgroup left in LDS, or on the order in which workgroups run, shows up here (and did, for the
    finest up leg's staging of the binary64 rows through LDS, since replaced by registers)."""
    p = circulation_problem(1024, seed=7, land_fraction=0.3)
    months = monthly_forcings(p)
    rows = [rows0] + [0] * (len(p.levels) - 1)
    single = Plan(acc, p.levels)
    ref = []
    try:
        single.set_tile_rows(rows)
        for f in months:
            single.set_forcing(f)
            ref.append((single.solve(1e-8, CG_METHOD_REFINED).cycles, single.read()[0]))
    finally:
        single.close()
    for rep in range(3):
        batch = BatchPlan(acc, p.levels, len(months))
        try:
            batch.set_tile_rows(rows)
            for i, f in enumerate(months):
                batch.set_forcing(i, f)
            print(f"OCEAN batched months repeat {rep} tile rows {rows0}", flush=True)
            infos = batch.solve_all(1e-8)
            for i in range(len(months)):
                assert infos[i].cycles == ref[i][0], (rep, i)
                assert np.array_equal(batch.read(i)[0], ref[i][1]), (rep, i)
        finally:
            batch.close()


def test_kernel_timer_leaves_the_answers_and_the_next_solve_unchanged(acc):
    """cg_ocean_plan_time_kernel on a batched plan, the batched kind (6: the fused finest up leg and
    outer step over every instance in one launch) included: positive times, the answers read after
    timing are those read before to the bit, and a solve after timing repeats the one before."""
    import cartogenesis_amd as cg
    p = circulation_problem(256, seed=3)
    forcings = monthly_forcings(p, months=range(3))
    batch = BatchPlan(acc, p.levels, 3)
    try:
        for i, f in enumerate(forcings):
            batch.set_forcing(i, f)
        print("OCEAN batch timer 512x256 x3", flush=True)
        infos = batch.solve_all(1e-8)
        before = [batch.read(i)[0].copy() for i in range(3)]
        lib = cg.lib()
        for kind in (0, 3, 5, 6):
            us = lib.cg_ocean_plan_time_kernel(batch.handle, kind, 5)
            assert us > 0, (kind, us)
        assert lib.cg_ocean_plan_time_kernel(batch.handle, 7, 5) < 0  # no such kind
        for i in range(3):
            assert np.array_equal(batch.read(i)[0].view(np.uint64), before[i].view(np.uint64))
        again = batch.solve_all(1e-8)
        for i in range(3):
            assert again[i].cycles == infos[i].cycles
            assert np.array_equal(batch.read(i)[0].view(np.uint64), before[i].view(np.uint64))
    finally:
        batch.close()
