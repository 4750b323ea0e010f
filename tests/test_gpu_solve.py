"""Whole solves on the device against the oracle.

- The reference method (binary32 values, binary64 residual, V(2,2), coarsest to 0.1 tol) is held
  to the oracle's restatement of OceanCirculation.solve (OceanCirculation.kt:473-493) to the bit:
  same values, same cycle count, same residual.
- The refined method (binary32 cycles inside binary64 refinement) is held to the binary64 oracle
  (the exact solution of the same operator, pinned against scipy direct solves in
  tests/test_oracle.py): relative L2 <= 1e-10 at tolerance 1e-12.
- At the full 1024 x 2048 size, where the oracle is slow, by properties that hold at any size: the
  residual recomputed on the CPU, exact linearity under a power-of-two scaling of the forcing, and
  bit-identical repeats.
"""
import numpy as np
import pytest

import oracle
from conftest import stommel_analytic, stommel_basin, to_oracle
from cartogenesis_amd import CG_METHOD_REFERENCE, CG_METHOD_REFINED, CG_OPT_USE_GRAPH
from cartogenesis_amd.ocean import OceanCirculation as OC, Plan
from cartogenesis_amd.problem import circulation_problem

pytestmark = pytest.mark.gpu

PARITY_L2 = 1e-10  # relative L2 against the binary64 oracle


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def oracle_residual(levels, x64):
    st = to_oracle(levels[0])
    return float(np.abs(oracle.residual64(st, x64)).max() / np.abs(oracle.balance_of(st)).max())


@pytest.mark.parametrize("rows,tol", [(64, 1e-3), (128, 1e-5), (256, 1e-3), (256, 1e-5)])
def test_reference_method_is_bit_identical_to_the_processor_solve(acc, rows, tol):
    p = circulation_problem(rows, seed=0)
    sol = acc.solve_levels(p.levels, None, tol, True, CG_METHOD_REFERENCE)
    x, cycles, rr = oracle.solve([to_oracle(s) for s in p.levels], np.zeros(p.levels[0].cells, np.float32), tol)
    print(f"OCEAN reference method {rows}x{2 * rows} tol {tol}: {sol.cycles} cycles, residual {sol.relative_residual}")
    assert sol.cycles == cycles
    assert sol.relative_residual == rr
    assert np.array_equal(sol.values.view(np.uint32), x.view(np.uint32))
    OC.require_solved("the circulation", sol, tol)


def test_reference_method_from_a_warm_start(acc):
    p = circulation_problem(128, seed=3, land_fraction=0.4)
    rng = np.random.default_rng(1)
    start = (rng.normal(size=p.levels[0].cells) * 100).astype(np.float32)
    sol = acc.solve_levels(p.levels, start, 1e-4, True, CG_METHOD_REFERENCE)
    x, cycles, rr = oracle.solve([to_oracle(s) for s in p.levels], start, 1e-4)
    assert sol.cycles == cycles and sol.relative_residual == rr
    assert np.array_equal(sol.values.view(np.uint32), x.view(np.uint32))


@pytest.mark.parametrize("rows,seed,land", [(64, 0, 0.3), (128, 0, 0.3), (256, 0, 0.3), (128, 7, 0.2), (128, 11, 0.4)])
def test_refined_method_matches_the_binary64_oracle(acc, rows, seed, land):
    p = circulation_problem(rows, seed=seed, land_fraction=land)
    sol = acc.solve_levels(p.levels, None, 1e-12, True, CG_METHOD_REFINED)
    ref64, _, ref_rr = oracle.solve64([to_oracle(s) for s in p.levels], 1e-13)
    err = rel_l2(sol.values64, ref64)
    print(f"OCEAN refined {rows}x{2 * rows} seed {seed}: {sol.cycles} cycles, residual {sol.relative_residual}, "
          f"rel L2 vs binary64 oracle {err}")
    assert ref_rr < 1e-13
    assert sol.relative_residual < 1e-12
    assert oracle_residual(p.levels, sol.values64) < 1e-12
    assert err <= PARITY_L2


def test_refined_graph_and_eager_agree_to_the_bit(acc):
    p = circulation_problem(128, seed=2)
    acc.set_option(CG_OPT_USE_GRAPH, 0)
    try:
        eager = acc.solve_levels(p.levels, None, 1e-10, True, CG_METHOD_REFINED)
    finally:
        acc.set_option(CG_OPT_USE_GRAPH, 1)
    graphed = acc.solve_levels(p.levels, None, 1e-10, True, CG_METHOD_REFINED)
    assert eager.cycles == graphed.cycles
    assert np.array_equal(eager.values64.view(np.uint64), graphed.values64.view(np.uint64))


def test_full_size_refined_solve_properties(acc):
    """1024 x 2048 (the benchmark's workload): converged by the CPU's own residual, linear to the
    bit under a forcing doubled, and bit-identical on repeat; replays of the step graph between
    eager work agree with each other."""
    p = circulation_problem(1024, seed=0)
    plan = Plan(acc, p.levels)
    try:
        info = plan.solve(1e-8, CG_METHOD_REFINED)
        x1, _ = plan.read()
        assert info.relative_residual < 1e-8
        assert oracle_residual(p.levels, x1) < 1e-8
        info2 = plan.solve(1e-8, CG_METHOD_REFINED)
        x2, _ = plan.read()
        assert info2.cycles == info.cycles
        assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
        plan.set_forcing(p.levels[0].forcing * np.float32(2.0))
        info3 = plan.solve(1e-8, CG_METHOD_REFINED)
        x3, _ = plan.read()
        assert info3.cycles == info.cycles
        assert np.array_equal((2.0 * x1).view(np.uint64), x3.view(np.uint64))
        print(f"OCEAN full size: {info.cycles} cycles to {info.relative_residual}")
    finally:
        plan.close()


@pytest.mark.parametrize("aspect", [1.0, 0.5])
def test_stommel_basin_on_the_device(acc, aspect):
    """OceanCirculationTest.kt:89-106: the discrete basin within 1% of Stommel's analytic one,
    solved on the device at the reference test's tolerance (1e-5)."""
    levels, across, down, dx, dy = stommel_basin(6_250.0, aspect)
    for method in (CG_METHOD_REFERENCE, CG_METHOD_REFINED):
        sol = acc.solve_levels(levels, None, 1e-5, True, method)
        stream = sol.values.reshape(down, across)
        cols = np.arange(1, across - 1)
        rows = np.arange(down)
        X, Y = np.meshgrid(cols * dx, (down - 0.5 - rows) * dy)
        exact = stommel_analytic(across, down, dx, dy, -1e-12, X, Y)
        worst = np.abs(stream[:, 1:-1] - exact).max()
        largest = np.abs(exact).max()
        print(f"OCEAN Stommel basin on device, aspect {aspect}, method {method}: {worst / largest}")
        assert worst / largest < 0.01


@pytest.mark.parametrize("rows", [64, 256, 512])
def test_fused_and_unfused_cycles_reach_the_oracle(acc, rows):
    """The fused cycle (one down and one up launch per grid, the bottom grids in one workgroup) and
    the unfused one (separate smoothing, residual, restriction and prolongation kernels) both reach
    the binary64 oracle; with two passes on every grid (the unfused cycle's only schedule) they
    differ only in the coarsest grid's summation order, so their step counts agree."""
    from cartogenesis_amd import CG_OPT_LEVEL_PASSES
    p = circulation_problem(rows, seed=5)
    ref64 = oracle.solve64([to_oracle(s) for s in p.levels], 1e-13)[0]
    out = {}
    default = acc.get_option(CG_OPT_LEVEL_PASSES)
    for fused in (1, 0):
        acc.set_option(7, fused)
        acc.set_option(CG_OPT_LEVEL_PASSES, 0x22222222)
        try:
            sol = acc.solve_levels(p.levels, None, 1e-11, True, CG_METHOD_REFINED)
        finally:
            acc.set_option(7, 1)
            acc.set_option(CG_OPT_LEVEL_PASSES, default)
        out[fused] = sol
        assert sol.relative_residual < 1e-11
        assert rel_l2(sol.values64, ref64) <= PARITY_L2
    assert abs(out[1].cycles - out[0].cycles) <= 1


@pytest.mark.parametrize("rows,graph", [(128, 1), (256, 0), (1024, 1)])
def test_outer_step_fused_into_the_up_leg_gives_the_same_bits(acc, rows, graph):
    """The finest up leg with the binary64 outer step fused in (CG_OPT_FUSE_OUTER, the default)
    computes exactly what the up leg and k_outer_residual compute one after the other: the same
    steps, residual and binary64 answer to the bit, with and without graph replay."""
    from cartogenesis_amd import CG_OPT_FUSE_OUTER
    p = circulation_problem(rows, seed=9)
    out = {}
    acc.set_option(CG_OPT_USE_GRAPH, graph)
    try:
        for fuse in (1, 0):
            acc.set_option(CG_OPT_FUSE_OUTER, fuse)
            plan = Plan(acc, p.levels)
            try:
                info = plan.solve(1e-10, CG_METHOD_REFINED)
                out[fuse] = (info.cycles, info.relative_residual, plan.read()[0])
            finally:
                plan.close()
    finally:
        acc.set_option(CG_OPT_FUSE_OUTER, 1)
        acc.set_option(CG_OPT_USE_GRAPH, 1)
    print(f"OCEAN fused outer {2 * rows}x{rows}: {out[1][0]} cycles, residual {out[1][1]}")
    assert out[1][0] == out[0][0] and out[1][1] == out[0][1]
    assert np.array_equal(out[1][2].view(np.uint64), out[0][2].view(np.uint64))


@pytest.mark.parametrize("rows,seed", [(64, 0), (256, 0), (128, 4)])
def test_fgmres_method_matches_the_binary64_oracle(acc, rows, seed):
    """FGMRES(m) right-preconditioned by the binary32 cycle, inside binary64 refinement: the same
    fixed point as the refined method, in fewer V-cycles."""
    from cartogenesis_amd import CG_METHOD_FGMRES
    p = circulation_problem(rows, seed=seed)
    ref64 = oracle.solve64([to_oracle(s) for s in p.levels], 1e-13)[0]
    sol = acc.solve_levels(p.levels, None, 1e-12, True, CG_METHOD_FGMRES)
    plain = acc.solve_levels(p.levels, None, 1e-12, True, CG_METHOD_REFINED)
    err = rel_l2(sol.values64, ref64)
    print(f"OCEAN fgmres {rows}x{2 * rows} seed {seed}: {sol.cycles} V-cycles in {sol.refinements} refinements "
          f"(plain {plain.cycles}), residual {sol.relative_residual}, rel L2 {err}")
    assert sol.relative_residual < 1e-12
    assert oracle_residual(p.levels, sol.values64) < 1e-12
    assert err <= PARITY_L2


def test_fgmres_full_size_converges_and_repeats(acc):
    from cartogenesis_amd import CG_METHOD_FGMRES
    p = circulation_problem(1024, seed=0)
    plan = Plan(acc, p.levels)
    try:
        info = plan.solve(1e-8, CG_METHOD_FGMRES)
        x1, _ = plan.read()
        assert info.relative_residual < 1e-8
        assert oracle_residual(p.levels, x1) < 1e-8
        info2 = plan.solve(1e-8, CG_METHOD_FGMRES)
        x2, _ = plan.read()
        assert info2.cycles == info.cycles
        assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
        plain = plan.solve(1e-8, CG_METHOD_REFINED)
        print(f"OCEAN full size fgmres: {info.cycles} V-cycles, {info.refinements} refinements; plain {plain.cycles}")
    finally:
        plan.close()


@pytest.mark.parametrize("schedule", [0x22222462, 0x22224662, 0x22222242])
def test_level_pass_schedules_reach_the_oracle(acc, schedule):
    """More passes on the coarser grids of the binary32 cycle change its convergence rate, never
    the fixed point."""
    from cartogenesis_amd import CG_OPT_LEVEL_PASSES
    p = circulation_problem(256, seed=0)
    ref64 = oracle.solve64([to_oracle(s) for s in p.levels], 1e-13)[0]
    default = acc.get_option(CG_OPT_LEVEL_PASSES)
    acc.set_option(CG_OPT_LEVEL_PASSES, schedule)
    try:
        sol = acc.solve_levels(p.levels, None, 1e-12, True, CG_METHOD_REFINED)
        acc.set_option(CG_OPT_LEVEL_PASSES, 0x22222222)
        base = acc.solve_levels(p.levels, None, 1e-12, True, CG_METHOD_REFINED)
    finally:
        acc.set_option(CG_OPT_LEVEL_PASSES, default)
    print(f"OCEAN schedule {schedule:#x}: {sol.cycles} cycles (V(2,2): {base.cycles})")
    assert sol.relative_residual < 1e-12
    assert rel_l2(sol.values64, ref64) <= PARITY_L2
    assert sol.cycles <= base.cycles


@pytest.mark.parametrize("rows,tol", [(64, 1e-6), (32, 1e-7)])
def test_krylov_rescue_matches_the_processor(acc, rows, tol):
    """solveWithKrylovRescue (OceanCirculation.kt:507-528): where binary32 cycles stall on their
    rounding floor (200 cycles), BiCGSTAB preconditioned by the reference cycle takes over. The
    cycles are bit-identical; BiCGSTAB's dot products are summed in another order (block partials),
    so its values agree to binary32 rounding, with the same iteration count."""
    from cartogenesis_amd import CG_METHOD_RESCUE
    p = circulation_problem(rows, seed=0)
    sol = acc.solve_levels(p.levels, None, tol, True, CG_METHOD_RESCUE)
    x, cycles, rr = oracle.solve([to_oracle(s) for s in p.levels], np.zeros(p.levels[0].cells, np.float32), tol,
                                 mode=oracle.RESCUE)
    print(f"OCEAN rescue {rows}x{2 * rows} tol {tol}: device {sol.cycles} / {sol.relative_residual}, "
          f"processor {cycles} / {rr}")
    assert cycles > 200, "the case must exercise the rescue"
    assert sol.cycles == cycles
    assert sol.relative_residual < tol and rr < tol
    assert sol.relative_residual == pytest.approx(rr, rel=1e-3)
    assert np.allclose(sol.values, x, rtol=1e-5, atol=1e-6 * np.abs(x).max())
    OC.require_solved("the circulation", sol, tol)


@pytest.mark.parametrize("rows", [128, 256])
def test_direct_bottom_and_cycled_bottom_reach_the_oracle(acc, rows):
    """CG_OPT_DIRECT_BOTTOM: the binary32 cycle ends at its 32 x 64 grid solved by the dense
    inverse (one chip-wide launch), or cycles that grid and the coarser ones in one workgroup. Both
    preconditioners leave the binary64 fixed point unchanged: each reaches the binary64 oracle, and
    the exact bottom solve needs no more cycles than the cycled one (one more allowed for the
    rounding of the last step), fused and unfused alike. The fused cycle runs two passes on every
    grid here, the unfused cycle's only schedule, so the two cycle counts are comparable."""
    from cartogenesis_amd import CG_OPT_DIRECT_BOTTOM, CG_OPT_FUSED, CG_OPT_LEVEL_PASSES
    p = circulation_problem(rows, seed=6)
    ref64 = oracle.solve64([to_oracle(s) for s in p.levels], 1e-13)[0]
    cycles = {}
    default = acc.get_option(CG_OPT_LEVEL_PASSES)
    acc.set_option(CG_OPT_LEVEL_PASSES, 0x22222222)
    try:
        for fused in (1, 0):
            for direct in (1, 0):
                acc.set_option(CG_OPT_FUSED, fused)
                acc.set_option(CG_OPT_DIRECT_BOTTOM, direct)
                print(f"OCEAN direct bottom {direct} fused {fused} {2 * rows}x{rows}", flush=True)
                sol = acc.solve_levels(p.levels, None, 1e-11, True, CG_METHOD_REFINED)
                assert sol.relative_residual < 1e-11
                assert rel_l2(sol.values64, ref64) <= PARITY_L2
                cycles[fused, direct] = sol.cycles
    finally:
        acc.set_option(CG_OPT_FUSED, 1)
        acc.set_option(CG_OPT_DIRECT_BOTTOM, 1)
        acc.set_option(CG_OPT_LEVEL_PASSES, default)
    for fused in (1, 0):
        assert cycles[fused, 1] <= cycles[fused, 0] + 1
    # The fused and unfused cycles share the direct bottom launch.
    assert abs(cycles[1, 1] - cycles[0, 1]) <= 1


def test_option_changes_rebuild_the_step_graphs(acc):
    """A plan's step graphs are captured under the context's settings; changing one
    (CG_OPT_DIRECT_BOTTOM here) between solves of the same plan rebuilds them: the solve with the
    setting restored repeats the first to the bit, and the other setting's answer, from a
    different preconditioner, agrees with it to the tolerance."""
    from cartogenesis_amd import CG_OPT_DIRECT_BOTTOM
    p = circulation_problem(128, seed=7)
    plan = Plan(acc, p.levels)
    try:
        print("OCEAN graph rebuild on option change 256x128", flush=True)
        first = plan.solve(1e-11, CG_METHOD_REFINED)
        x_first, _ = plan.read()
        acc.set_option(CG_OPT_DIRECT_BOTTOM, 0)
        other = plan.solve(1e-11, CG_METHOD_REFINED)
        x_other, _ = plan.read()
        acc.set_option(CG_OPT_DIRECT_BOTTOM, 1)
        again = plan.solve(1e-11, CG_METHOD_REFINED)
        x_again, _ = plan.read()
    finally:
        acc.set_option(CG_OPT_DIRECT_BOTTOM, 1)
        plan.close()
    assert again.cycles == first.cycles
    assert np.array_equal(x_again, x_first)
    assert other.relative_residual < 1e-11
    assert rel_l2(x_other, x_first) <= 1e-9


def test_error_bound_decides_where_the_residual_alone_stops_short(acc):
    """A month whose stress band vanishes (cos(2 pi 3 / 12) = 0: the forcing is the noise alone) has a
    residual under tol long before its answer is within 1e-10 of the exact one. With the error bound
    (CG_OPT_ERROR_PICO, default 1e-10) the solve runs on until its estimated error is under it, and
    the true error (against the binary64 oracle) is; with the residual alone (0) it stops earlier."""
    from cartogenesis_amd import CG_OPT_ERROR_PICO
    from cartogenesis_amd.problem import monthly_forcings
    p = circulation_problem(512, seed=1, land_fraction=0.2)
    f3 = monthly_forcings(p, months=(3,))[0]
    ladder = [to_oracle(s) for s in p.levels]
    ladder[0] = ladder[0].with_forcing(f3)
    ref = oracle.solve64(ladder, 1e-13)[0]
    plan = Plan(acc, p.levels)
    try:
        plan.set_forcing(f3)
        bounded = plan.solve(1e-8, CG_METHOD_REFINED)
        xb = plan.read()[0]
        acc.set_option(CG_OPT_ERROR_PICO, 0)
        try:
            plain = plan.solve(1e-8, CG_METHOD_REFINED)
            xp = plan.read()[0]
        finally:
            acc.set_option(CG_OPT_ERROR_PICO, 100)
    finally:
        plan.close()
    eb, ep = rel_l2(xb, ref), rel_l2(xp, ref)
    print(f"OCEAN error bound 1024x512 month 3: {bounded.cycles} cycles, rel L2 {eb:.3e}; residual alone "
          f"{plain.cycles} cycles, rel L2 {ep:.3e}")
    assert bounded.relative_residual < 1e-8 and plain.relative_residual < 1e-8
    assert eb <= PARITY_L2
    assert plain.cycles < bounded.cycles and ep > eb


def test_warm_start_at_its_answer_runs_no_step(acc):
    p = circulation_problem(256, seed=2)
    plan = Plan(acc, p.levels)
    try:
        first = plan.solve(1e-9, CG_METHOD_REFINED)
        x1 = plan.read()[0]
        again = plan.solve(1e-9, CG_METHOD_REFINED, warm=True)
        x2 = plan.read()[0]
    finally:
        plan.close()
    assert first.cycles > 0 and again.cycles == 0
    assert again.relative_residual < 1e-9
    assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
