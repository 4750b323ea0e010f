"""The C ABI and the host-side code, without a GPU: the library loads and exports every entry point
include/cgocean.h declares; the host builders agree with the oracle to the bit; the synthetic
generator keeps its promises; the Python mirror keeps the reference's failure semantics."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
import cartogenesis_amd as cg
from cartogenesis_amd.ocean import HipOcean, OceanCirculation as OC, OceanSolveFailure, Solution
from cartogenesis_amd.problem import beta_rows, bottom_drag, circulation_problem, generate
from test_oracle import beta_rows_numpy, drag_numpy

INCLUDE = Path(__file__).resolve().parents[1] / "include"
HEADER = INCLUDE / "cgocean.h"


def declared_functions():
    """Every entry point include/*.h declares."""
    text = "".join(h.read_text() for h in sorted(INCLUDE.glob("*.h")))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_\w+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = cg.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), f"{name} is declared in include/ but not exported"
    assert set(names) == set(cg.SIGNATURES), "the ctypes table and the header disagree"


def test_jni_shim_and_kotlin_class_agree_with_each_other_and_the_header():
    """The JNI binding cannot be compiled here (no JDK): its C natives must be exactly the Kotlin
    class's `external fun`s, and every library call it makes must be declared in the header."""
    jni = Path(__file__).resolve().parents[1] / "cartogenesis_amd" / "jni"
    c_text = (jni / "cgocean_jni.c").read_text()
    kt_text = (jni / "HipOcean.kt").read_text()
    natives = set(re.findall(r"Java_com_cartogenesis_desktop_HipOcean_(\w+)\(", c_text))
    externals = set(re.findall(r"external fun (\w+)\(", kt_text))
    assert natives and natives == externals
    assert "package com.cartogenesis.desktop" in kt_text and "class HipOcean" in kt_text
    body = re.sub(r"/\*.*?\*/", "", c_text, flags=re.S)
    called = set(re.findall(r"\b(cg_\w+)\s*\(", body)) - {"cg_solution"}
    assert called and called <= set(declared_functions())


def test_open_without_a_device_declines_with_a_reason():
    res = HipOcean.create_or_null(0)
    if res.accelerator is not None:
        res.accelerator.close()
        pytest.skip("a GPU is present")
    assert res.unavailable_because and "device" in res.unavailable_because


def test_null_arguments_are_rejected_not_dereferenced():
    lib = cg.lib()
    assert lib.cg_ocean_set_option(None, 1, 0) == cg.CG_ERR_ARG
    assert lib.cg_ocean_plan_destroy(None) == cg.CG_ERR_ARG
    assert lib.cg_ocean_last_error(None) == b"no context"
    assert lib.cg_beta_rows(0, 40075.0, np.zeros(1)) == cg.CG_ERR_ARG


def test_constants_and_beta_match_the_numpy_restatement():
    assert bottom_drag() == drag_numpy()
    for down in (8, 256, 1024, 960):
        assert np.array_equal(beta_rows(down), beta_rows_numpy(down))


@pytest.mark.parametrize("rows,seed", [(64, 0), (128, 9)])
def test_host_stencil_and_ladder_match_the_oracle_to_the_bit(rows, seed):
    p = circulation_problem(rows, seed=seed)
    o = oracle.stencil(p.cols, p.rows, p.dx, p.dy, p.water, beta_rows_numpy(p.rows), drag_numpy(), p.balance)
    f = p.levels[0]
    for a, b in ((f.east_weight, o.east), (f.west_weight, o.west), (f.north_weight, o.north),
                 (f.south_weight, o.south), (f.forcing, o.forcing)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(f.centre_weight.view(np.uint64), o.centre.view(np.uint64))
    shapes = oracle.level_shapes(p.cols, p.rows, p.dx, p.dy)
    assert [(s.cells_across, s.cells_down) for s in p.levels] == [(a, d) for a, d, _, _ in shapes]
    water = p.water
    a, d = p.cols, p.rows
    for s in p.levels[1:]:
        water = oracle.coarse_water(water, a, d, s.cells_across, s.cells_down, True)
        a, d = s.cells_across, s.cells_down
        assert np.array_equal(s.is_water, water)
        assert not s.forcing.any(), "coarse circulation stencils carry no forcing (OS:434)"


def test_generator_promises():
    cols, rows = 512, 256
    water, depth, balance = generate(cols, rows, 0, 0.3)
    assert water.sum() == cols * rows - round(0.3 * cols * rows)
    wet = water != 0
    assert depth[wet].min() >= 200.0 and depth[wet].max() <= 6000.0 and not depth[~wet].any()
    assert not balance[~wet].any() and np.isfinite(balance).all() and np.abs(balance[wet]).max() > 0
    w2, d2, b2 = generate(cols, rows, 0, 0.3)
    assert np.array_equal(water, w2) and np.array_equal(balance, b2)
    w3, _, b3 = generate(cols, rows, 1, 0.3)
    assert not np.array_equal(water, w3)
    # Seasons: months 0 and 6 carry opposite bands over the same noise, month 3 (amplitude
    # cos(pi/2)) the noise alone, so January plus July is twice April.
    _, _, jan = generate(cols, rows, 0, 0.3, 0)
    _, _, jul = generate(cols, rows, 0, 0.3, 6)
    _, _, apr = generate(cols, rows, 0, 0.3, 3)
    assert np.allclose(jan + jul, 2 * apr, rtol=0, atol=1e-9 * np.abs(jan).max())
    assert np.abs(jan - jul).max() > 0.1 * np.abs(jan).max()
    for land in (0.2, 0.4):
        w, _, _ = generate(cols, rows, 2, land)
        assert (w == 0).sum() == round(land * cols * rows)


def test_generator_mask_is_periodic_noise_not_stripes():
    water, _, _ = generate(256, 128, 3, 0.3)
    m = water.reshape(128, 256).astype(float)
    # Land clusters (value noise), so neighbouring cells agree far more often than chance.
    agree = (m[:, 1:] == m[:, :-1]).mean()
    assert agree > 0.9
    # Columns wrap: the seam is no more a boundary than any other column pair.
    assert (m[:, 0] == m[:, -1]).mean() > 0.85


def test_require_solved_semantics():
    """OceanAcceleratorTest.kt:151-159 and OceanCirculation.kt:140-151."""
    stopped = Solution(np.ones(4, np.float32), OC.MOST_CYCLES, 0.2)
    with pytest.raises(OceanSolveFailure):
        OC.require_solved("a stopped solve", stopped, OC.RESIDUAL_TOLERANCE)
    solved = Solution(np.ones(4, np.float32), 3, 1e-4)
    assert OC.require_solved("a finished solve", solved, OC.RESIDUAL_TOLERANCE) is solved
    nan = Solution(np.array([1, np.nan, 1, 1], np.float32), 3, 1e-4)
    with pytest.raises(OceanSolveFailure, match="cell 1"):
        OC.require_solved("the circulation", nan, OC.RESIDUAL_TOLERANCE)
    with pytest.raises(OceanSolveFailure, match="the circulation"):
        OC.require_solved("the circulation", Solution(np.ones(4, np.float32), 3, float("nan")), 1e-3)


def test_with_balance_shares_the_weight_arrays():
    p = circulation_problem(64, seed=0)
    again = OC.with_balance(p.levels[0], np.zeros(p.levels[0].cells))
    assert again.east_weight is p.levels[0].east_weight
    assert not again.forcing.any()
    back = OC.with_balance(p.levels[0], OC.balance_of(p.levels[0]))
    assert np.array_equal(back.forcing, p.levels[0].forcing)
