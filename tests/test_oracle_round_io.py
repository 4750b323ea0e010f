"""The oracle's input and output steps of a circulation round, pinned (CPU only).

- waterOn (OceanStage.kt:551-573): against a brute force in exact rational arithmetic (a solve cell
  is water when no land map cell overlaps it by more than 1e-6 of a map cell), at finer, coarser,
  equal and non-dividing map/solve ratios; and the reference's promise that a one-cell land strip
  of the map stays land on the solve grid.
- curlForcing (OceanStage.kt:689-708): closed forms, exact for the differences it takes: a stress
  linear in the row (edge rows one-sided give the same slope), a northward stress that is a sine of
  the column (the wrap), the sign convention (row 0 is north), the 500 m layer versus per-cell
  depth, and linearity; plus the synthetic workload's balance being its curl.
- velocities (OceanCirculation.kt:801-828): a stream function linear in the row (uniform eastward
  flow), a sine in the column (the wrap), the wall beyond each pole, zero on land, and the discrete
  divergence of the velocities vanishing wherever all four neighbours are water.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle

RHO = 1025.0
LAYER = 500.0


def water_on_exact(land, ma, md, across, down):
    """A solve cell [c a/A, (c+1) a/A) x [r d/D, (r+1) d/D) in map-cell units is land when any map
    cell it overlaps by more than 1e-6 (of a map cell, along each axis) is land."""
    touch = Fraction(1, 10**6)
    L = np.asarray(land, dtype=bool).reshape(md, ma)
    out = np.zeros((down, across), np.uint8)

    def span(i, n_map, n_solve):
        lo, hi = Fraction(i * n_map, n_solve), Fraction((i + 1) * n_map, n_solve)
        return [m for m in range(n_map) if min(hi, m + 1) - max(lo, m) > touch]

    rows = [span(r, md, down) for r in range(down)]
    cols = [span(c, ma, across) for c in range(across)]
    for r in range(down):
        for c in range(across):
            out[r, c] = 0 if L[np.ix_(rows[r], cols[c])].any() else 1
    return out.reshape(-1)


@pytest.mark.parametrize("ma,md,across,down", [
    (64, 32, 32, 16),    # map twice as fine
    (32, 16, 64, 32),    # map twice as coarse
    (40, 20, 40, 20),    # equal
    (60, 30, 40, 20),    # 3:2
    (50, 25, 30, 15),    # 5:3
    (37, 19, 16, 8),     # no common divisor
    (16, 8, 37, 19),     # coarser map, no common divisor
])
def test_water_on_matches_the_exact_overlap_rule(ma, md, across, down):
    rng = np.random.default_rng(ma * 1000 + across)
    land = (rng.random(ma * md) < 0.25).astype(np.uint8)
    got = oracle.water_on(land, ma, md, across, down)
    assert np.array_equal(got, water_on_exact(land, ma, md, across, down))


def test_water_on_keeps_a_one_cell_land_strip():
    ma, md, across, down = 90, 45, 40, 20
    land = np.zeros((md, ma), np.uint8)
    land[:, 31] = 1  # a meridional strip one map cell wide
    w = oracle.water_on(land.reshape(-1), ma, md, across, down).reshape(down, across)
    assert (w == 0).any(axis=1).all(), "every solve row must keep the strip as land"
    assert w.mean() > 0.9
    assert oracle.water_on(np.zeros(ma * md, np.uint8), ma, md, across, down).all()


def _grid(across, down):
    width = 40075e3 / across
    height = 20037.5e3 / down
    return width, height


def test_curl_of_a_stress_linear_in_the_row():
    across, down = 24, 12
    w, h = _grid(across, down)
    s = 1e-3  # N/m^2 per row, growing southward
    east = np.repeat(np.arange(down) * s, across).astype(np.float64)
    north = np.zeros(across * down)
    F = oracle.curl_forcing(across, down, w, h, east, north)
    # d tau_e / dy with y north is -s / h; the curl is +s / h, over rho times the 500 m layer.
    expect = (s / h) / (RHO * LAYER)
    assert np.allclose(F, expect, rtol=1e-12, atol=0)
    F2 = oracle.curl_forcing(across, down, w, h, east * 3.0, north)
    assert np.allclose(F2, 3.0 * F, rtol=1e-13, atol=0)


def test_curl_of_a_northward_stress_wrapping_in_columns():
    across, down = 32, 8
    w, h = _grid(across, down)
    k = 2 * np.pi / across
    col = np.tile(np.arange(across), down)
    north = 0.05 * np.sin(k * col)
    east = np.zeros(across * down)
    depth = np.full(across * down, 4000.0)
    F = oracle.curl_forcing(across, down, w, h, east, north, depth)
    # (sin(k(c+1)) - sin(k(c-1))) / 2w = cos(kc) sin(k) / w, exactly in real arithmetic.
    expect = 0.05 * np.cos(k * col) * np.sin(k) / w / (RHO * 4000.0)
    assert np.allclose(F, expect, rtol=0, atol=1e-12 * np.abs(expect).max())


def test_curl_depth_divides_per_cell():
    across, down = 16, 8
    w, h = _grid(across, down)
    rng = np.random.default_rng(3)
    east = rng.normal(0, 0.1, across * down)
    north = rng.normal(0, 0.1, across * down)
    depth = rng.uniform(200, 6000, across * down)
    F500 = oracle.curl_forcing(across, down, w, h, east, north)
    Fd = oracle.curl_forcing(across, down, w, h, east, north, depth)
    assert np.allclose(Fd * depth, F500 * LAYER, rtol=1e-13, atol=0)


def test_synthetic_balance_is_the_curl_of_the_synthetic_stress():
    from cartogenesis_amd.problem import generate, stress
    for month in (-1, 0, 4):
        across, down = 64, 32
        water, depth, balance = generate(across, down, 5, 0.3, month)
        e, n = stress(across, down, 5, month)
        assert np.all(n == 0)
        w, h = _grid(across, down)
        F = oracle.curl_forcing(across, down, w, h, e, n, np.where(water != 0, depth, 1.0))
        wet = water != 0
        assert np.array_equal(F[wet], balance[wet])
        assert np.all(balance[~wet] == 0)


def _stencil(water, across, down):
    z = np.zeros(across * down, np.float32)
    return oracle.Stencil(across, down, water, z, z, z, z, z, np.ones(across * down))


def test_velocities_of_a_stream_linear_in_the_row():
    across, down = 16, 10
    w, h = _grid(across, down)
    a = 0.25  # exactly representable steps
    psi = np.repeat(np.arange(down, dtype=np.float32) * np.float32(a), across)
    water = np.ones(across * down, np.uint8)
    e, n = oracle.velocities(_stencil(water, across, down), psi, w, h)
    E = e.reshape(down, across)
    # psi grows southward: u = -d psi/dy = a / h in the interior.
    assert np.all(E[1:-1] == np.float32(a / h))
    assert np.all(n == 0)
    # Row 0: psi beyond the pole is -psi[0] = 0, so u = -(0 - a) / 2h.
    assert np.all(E[0] == np.float32(a / (2 * h)))
    # Last row: psi beyond is -psi[last]: u = -(psi[l-1] + psi[l]) / 2h.
    last = (down - 1) * a
    assert np.all(E[-1] == np.float32(-((last - a) - (-last)) / (2 * h)))


def test_velocities_wrap_in_columns_and_vanish_on_land():
    across, down = 32, 6
    w, h = _grid(across, down)
    k = 2 * np.pi / across
    col = np.tile(np.arange(across), down)
    psi = (1e4 * np.sin(k * col)).astype(np.float32)
    water = np.ones(across * down, np.uint8)
    water[across + 7] = 0
    e, n = oracle.velocities(_stencil(water, across, down), psi, w, h)
    expect = 1e4 * np.cos(k * col) * np.sin(k) / w
    wet = water != 0
    assert np.allclose(n[wet], expect[wet], rtol=0, atol=1e-6 * np.abs(expect).max())
    assert e[across + 7] == 0 and n[across + 7] == 0
    E, P = e.reshape(down, across), psi.reshape(down, across).astype(np.float64)
    assert np.all(E[1:-1] == 0)
    # Beyond either pole psi is minus the edge row's (a wall): u = +-psi_edge / dy there.
    assert np.allclose(E[0], P[0] / h, rtol=1e-6) and np.allclose(E[-1], -P[-1] / h, rtol=1e-6)


def test_velocities_are_divergence_free_between_water():
    across, down = 48, 24
    w, h = _grid(across, down)
    rng = np.random.default_rng(11)
    psi = rng.normal(0, 1e5, across * down).astype(np.float32)
    water = (rng.random(across * down) > 0.1).astype(np.uint8)
    e, n = oracle.velocities(_stencil(water, across, down), psi, w, h)
    E, N, W = e.reshape(down, across).astype(np.float64), n.reshape(down, across).astype(np.float64), \
        water.reshape(down, across) != 0
    # (u_E - u_W) / 2dx + (v_N - v_S) / 2dy over cells whose four neighbours are water and which are
    # two rows from either pole.
    div = (np.roll(E, -1, 1) - np.roll(E, 1, 1)) / (2 * w) + (np.roll(N, 1, 0) - np.roll(N, -1, 0)) / (2 * h)
    ok = W & np.roll(W, -1, 1) & np.roll(W, 1, 1) & np.roll(W, 1, 0) & np.roll(W, -1, 0)
    ok[:2] = False
    ok[-2:] = False
    scale = np.abs(psi).max() / (w * h)
    assert ok.sum() > 100
    assert np.abs(div[ok]).max() < 1e-5 * scale
    # A swapped sign in one component breaks it.
    bad = (np.roll(E, -1, 1) - np.roll(E, 1, 1)) / (2 * w) - (np.roll(N, 1, 0) - np.roll(N, -1, 0)) / (2 * h)
    assert np.abs(bad[ok]).max() > 1e-2 * scale
