"""The synthetic circulation workloads, built by the library's host code.

A problem is the reference's circulation on a `rows` x `2 rows` latitude-longitude grid of an
Earth-sized world (WorldGenConfig.forRows, WGC:3381-3382): the ladder of stencils
OceanStage.circulateUnder builds (OceanStage.kt:425-435, circulationStencil OceanStage.kt:579-594),
with the balance curl_z(tau) / (rho H) of a seeded synthetic world (cg_problem_generate).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List

import numpy as np

from . import check, lib
from .ocean import OceanCirculation, OceanStencil

EARTH_EQUATOR_KM = 40075.0  # WorldScale.EARTH_EQUATOR_KM (WGC)


def bottom_drag() -> float:
    return float(lib().cg_bottom_drag())


def beta_rows(down: int, world_width_km: float = EARTH_EQUATOR_KM) -> np.ndarray:
    out = np.empty(down, np.float64)
    check(lib().cg_beta_rows(down, float(world_width_km), out), what="cg_beta_rows")
    return out


def generate(across: int, down: int, seed: int, land_fraction: float = 0.3, month: int = -1):
    """(water uint8, depth float64, balance float64) of the synthetic world."""
    n = across * down
    water = np.empty(n, np.uint8)
    depth = np.empty(n, np.float64)
    balance = np.empty(n, np.float64)
    check(lib().cg_problem_generate(across, down, int(seed), float(land_fraction), int(month),
                                    water.ctypes.data_as(ctypes.c_void_p), depth.ctypes.data_as(ctypes.c_void_p),
                                    balance.ctypes.data_as(ctypes.c_void_p)), what="cg_problem_generate")
    return water, depth, balance


def stress(across: int, down: int, seed: int, month: int = -1):
    """(eastward, northward) wind stress of the synthetic world, N/m^2: generate's balance is its curl
    over rho H."""
    n = across * down
    e = np.empty(n, np.float64)
    nn = np.empty(n, np.float64)
    check(lib().cg_problem_stress(across, down, int(seed), int(month), e.ctypes.data_as(ctypes.c_void_p),
                                  nn.ctypes.data_as(ctypes.c_void_p)), what="cg_problem_stress")
    return e, nn


@dataclass
class Problem:
    rows: int
    cols: int
    seed: int
    land_fraction: float
    month: int
    water: np.ndarray
    depth: np.ndarray
    balance: np.ndarray
    levels: List[OceanStencil]
    dx: float
    dy: float


def circulation_levels(cols: int, rows: int, water: np.ndarray, balance, world_width_km: float = EARTH_EQUATOR_KM):
    """The circulation's ladder as circulateUnder builds it (OceanStage.kt:425-435)."""
    width_m = world_width_km * 1000.0
    height_m = width_m * 0.5
    dx, dy = width_m / cols, height_m / rows
    drag = bottom_drag()
    finest = OceanCirculation.stencil(cols, rows, dx, dy, water, beta_rows(rows, world_width_km), drag, balance)

    def coarser(a, d, w):
        return OceanCirculation.stencil(a, d, width_m / a, height_m / d, w, beta_rows(d, world_width_km), drag, None)

    return OceanCirculation.levels(finest, dx, dy, True, coarser), dx, dy


def circulation_problem(rows: int, cols: int = 0, seed: int = 0, land_fraction: float = 0.3,
                        month: int = -1) -> Problem:
    cols = cols or 2 * rows
    water, depth, balance = generate(cols, rows, seed, land_fraction, month)
    levels, dx, dy = circulation_levels(cols, rows, water, balance)
    return Problem(rows, cols, seed, land_fraction, month, water, depth, balance, levels, dx, dy)


def monthly_forcings(problem: Problem, months=range(12)) -> List[np.ndarray]:
    """The finest binary32 forcing (withBalance of the month's curl) of each month of one synthetic
    world: its mask and depth are the year's, the stress band's amplitude and latitude the month's."""
    out = []
    for m in months:
        _, _, balance = generate(problem.cols, problem.rows, problem.seed, problem.land_fraction, int(m))
        out.append(OceanCirculation.with_balance(problem.levels[0], balance).forcing)
    return out
