"""Host-side mirror of the reference's ocean-solver interface, over the C ABI of libcgocean.so.

Names, argument meanings and failure behaviour follow the reference's Kotlin:
  OceanStencil            OceanCirculation.kt:24-42
  OceanCirculation.*      OceanCirculation.kt:82-829 (stencil, withBalance, coarseWater, levels,
                          Solution, OceanSolveFailure, requireSolved, solve)
  OceanHeat               OceanHeat.kt:61-281 (diffusivity, stencil)
  OceanStage (heat grids) OceanStage.kt:449-498, 857-860, 1199-1223 (heat_levels, resample, coarse_mean)
  OceanAccelerator        OceanAccelerator.kt:15-42
  HipOcean                desktop/.../GpuOcean.kt:26-251 (createOrNull, name, solve, residency)
The relaxation, the residual and the whole solve run on the GPU through the HIP library; nothing
here computes them on the CPU.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from . import (CG_CANCELLED, CG_DECLINED, CG_METHOD_REFERENCE, CG_METHOD_REFINED, CG_OK, CG_OPT_MIN_CELLS,
               CgSolution, check, lib)


def _flat(a, dtype, n: int) -> np.ndarray:
    """a itself when it already is a flat contiguous array of dtype and size n (so that arrays
    shared between stencils keep their identity), else a converted copy."""
    if isinstance(a, np.ndarray) and a.dtype == dtype and a.ndim == 1 and a.size == n and a.flags.c_contiguous:
        return a
    return np.ascontiguousarray(a, dtype=dtype).reshape(n)


class OceanStencil:
    """One steady linear problem on one grid (OceanCirculation.kt:24-42): binary32 neighbour
    weights and forcing, the binary64 centre weight, a water mask; row-major, columns wrapping."""

    def __init__(self, cells_across: int, cells_down: int, is_water, east_weight, west_weight, north_weight,
                 south_weight, forcing, centre_weight):
        n = int(cells_across) * int(cells_down)
        self.cells_across = int(cells_across)
        self.cells_down = int(cells_down)
        self.is_water = _flat(is_water, np.uint8, n)
        self.east_weight = _flat(east_weight, np.float32, n)
        self.west_weight = _flat(west_weight, np.float32, n)
        self.north_weight = _flat(north_weight, np.float32, n)
        self.south_weight = _flat(south_weight, np.float32, n)
        self.forcing = _flat(forcing, np.float32, n)
        self.centre_weight = _flat(centre_weight, np.float64, n)

    @property
    def cells(self) -> int:
        return self.cells_across * self.cells_down


@dataclass
class Solution:
    """OceanCirculation.Solution (OceanCirculation.kt:125), with the binary64 answer when the
    method carries one and the refinement count."""
    values: np.ndarray
    cycles: int
    relative_residual: float
    values64: Optional[np.ndarray] = None
    refinements: int = 0


@dataclass
class WaterBodies:
    """OceanCirculation.WaterBodies (OceanCirculation.kt:667-668): body per cell, -1 on land."""
    body_of: np.ndarray
    count: int


class OceanSolveFailure(RuntimeError):
    """OceanCirculation.OceanSolveFailure (OceanCirculation.kt:133)."""


class OceanCirculation:
    """The reference's solver object (OceanCirculation.kt:82-829), with the heavy work on the GPU."""

    SERIES_BELOW_CELL_PECLET = 1e-4
    RESIDUAL_TOLERANCE = 1e-3
    PASSES_EACH_WAY = 2
    COARSEST_ROWS = 8
    COARSEST_TOLERANCE_SHARE = 0.1
    COARSEST_PASSES_PER_CHECK = 20
    MOST_CYCLES = 200
    KRYLOV_RESTARTS = 3
    PRECONDITIONER_COARSEST_PASSES = 200

    @staticmethod
    def require_solved(what: str, solution: Solution, tolerance: float) -> Solution:
        """requireSolved (OceanCirculation.kt:140-151)."""
        residual_finite = np.isfinite(solution.relative_residual)
        bad = np.flatnonzero(~np.isfinite(solution.values))
        first_bad = int(bad[0]) if bad.size else -1
        if not residual_finite or first_bad >= 0 or solution.relative_residual >= tolerance:
            msg = (f"{what} did not solve: relative residual {solution.relative_residual} after {solution.cycles} "
                   f"iterations against a tolerance of {tolerance}")
            if first_bad >= 0:
                msg += f", and cell {first_bad} holds {solution.values[first_bad]}"
            raise OceanSolveFailure(msg)
        return solution

    @staticmethod
    def stencil(cells_across, cells_down, cell_width_meters, cell_height_meters, is_water, beta_per_meter_second,
                bottom_drag_per_second, balance_per_second_squared=None) -> OceanStencil:
        """OceanCirculation.stencil (OceanCirculation.kt:161-219), built by the library's host code."""
        n = cells_across * cells_down
        water = np.ascontiguousarray(is_water, dtype=np.uint8).reshape(n)
        beta = np.ascontiguousarray(beta_per_meter_second, dtype=np.float64)
        e, w, no, s, f = (np.empty(n, np.float32) for _ in range(5))
        c = np.empty(n, np.float64)
        bal = None
        if balance_per_second_squared is not None:
            bal = np.ascontiguousarray(balance_per_second_squared, dtype=np.float64).reshape(n)
        check(lib().cg_circulation_stencil(cells_across, cells_down, float(cell_width_meters),
                                           float(cell_height_meters), water, beta, float(bottom_drag_per_second),
                                           None if bal is None else bal.ctypes.data_as(ctypes.c_void_p),
                                           e, w, no, s, c, f), what="cg_circulation_stencil")
        return OceanStencil(cells_across, cells_down, water, e, w, no, s, f, c)

    @staticmethod
    def with_balance(stencil: OceanStencil, balance) -> OceanStencil:
        """withBalance (OceanCirculation.kt:222-231): the same operator, another right-hand side."""
        bal = np.asarray(balance, dtype=np.float64).reshape(stencil.cells)
        f = np.zeros(stencil.cells, np.float32)
        wet = stencil.is_water != 0
        f[wet] = (bal[wet] / stencil.centre_weight[wet]).astype(np.float32)
        return OceanStencil(stencil.cells_across, stencil.cells_down, stencil.is_water, stencil.east_weight,
                            stencil.west_weight, stencil.north_weight, stencil.south_weight, f,
                            stencil.centre_weight)

    @staticmethod
    def balance_of(stencil: OceanStencil) -> np.ndarray:
        """balanceOf (OceanCirculation.kt:325-327)."""
        return stencil.forcing.astype(np.float64) * stencil.centre_weight

    @staticmethod
    def coarse_water(fine, fine_across, fine_down, coarse_across, coarse_down, every_cell_water) -> np.ndarray:
        """coarseWater (OceanCirculation.kt:404-423)."""
        out = np.empty(coarse_across * coarse_down, np.uint8)
        check(lib().cg_coarse_water(np.ascontiguousarray(fine, dtype=np.uint8), fine_across, fine_down,
                                    coarse_across, coarse_down, int(bool(every_cell_water)), out),
              what="cg_coarse_water")
        return out

    @staticmethod
    def level_shapes(across, down, dx, dy, max_levels=24):
        a = np.empty(max_levels, np.int32)
        d = np.empty(max_levels, np.int32)
        w = np.empty(max_levels)
        h = np.empty(max_levels)
        n = check(lib().cg_level_shapes(across, down, float(dx), float(dy), max_levels, a, d, w, h),
                  what="cg_level_shapes")
        return [(int(a[i]), int(d[i]), float(w[i]), float(h[i])) for i in range(n)]

    @staticmethod
    def levels(finest: OceanStencil, cell_width_meters: float, cell_height_meters: float, every_cell_water: bool,
               coarser: Callable[[int, int, np.ndarray], OceanStencil]) -> List[OceanStencil]:
        """levels (OceanCirculation.kt:434-462): the grids of a V-cycle, finest first."""
        out = [finest]
        shapes = OceanCirculation.level_shapes(finest.cells_across, finest.cells_down, cell_width_meters,
                                               cell_height_meters)
        water, a, d = finest.is_water, finest.cells_across, finest.cells_down
        for (na, nd, _, _) in shapes[1:]:
            water = OceanCirculation.coarse_water(water, a, d, na, nd, every_cell_water)
            a, d = na, nd
            out.append(coarser(a, d, water))
        return out


    @staticmethod
    def water_bodies(is_water, cells_across: int, cells_down: int) -> WaterBodies:
        """waterBodies (OceanCirculation.kt:670-697)."""
        out = np.empty(cells_across * cells_down, np.int32)
        count = check(lib().cg_water_bodies(np.ascontiguousarray(is_water, dtype=np.uint8).reshape(-1), cells_across,
                                            cells_down, out), what="cg_water_bodies")
        return WaterBodies(out, int(count))


def resample(field, from_across: int, from_down: int, to_across: int, to_down: int) -> np.ndarray:
    """OceanStage.resample (OceanStage.kt:1199-1223)."""
    out = np.empty(to_across * to_down, np.float32)
    check(lib().cg_resample(np.ascontiguousarray(field, dtype=np.float32).reshape(-1), from_across, from_down,
                            to_across, to_down, out), what="cg_resample")
    return out


def coarse_mean(fine, fine_across: int, fine_down: int, coarse_across: int, coarse_down: int) -> np.ndarray:
    """OceanStage.coarseMean (OceanStage.kt:857-860)."""
    out = np.empty(coarse_across * coarse_down, np.float32)
    check(lib().cg_coarse_mean(np.ascontiguousarray(fine, dtype=np.float32).reshape(-1), fine_across, fine_down,
                               coarse_across, coarse_down, out), what="cg_coarse_mean")
    return out


class OceanHeat:
    """The sea's heat balance (OceanHeat.kt:61-281): u.grad T - K lap T + (T - T_lat)/tau +
    (w/h)(T - T_sub) = 0, exponentially fitted per face, solved by the same multigrid."""

    MIDLATITUDE_DIFFUSIVITY_M2_PER_S = 2500.0
    MIDLATITUDE_DEGREES = 45.0
    # OceanStage.RELAXATION_SECONDS (OceanStage.kt:113-114): a 50 m mixed layer's heat capacity,
    # 4e6 J/m^3/C, over the 25 W/m^2/C surface exchange (EnergyBalance.kt:688, 738, 754, 832).
    RELAXATION_SECONDS = 50.0 * 4.0e6 / 25.0
    HEAT_SOLVE_RESTARTS = 3

    @staticmethod
    def diffusivity(latitude_degrees: float, radius_meters: float) -> float:
        """diffusivity (OceanHeat.kt:128-129)."""
        return float(lib().cg_heat_diffusivity(float(latitude_degrees), float(radius_meters)))

    @staticmethod
    def stencil(cells_across, cells_down, cell_width_meters, cell_height_meters, is_water, stream, target_c,
                relaxation_seconds, with_target, radius_meters=None, constant_diffusivity=None,
                entrainment_per_s=None, subsurface_c=None) -> OceanStencil:
        """stencil (OceanHeat.kt:151-228). The diffusivity is diffusivity(latitude, radius_meters),
        or constant_diffusivity where given (the reference's diffusivityAt lambda)."""
        n = cells_across * cells_down
        water = np.ascontiguousarray(is_water, dtype=np.uint8).reshape(n)
        e, w, no, s, f = (np.empty(n, np.float32) for _ in range(5))
        c = np.empty(n, np.float64)

        def opt(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(n)

        tgt, ent, sub = opt(target_c), opt(entrainment_per_s), opt(subsurface_c)
        keep = [a for a in (tgt, ent, sub) if a is not None]
        ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().cg_heat_stencil(cells_across, cells_down, float(cell_width_meters), float(cell_height_meters),
                                    water, np.ascontiguousarray(stream, dtype=np.float32).reshape(n), ptr(tgt),
                                    float(relaxation_seconds), int(bool(with_target)),
                                    float(radius_meters or 0.0), float(constant_diffusivity or 0.0), ptr(ent),
                                    ptr(sub), e, w, no, s, c, f), what="cg_heat_stencil")
        del keep
        return OceanStencil(cells_across, cells_down, water, e, w, no, s, f, c)

    @staticmethod
    def levels(across, down, width_m, height_m, is_water, stream, target_c, radius_meters=None,
               constant_diffusivity=None, entrainment_per_s=None, subsurface_c=None,
               relaxation_seconds=None) -> List[OceanStencil]:
        """The heat's grids as OceanStage.circulateUnder builds them (OceanStage.kt:463-479): the
        finest with its target, the coarse ones with the stream resampled, the entrainment's block
        means and no target; coarse masks keep any water (everyCellWater = false)."""
        tau = OceanHeat.RELAXATION_SECONDS if relaxation_seconds is None else relaxation_seconds
        stream = np.ascontiguousarray(stream, dtype=np.float32).reshape(-1)
        finest = OceanHeat.stencil(across, down, width_m, height_m, is_water, stream, target_c, tau, True,
                                   radius_meters, constant_diffusivity, entrainment_per_s, subsurface_c)

        def coarser(a, d, w):
            ent = None if entrainment_per_s is None else coarse_mean(entrainment_per_s, across, down, a, d)
            return OceanHeat.stencil(a, d, width_m * across / a, height_m * down / d, w,
                                     resample(stream, across, down, a, d), np.zeros(a * d, np.float32), tau, False,
                                     radius_meters, constant_diffusivity, ent, None)

        return OceanCirculation.levels(finest, width_m, height_m, False, coarser)


class OceanAccelerator:
    """OceanAccelerator (OceanAccelerator.kt:15-42): relax on a device, or return None to decline."""

    name: str = ""

    def solve(self, stencil: OceanStencil, start: np.ndarray, passes: int) -> Optional[np.ndarray]:
        raise NotImplementedError


class HipOcean(OceanAccelerator):
    """The MI355X accelerator: GpuOcean's role (GpuOcean.kt:26-251) over the HIP library.

    Stencils are made resident on first sight and kept in an LRU of MOST_RESIDENT_STENCILS keyed on
    the identity of their east-weight array, as GpuOcean keys its residency (GpuOcean.kt:104-137);
    a stencil rebuilt by with_balance shares that array and so hits the cache. Evicting a stencil
    that a live Plan was built on is safe: the library keeps its memory until that plan is closed."""

    MOST_RESIDENT_STENCILS = 24
    SMALLEST_GRID_WORTH_THE_CARD_CELLS = 1 << 16

    @dataclass
    class Result:
        accelerator: Optional["HipOcean"]
        unavailable_because: Optional[str]

    def __init__(self, ctx: ctypes.c_void_p, name: str):
        self._ctx = ctx
        self.name = name
        self._resident: "OrderedDict[int, tuple]" = OrderedDict()

    @staticmethod
    def create_or_null(device: int = 0) -> "HipOcean.Result":
        """createOrNull (GpuOcean.kt:189-202): an accelerator, or the reason there is none."""
        ctx = ctypes.c_void_p()
        reason = ctypes.create_string_buffer(512)
        status = lib().cg_ocean_open(device, ctypes.byref(ctx), reason, len(reason))
        if status != CG_OK or not ctx:
            return HipOcean.Result(None, reason.value.decode() or f"status {status}")
        name = ctypes.create_string_buffer(256)
        lib().cg_ocean_name(ctx, name, len(name))
        return HipOcean.Result(HipOcean(ctx, name.value.decode()), None)

    @property
    def ctx(self):
        return self._ctx

    def close(self) -> None:
        if self._ctx:
            for (_, handle) in self._resident.values():
                lib().cg_ocean_stencil_release(self._ctx, handle)
            self._resident.clear()
            lib().cg_ocean_close(self._ctx)
            self._ctx = None

    def set_option(self, option: int, value: int) -> None:
        check(lib().cg_ocean_set_option(self._ctx, option, int(value)), self._ctx, "cg_ocean_set_option")

    def get_option(self, option: int) -> int:
        return int(lib().cg_ocean_get_option(self._ctx, option))

    def holding(self, stencil: OceanStencil) -> int:
        """The device handle of a stencil, uploading it the first time it is seen."""
        key = id(stencil.east_weight)
        hit = self._resident.get(key)
        if hit is not None and hit[0] is stencil.east_weight:
            self._resident.move_to_end(key)
            return hit[1]
        while len(self._resident) >= self.MOST_RESIDENT_STENCILS:
            _, (_, old) = self._resident.popitem(last=False)
            lib().cg_ocean_stencil_release(self._ctx, old)
        handle = ctypes.c_uint64(0)
        check(lib().cg_ocean_stencil_upload(self._ctx, stencil.cells_across, stencil.cells_down, stencil.is_water,
                                            stencil.east_weight, stencil.west_weight, stencil.north_weight,
                                            stencil.south_weight,
                                            stencil.centre_weight.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.byref(handle)), self._ctx, "cg_ocean_stencil_upload")
        self._resident[key] = (stencil.east_weight, handle.value)
        return handle.value

    def solve(self, stencil: OceanStencil, start: np.ndarray, passes: int, cancel=None) -> Optional[np.ndarray]:
        """OceanAccelerator.solve (OceanAccelerator.kt:37-41): `passes` red-black passes from start,
        a new array back, or None when the device declines (or saw `cancel` set: a ctypes c_int)."""
        across, down = stencil.cells_across, stencil.cells_down
        if across <= 0 or across % 2 != 0 or down <= 0 or passes < 0:
            return None
        start = np.ascontiguousarray(start, dtype=np.float32)
        if start.size != across * down or stencil.is_water.size != across * down:
            return None
        if across * down < lib().cg_ocean_get_option(self._ctx, CG_OPT_MIN_CELLS):
            return None
        handle = self.holding(stencil)
        out = np.empty(across * down, np.float32)
        status = lib().cg_ocean_relax(self._ctx, handle, stencil.forcing, start, out, int(passes),
                                      None if cancel is None else ctypes.cast(ctypes.pointer(cancel), ctypes.c_void_p))
        if status in (CG_DECLINED, CG_CANCELLED):
            return None
        check(status, self._ctx, "cg_ocean_relax")
        return out

    def solve_levels(self, levels: List[OceanStencil], start: Optional[np.ndarray], tolerance: float,
                     pole_is_wall: bool = True, method: int = CG_METHOD_REFERENCE, cancel=None) -> Solution:
        """The whole solve on the device: OceanCirculation.solve (OceanCirculation.kt:473-493) for
        CG_METHOD_REFERENCE, or the binary32-in-binary64 refinement for CG_METHOD_REFINED."""
        handles = (ctypes.c_uint64 * len(levels))(*[self.holding(s) for s in levels])
        n = levels[0].cells
        out = np.empty(n, np.float32)
        out64 = np.empty(n, np.float64)
        info = CgSolution()
        st = None if start is None else np.ascontiguousarray(start, dtype=np.float32)
        check(lib().cg_ocean_solve(self._ctx, handles, len(levels), levels[0].forcing,
                                   None if st is None else st.ctypes.data_as(ctypes.c_void_p), float(tolerance),
                                   int(bool(pole_is_wall)), int(method), out.ctypes.data_as(ctypes.c_void_p),
                                   out64.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info),
                                   None if cancel is None else ctypes.cast(ctypes.pointer(cancel), ctypes.c_void_p)),
              self._ctx, "cg_ocean_solve")
        return Solution(out, info.cycles, info.relative_residual, out64, info.refinements)


    def water_on(self, map_land, map_across: int, map_down: int, across: int, down: int) -> np.ndarray:
        """OceanStage.waterOn (OceanStage.kt:551-573) on the device: the solve grid's water mask from
        a map's land mask (non-zero for land)."""
        out = np.empty(across * down, np.uint8)
        land = np.ascontiguousarray(map_land, dtype=np.uint8).reshape(map_across * map_down)
        check(lib().cg_ocean_water_on(self._ctx, land, map_across, map_down, across, down, out),
              self._ctx, "cg_ocean_water_on")
        return out

    def curl_forcing(self, across: int, down: int, width_m: float, height_m: float, eastward, northward,
                     depth=None) -> np.ndarray:
        """OceanStage.curlForcing (OceanStage.kt:689-708) on the device; depth None takes the
        reference's 500 m wind-driven layer."""
        n = across * down
        out = np.empty(n, np.float64)
        d = None if depth is None else np.ascontiguousarray(depth, dtype=np.float64).reshape(n)
        check(lib().cg_ocean_curl_forcing(self._ctx, across, down, float(width_m), float(height_m),
                                          np.ascontiguousarray(eastward, dtype=np.float64).reshape(n),
                                          np.ascontiguousarray(northward, dtype=np.float64).reshape(n),
                                          None if d is None else d.ctypes.data_as(ctypes.c_void_p), out),
              self._ctx, "cg_ocean_curl_forcing")
        return out


class Plan:
    """A device-resident solve (cg_ocean_plan_*): ladder, buffers, forcing and answer stay in HBM."""

    def __init__(self, acc: HipOcean, levels: List[OceanStencil], pole_is_wall: bool = True):
        self.acc = acc
        self.levels = levels
        handles = (ctypes.c_uint64 * len(levels))(*[acc.holding(s) for s in levels])
        self._p = ctypes.c_void_p()
        check(lib().cg_ocean_plan_create(acc.ctx, handles, len(levels), int(bool(pole_is_wall)),
                                         ctypes.byref(self._p)), acc.ctx, "cg_ocean_plan_create")
        self.set_forcing(levels[0].forcing)

    @property
    def handle(self):
        return self._p

    def set_forcing(self, forcing) -> None:
        check(lib().cg_ocean_plan_set_forcing(self._p, np.ascontiguousarray(forcing, dtype=np.float32)),
              self.acc.ctx, "cg_ocean_plan_set_forcing")

    def solve(self, tol: float, method: int = CG_METHOD_REFINED, warm: bool = False) -> CgSolution:
        info = CgSolution()
        check(lib().cg_ocean_plan_solve(self._p, float(tol), int(method), int(warm), ctypes.byref(info)),
              self.acc.ctx, "cg_ocean_plan_solve")
        return info

    def set_stress(self, width_m: float, height_m: float, eastward, northward, depth=None) -> None:
        """The finest forcing from a wind stress, on the device: curlForcing then withBalance
        (OceanStage.kt:689-708, OceanCirculation.kt:222-231)."""
        n = self.levels[0].cells
        d = None if depth is None else np.ascontiguousarray(depth, dtype=np.float64).reshape(n)
        check(lib().cg_ocean_plan_set_stress(self._p, float(width_m), float(height_m),
                                             np.ascontiguousarray(eastward, dtype=np.float64).reshape(n),
                                             np.ascontiguousarray(northward, dtype=np.float64).reshape(n),
                                             None if d is None else d.ctypes.data_as(ctypes.c_void_p)),
              self.acc.ctx, "cg_ocean_plan_set_stress")

    def velocities(self, width_m: float, height_m: float):
        """OceanCirculation.velocities (OceanCirculation.kt:801-828) of the last answer, on the device:
        (eastward, northward) in m/s, binary32."""
        n = self.levels[0].cells
        e = np.empty(n, np.float32)
        nn = np.empty(n, np.float32)
        check(lib().cg_ocean_plan_velocities(self._p, float(width_m), float(height_m),
                                             e.ctypes.data_as(ctypes.c_void_p), nn.ctypes.data_as(ctypes.c_void_p)),
              self.acc.ctx, "cg_ocean_plan_velocities")
        return e, nn

    def set_bodies(self, bodies: Optional[WaterBodies]) -> None:
        """The finest grid's water bodies for correctBodyMeans (OceanCirculation.kt:709-731); None
        removes them."""
        if bodies is None or bodies.count == 0:
            check(lib().cg_ocean_plan_set_bodies(self._p, None, 0), self.acc.ctx, "cg_ocean_plan_set_bodies")
            return
        b = np.ascontiguousarray(bodies.body_of, dtype=np.int32).reshape(-1)
        check(lib().cg_ocean_plan_set_bodies(self._p, b.ctypes.data_as(ctypes.c_void_p), int(bodies.count)),
              self.acc.ctx, "cg_ocean_plan_set_bodies")

    def correct_bodies(self, values) -> np.ndarray:
        """correctBodyMeans on the device against the plan's finest forcing."""
        out = np.empty(self.levels[0].cells, np.float32)
        check(lib().cg_ocean_plan_correct_bodies(self._p, np.ascontiguousarray(values, dtype=np.float32), out),
              self.acc.ctx, "cg_ocean_plan_correct_bodies")
        return out

    def autotune(self, reps: int = 20) -> List[int]:
        """The plan's tile heights per grid, chosen by timing (cg_ocean_plan_autotune); 0 where a grid
        is not tiled."""
        rows = (ctypes.c_int * len(self.levels))()
        check(lib().cg_ocean_plan_autotune(self._p, int(reps), rows), self.acc.ctx, "cg_ocean_plan_autotune")
        return list(rows)

    def set_tile_rows(self, rows) -> None:
        """Tile heights per grid (0 for the default rule), as autotune returns them."""
        arr = (ctypes.c_int * len(self.levels))(*[int(r) for r in rows])
        check(lib().cg_ocean_plan_set_tile_rows(self._p, arr), self.acc.ctx, "cg_ocean_plan_set_tile_rows")

    def set_start(self, start) -> None:
        """The values a warm solve starts from."""
        check(lib().cg_ocean_plan_set_start(self._p, np.ascontiguousarray(start, dtype=np.float32)),
              self.acc.ctx, "cg_ocean_plan_set_start")

    def read(self):
        n = self.levels[0].cells
        out64 = np.empty(n, np.float64)
        out32 = np.empty(n, np.float32)
        check(lib().cg_ocean_plan_read(self._p, out64.ctypes.data_as(ctypes.c_void_p),
                                       out32.ctypes.data_as(ctypes.c_void_p)), self.acc.ctx, "cg_ocean_plan_read")
        return out64, out32

    def close(self) -> None:
        if self._p:
            lib().cg_ocean_plan_destroy(self._p)
            self._p = None


class BatchPlan:
    """Several right-hand sides on one resident ladder, solved together (cg_ocean_plan_create_batch):
    each kernel of the cycle covers every instance in one launch. Forcings are per instance;
    solve_all runs the refined method from zero for all of them."""

    def __init__(self, acc: HipOcean, levels: List[OceanStencil], instances: int, pole_is_wall: bool = True):
        self.acc = acc
        self.levels = levels
        self.instances = int(instances)
        handles = (ctypes.c_uint64 * len(levels))(*[acc.holding(s) for s in levels])
        self._p = ctypes.c_void_p()
        check(lib().cg_ocean_plan_create_batch(acc.ctx, handles, len(levels), int(bool(pole_is_wall)), self.instances,
                                               ctypes.byref(self._p)), acc.ctx, "cg_ocean_plan_create_batch")

    @property
    def handle(self):
        return self._p

    def set_forcing(self, instance: int, forcing) -> None:
        check(lib().cg_ocean_plan_set_forcing_at(self._p, int(instance), np.ascontiguousarray(forcing, dtype=np.float32)),
              self.acc.ctx, "cg_ocean_plan_set_forcing_at")

    def set_tile_rows(self, rows) -> None:
        """As Plan.set_tile_rows."""
        arr = (ctypes.c_int * len(self.levels))(*[int(r) for r in rows])
        check(lib().cg_ocean_plan_set_tile_rows(self._p, arr), self.acc.ctx, "cg_ocean_plan_set_tile_rows")

    def autotune(self, reps: int = 20) -> List[int]:
        """As Plan.autotune, timing the launches of all instances."""
        rows = (ctypes.c_int * len(self.levels))()
        check(lib().cg_ocean_plan_autotune(self._p, int(reps), rows), self.acc.ctx, "cg_ocean_plan_autotune")
        return list(rows)

    def solve_all(self, tol: float) -> List[CgSolution]:
        infos = (CgSolution * self.instances)()
        check(lib().cg_ocean_plan_solve_all(self._p, float(tol), CG_METHOD_REFINED, infos), self.acc.ctx,
              "cg_ocean_plan_solve_all")
        return list(infos)

    def read(self, instance: int):
        n = self.levels[0].cells
        out64 = np.empty(n, np.float64)
        out32 = np.empty(n, np.float32)
        check(lib().cg_ocean_plan_read_at(self._p, int(instance), out64.ctypes.data_as(ctypes.c_void_p),
                                          out32.ctypes.data_as(ctypes.c_void_p)), self.acc.ctx, "cg_ocean_plan_read_at")
        return out64, out32

    def close(self) -> None:
        if self._p:
            lib().cg_ocean_plan_destroy(self._p)
            self._p = None


class PlanGroup:
    """Plans of one accelerator solved together by the refined method from zero (cg_ocean_group_*):
    the worlds of uneven size of one batch, each a Plan or a BatchPlan of its own ladder, stepped by
    one host loop: by default (CG_OPT_GROUP_MERGED) one graph launch a step for the plans still
    running, each kernel launch covering one depth of every such plan's ladder; else each plan's step
    graph on a stream of the group's own. Each instance stops at the step, and with the answer, it
    would reach alone. Close the group before its plans."""

    def __init__(self, acc: HipOcean, plans):
        self.acc = acc
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle for p in self.plans])
        self._g = ctypes.c_void_p()
        check(lib().cg_ocean_group_create(arr, len(self.plans), ctypes.byref(self._g)), acc.ctx,
              "cg_ocean_group_create")

    def solve(self, tol: float, method: int = CG_METHOD_REFINED) -> List[List[CgSolution]]:
        """One list of results per plan (one per instance); method CG_METHOD_REFINED or, with every
        plan's coarse network set, CG_METHOD_LEARNED."""
        counts = [getattr(p, "instances", 1) for p in self.plans]
        infos = (CgSolution * sum(counts))()
        check(lib().cg_ocean_group_solve(self._g, float(tol), int(method), infos), self.acc.ctx,
              "cg_ocean_group_solve")
        out, at = [], 0
        for c in counts:
            out.append(list(infos[at:at + c]))
            at += c
        return out

    def close(self) -> None:
        if self._g:
            lib().cg_ocean_group_destroy(self._g)
            self._g = None
