"""ORACLE: test infrastructure only.

ctypes bindings of ``ocean_oracle.c``, the plain-C restatement of the reference's
ocean-circulation multigrid (OceanCirculation.kt). Only the tests, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this module, and only as the checker or the timed CPU
baseline; the product library never loads it.

Every function here forwards to a C function whose comment cites the reference lines it restates.
The checks that pin each function are in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboceanoracle.so"

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int, _c_double, _c_long = ctypes.c_int, ctypes.c_double, ctypes.c_long
_vpp = ctypes.POINTER(ctypes.c_void_p)

_lib = None


def build() -> Path:
    """Compiles the oracle with its Makefile (gcc, -ffp-contract=off, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = ctypes.CDLL(str(LIB_PATH))
    L.or_set_threads.argtypes = [_c_int]
    L.or_stencil.argtypes = [_c_int, _c_int, _c_double, _c_double, _u8p, _f64p, _c_double, _f64p,
                             _f32p, _f32p, _f32p, _f32p, _f64p, _f32p]
    L.or_relax.argtypes = [_c_int, _c_int, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _c_int]
    L.or_residual.argtypes = [_c_int, _c_int, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p, _f32p, _f64p]
    L.or_residual64.argtypes = [_c_int, _c_int, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p, _f64p, _f64p]
    L.or_relax64_flat.argtypes = [_c_int, _c_int, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p, _c_int]
    L.or_largest.argtypes = [_c_long, _f64p]
    L.or_largest.restype = _c_double
    L.or_restrict.argtypes = [_f64p, _c_int, _c_int, _c_int, _c_int, _f64p]
    L.or_prolong_add.argtypes = [_f32p, _c_int, _c_int, _f32p, _c_int, _c_int, _u8p, _c_int]
    L.or_coarse_water.argtypes = [_u8p, _c_int, _c_int, _c_int, _c_int, _c_int, _u8p]
    L.or_level_shapes.argtypes = [_c_int, _c_int, _c_double, _c_double, _c_int, _i32p, _i32p, _f64p, _f64p]
    L.or_level_shapes.restype = _c_int
    L.or_relax_to_convergence.argtypes = [_c_int, _c_int, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p,
                                          _f32p, _c_double]
    L.or_solve.argtypes = [_c_int, _i32p, _i32p, _vpp, _vpp, _vpp, _vpp, _vpp, _vpp, _vpp, _f32p,
                           _c_double, _c_int, _c_int, ctypes.POINTER(_c_double)]
    L.or_solve.restype = _c_int
    L.or_solve64.argtypes = [_c_int, _i32p, _i32p, _vpp, _vpp, _vpp, _vpp, _vpp, _vpp, _vpp, _f64p,
                             _c_double, _c_int, _c_int, _c_int, ctypes.c_void_p, ctypes.POINTER(_c_double)]
    L.or_solve64.restype = _c_int
    L.or_velocities.argtypes = [_c_int, _c_int, _u8p, _f32p, _c_double, _c_double, _f32p, _f32p]
    L.or_water_on.argtypes = [_u8p, _c_int, _c_int, _c_int, _c_int, _u8p]
    L.or_water_bodies.argtypes = [_u8p, _c_int, _c_int, _i32p]
    L.or_water_bodies.restype = _c_int
    L.or_correct_body_means_flat.argtypes = [_c_int, _c_int, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p,
                                             _f32p, _i32p, _c_int]
    L.or_solve_bodies.argtypes = [_c_int, _i32p, _i32p, _vpp, _vpp, _vpp, _vpp, _vpp, _vpp, _vpp, _f32p,
                                  _c_double, _c_int, _c_int, ctypes.c_void_p, _c_int, _c_int,
                                  ctypes.POINTER(_c_double)]
    L.or_solve_bodies.restype = _c_int
    L.or_deformation_radius.argtypes = [_c_double, _c_double]
    L.or_deformation_radius.restype = _c_double
    L.or_heat_diffusivity.argtypes = [_c_double, _c_double]
    L.or_heat_diffusivity.restype = _c_double
    _vp = ctypes.c_void_p
    L.or_heat_stencil.argtypes = [_c_int, _c_int, _c_double, _c_double, _u8p, _f32p, _f32p, _c_double, _c_int,
                                  _c_double, _c_double, _vp, _vp, _f32p, _f32p, _f32p, _f32p, _f64p, _f32p]
    L.or_resample.argtypes = [_f32p, _c_int, _c_int, _c_int, _c_int, _f32p]
    L.or_coarse_mean.argtypes = [_f32p, _c_int, _c_int, _c_int, _c_int, _f32p]
    L.or_curl_forcing.argtypes = [_c_int, _c_int, _c_double, _c_double, _f64p, _f64p, ctypes.c_void_p, _f64p]
    _lib = L
    return L


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


class Stencil:
    """OceanStencil (OC:24-42) as flat numpy arrays: binary32 weights and forcing, binary64 centre."""

    def __init__(self, across, down, water, east, west, north, south, forcing, centre):
        self.across, self.down = int(across), int(down)
        n = self.across * self.down
        self.water = np.ascontiguousarray(water, dtype=np.uint8).reshape(n)
        self.east = np.ascontiguousarray(east, dtype=np.float32).reshape(n)
        self.west = np.ascontiguousarray(west, dtype=np.float32).reshape(n)
        self.north = np.ascontiguousarray(north, dtype=np.float32).reshape(n)
        self.south = np.ascontiguousarray(south, dtype=np.float32).reshape(n)
        self.forcing = np.ascontiguousarray(forcing, dtype=np.float32).reshape(n)
        self.centre = np.ascontiguousarray(centre, dtype=np.float64).reshape(n)

    @property
    def cells(self) -> int:
        return self.across * self.down

    def with_forcing(self, forcing) -> "Stencil":
        return Stencil(self.across, self.down, self.water, self.east, self.west, self.north, self.south,
                       forcing, self.centre)


def stencil(across, down, dx, dy, water, beta, drag, balance=None) -> Stencil:
    """OceanCirculation.stencil + withBalance (OC:161-231)."""
    n = across * down
    water = np.ascontiguousarray(water, dtype=np.uint8).reshape(n)
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    bal = np.zeros(n) if balance is None else np.ascontiguousarray(balance, dtype=np.float64).reshape(n)
    e, w, no, s, f = (np.empty(n, np.float32) for _ in range(5))
    c = np.empty(n, np.float64)
    lib().or_stencil(across, down, dx, dy, water, beta, drag, bal, e, w, no, s, c, f)
    return Stencil(across, down, water, e, w, no, s, f, c)


def with_balance(st: Stencil, balance) -> Stencil:
    """withBalance (OC:222-231)."""
    bal = np.asarray(balance, dtype=np.float64).reshape(st.cells)
    f = np.where(st.water != 0, (bal / st.centre).astype(np.float32), np.float32(0)).astype(np.float32)
    return st.with_forcing(f)


def balance_of(st: Stencil) -> np.ndarray:
    """balanceOf (OC:325-327)."""
    return st.forcing.astype(np.float64) * st.centre


def relax(st: Stencil, values, passes: int) -> np.ndarray:
    """relax (OC:241-257) on a copy of values."""
    x = np.array(values, dtype=np.float32).reshape(st.cells).copy()
    lib().or_relax(st.across, st.down, st.water, st.east, st.west, st.north, st.south, st.forcing, x, int(passes))
    return x


def residual(st: Stencil, values) -> np.ndarray:
    """residual (OC:292-315) of binary32 values."""
    x = np.ascontiguousarray(values, dtype=np.float32).reshape(st.cells)
    r = np.empty(st.cells)
    lib().or_residual(st.across, st.down, st.water, st.east, st.west, st.north, st.south, st.forcing,
                      st.centre, x, r)
    return r


def residual64(st: Stencil, values) -> np.ndarray:
    """The residual of binary64 values, same arithmetic order, the forcing promoted."""
    x = np.ascontiguousarray(values, dtype=np.float64).reshape(st.cells)
    r = np.empty(st.cells)
    lib().or_residual64(st.across, st.down, st.water, st.east, st.west, st.north, st.south, st.forcing,
                        st.centre, x, r)
    return r


def relax64(st: Stencil, values, passes: int) -> np.ndarray:
    x = np.array(values, dtype=np.float64).reshape(st.cells).copy()
    lib().or_relax64_flat(st.across, st.down, st.water, st.east, st.west, st.north, st.south, st.forcing,
                          x, int(passes))
    return x


def largest(v) -> float:
    """largest (OC:318-322)."""
    a = np.ascontiguousarray(v, dtype=np.float64)
    return float(lib().or_largest(a.size, a))


def restrict(fine, fa, fd, ca, cd) -> np.ndarray:
    """restrict (OC:334-346)."""
    out = np.empty(ca * cd)
    lib().or_restrict(np.ascontiguousarray(fine, dtype=np.float64), fa, fd, ca, cd, out)
    return out


def prolong_add(coarse, ca, cd, fine, fa, fd, water, pole_is_wall: bool) -> np.ndarray:
    """prolongAdd (OC:354-392) on a copy of fine."""
    out = np.array(fine, dtype=np.float32).reshape(fa * fd).copy()
    lib().or_prolong_add(np.ascontiguousarray(coarse, dtype=np.float32), ca, cd, out, fa, fd,
                         np.ascontiguousarray(water, dtype=np.uint8), int(pole_is_wall))
    return out


def coarse_water(fine, fa, fd, ca, cd, every_cell_water: bool) -> np.ndarray:
    """coarseWater (OC:404-423)."""
    out = np.empty(ca * cd, np.uint8)
    lib().or_coarse_water(np.ascontiguousarray(fine, dtype=np.uint8), fa, fd, ca, cd, int(every_cell_water), out)
    return out


def level_shapes(across, down, dx, dy, max_levels=24):
    """The (across, down, cell width, cell height) sequence levels walks (OC:434-462)."""
    a = np.empty(max_levels, np.int32)
    d = np.empty(max_levels, np.int32)
    w = np.empty(max_levels)
    h = np.empty(max_levels)
    n = lib().or_level_shapes(across, down, dx, dy, max_levels, a, d, w, h)
    return [(int(a[i]), int(d[i]), float(w[i]), float(h[i])) for i in range(n)]


def levels(finest: Stencil, dx, dy, every_cell_water: bool, coarser):
    """levels (OC:434-462): coarser(across, down, water) builds each coarse stencil."""
    out = [finest]
    shapes = level_shapes(finest.across, finest.down, dx, dy)
    water, a, d = finest.water, finest.across, finest.down
    for (na, nd, _, _) in shapes[1:]:
        water = coarse_water(water, a, d, na, nd, every_cell_water)
        a, d = na, nd
        out.append(coarser(a, d, water))
    return out


def relax_to_convergence(st: Stencil, values, tol) -> np.ndarray:
    """relaxToConvergence (OC:784-794) on a copy of values."""
    x = np.array(values, dtype=np.float32).reshape(st.cells).copy()
    lib().or_relax_to_convergence(st.across, st.down, st.water, st.east, st.west, st.north, st.south,
                                  st.forcing, st.centre, x, float(tol))
    return x


def _ladder_args(ladder):
    n = len(ladder)
    acr = np.array([s.across for s in ladder], np.int32)
    dwn = np.array([s.down for s in ladder], np.int32)

    def ptrs(name):
        arr = (ctypes.c_void_p * n)(*[getattr(s, name).ctypes.data for s in ladder])
        return ctypes.cast(arr, _vpp), arr

    keep = []
    out = [n, acr, dwn]
    for name in ("water", "east", "west", "north", "south", "forcing", "centre"):
        p, arr = ptrs(name)
        keep.append(arr)
        out.append(p)
    return out, keep


SOLVE, RESCUE, VCYCLE, KRYLOV = 0, 1, 2, 3


def solve_bodies(ladder, start, tol, pole_is_wall, mode, body_of=None, body_count=0, most_iterations=200):
    """solve / solveWithKrylovRescue / solveByKrylov with the finest grid's water bodies
    (OC:473-607, correctBodyMeans OC:709-731). Returns (values, cycles, relative_residual)."""
    args, keep = _ladder_args(ladder)
    x = np.array(start, dtype=np.float32).reshape(ladder[0].cells).copy()
    rr = ctypes.c_double(0.0)
    b = None if body_of is None else np.ascontiguousarray(body_of, dtype=np.int32)
    cycles = lib().or_solve_bodies(*args, x, float(tol), int(pole_is_wall), int(mode),
                                   None if b is None else b.ctypes.data_as(ctypes.c_void_p), int(body_count),
                                   int(most_iterations), ctypes.byref(rr))
    del keep
    if cycles < 0:
        raise ValueError("oracle solve rejected its arguments")
    return x, int(cycles), float(rr.value)


def solve(ladder, start, tol, pole_is_wall=True, mode=SOLVE):
    """solve / solveWithKrylovRescue / one vCycle / solveByKrylov (OC:473-607), processor relax.

    Returns (values, cycles, relative_residual)."""
    args, keep = _ladder_args(ladder)
    x = np.array(start, dtype=np.float32).reshape(ladder[0].cells).copy()
    rr = ctypes.c_double(0.0)
    cycles = lib().or_solve(*args, x, float(tol), int(pole_is_wall), int(mode), ctypes.byref(rr))
    del keep
    if cycles < 0:
        raise ValueError("oracle solve rejected its arguments")
    return x, int(cycles), float(rr.value)


def solve64(ladder, tol, pole_is_wall=True, max_cycles=400, coarsest_passes=400, start=None, history=False):
    """The binary64 twin of solve: the exact-arithmetic fixed point of the same float operator.

    Returns (values64, cycles, relative_residual[, history])."""
    args, keep = _ladder_args(ladder)
    x = np.zeros(ladder[0].cells) if start is None else np.array(start, dtype=np.float64).copy()
    hist = np.zeros(max_cycles)
    rr = ctypes.c_double(0.0)
    cycles = lib().or_solve64(*args, x, float(tol), int(pole_is_wall), int(max_cycles), int(coarsest_passes),
                              hist.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rr))
    del keep
    if history:
        return x, int(cycles), float(rr.value), hist[:cycles].copy()
    return x, int(cycles), float(rr.value)


def velocities(st: Stencil, stream, dx, dy):
    """velocities (OC:801-828)."""
    e = np.empty(st.cells, np.float32)
    n = np.empty(st.cells, np.float32)
    lib().or_velocities(st.across, st.down, st.water, np.ascontiguousarray(stream, dtype=np.float32), dx, dy, e, n)
    return e, n


def water_on(map_land, map_across, map_down, across, down):
    """OceanStage.waterOn (OS:551-573): the solve grid's mask from a map's land mask."""
    out = np.empty(across * down, np.uint8)
    land = np.ascontiguousarray(map_land, dtype=np.uint8).reshape(map_across * map_down)
    lib().or_water_on(land, map_across, map_down, across, down, out)
    return out


def curl_forcing(across, down, width_m, height_m, eastward, northward, depth=None):
    """OceanStage.curlForcing (OS:689-708); depth None takes the 500 m wind-driven layer."""
    n = across * down
    e = np.ascontiguousarray(eastward, dtype=np.float64).reshape(n)
    nn = np.ascontiguousarray(northward, dtype=np.float64).reshape(n)
    d = None if depth is None else np.ascontiguousarray(depth, dtype=np.float64).reshape(n)
    out = np.empty(n)
    lib().or_curl_forcing(across, down, float(width_m), float(height_m), e, nn,
                          None if d is None else d.ctypes.data_as(ctypes.c_void_p), out)
    return out


def water_bodies(water, across, down):
    """waterBodies (OC:670-697): (body_of int32, count)."""
    out = np.empty(across * down, np.int32)
    count = lib().or_water_bodies(np.ascontiguousarray(water, dtype=np.uint8).reshape(-1), across, down, out)
    return out, int(count)


def correct_body_means(st: Stencil, values, body_of, count):
    """correctBodyMeans (OC:709-731) on a copy of values."""
    x = np.array(values, dtype=np.float32).reshape(st.cells).copy()
    lib().or_correct_body_means_flat(st.across, st.down, st.water, st.east, st.west, st.north, st.south,
                                     st.forcing, st.centre, x, np.ascontiguousarray(body_of, dtype=np.int32),
                                     int(count))
    return x


def deformation_radius(latitude_deg, radius_m):
    """OceanHeat.deformationRadiusMeters (OceanHeat.kt:98-105)."""
    return float(lib().or_deformation_radius(float(latitude_deg), float(radius_m)))


def heat_diffusivity(latitude_deg, radius_m):
    """OceanHeat.diffusivity (OceanHeat.kt:128-129)."""
    return float(lib().or_heat_diffusivity(float(latitude_deg), float(radius_m)))


def heat_stencil(across, down, dx, dy, water, stream, target, relaxation_s, with_target, radius_m=0.0,
                 constant_k=0.0, entrainment=None, subsurface=None) -> Stencil:
    """OceanHeat.stencil (OceanHeat.kt:151-228); constant_k > 0 holds the diffusivity constant."""
    n = across * down
    e, w, nn, s = (np.empty(n, np.float32) for _ in range(4))
    c = np.empty(n, np.float64)
    f = np.empty(n, np.float32)
    ent = None if entrainment is None else np.ascontiguousarray(entrainment, dtype=np.float32).reshape(n)
    sub = None if subsurface is None else np.ascontiguousarray(subsurface, dtype=np.float32).reshape(n)
    water = np.ascontiguousarray(water, dtype=np.uint8).reshape(n)
    lib().or_heat_stencil(across, down, float(dx), float(dy), water,
                          np.ascontiguousarray(stream, dtype=np.float32).reshape(n),
                          np.ascontiguousarray(target, dtype=np.float32).reshape(n), float(relaxation_s),
                          int(bool(with_target)), float(radius_m), float(constant_k),
                          None if ent is None else ent.ctypes.data_as(ctypes.c_void_p),
                          None if sub is None else sub.ctypes.data_as(ctypes.c_void_p), e, w, nn, s, c, f)
    return Stencil(across, down, water, e, w, nn, s, f, c)


def resample(field, fa, fd, ta, td):
    """OceanStage.resample (OceanStage.kt:1199-1223)."""
    out = np.empty(ta * td, np.float32)
    lib().or_resample(np.ascontiguousarray(field, dtype=np.float32).reshape(-1), fa, fd, ta, td, out)
    return out


def coarse_mean(fine, fa, fd, ca, cd):
    """OceanStage.coarseMean (OceanStage.kt:857-860)."""
    out = np.empty(ca * cd, np.float32)
    lib().or_coarse_mean(np.ascontiguousarray(fine, dtype=np.float32).reshape(-1), fa, fd, ca, cd, out)
    return out
