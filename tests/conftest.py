import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# torch bundles its own HIP runtime under the soname the library links (libamdhip64.so.7); one
# runtime serves a process, the first loaded. The learned preconditioner's tests use torch's, so
# torch is loaded before the library is: every test of the session then runs on that one runtime.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def acc():
    """The HIP accelerator on device 0; a GPU test fails (not skips) when there is none."""
    from cartogenesis_amd import ocean
    result = ocean.HipOcean.create_or_null(0)
    if result.accelerator is None:
        pytest.fail(f"no MI355X accelerator: {result.unavailable_because}")
    yield result.accelerator
    result.accelerator.close()


def to_oracle(st):
    import oracle
    return oracle.Stencil(st.cells_across, st.cells_down, st.is_water, st.east_weight, st.west_weight,
                          st.north_weight, st.south_weight, st.forcing, st.centre_weight)


# Reference test constants (OceanCirculationTest.kt:22-26, GpuOceanTest.kt:302-313).
STOMMEL_BETA = 6.61e-11
STOMMEL_DRAG = 9.91e-7
UNIT_ROUNDOFF = 1.0 / (1 << 24)
ROUNDED_OPERATIONS_PER_UPDATE = 8.0
BATCH_PASSES = 50


def stommel_basin(spacing, aspect, amplitude=-1e-12, drag=STOMMEL_DRAG, beta=STOMMEL_BETA):
    """OceanCirculationTest.basin (OceanCirculationTest.kt:42-60): a 3200 x 1600 km basin between
    two land columns, forced by amplitude sin(pi y / b). Returns (levels, across, down, dx, dy)."""
    from cartogenesis_amd.ocean import OceanCirculation as OC
    dx = spacing
    dy = spacing * aspect
    across = int(3_200e3 / dx)
    down = int(1_600e3 / dy)
    col = np.arange(across * down) % across
    water = ((col != 0) & (col != across - 1)).astype(np.uint8)
    b = down * dy
    row = np.arange(across * down) // across
    y = (down - 0.5 - row) * dy
    forcing = amplitude * np.sin(np.pi * y / b)
    finest = OC.stencil(across, down, dx, dy, water, np.full(down, beta), drag, forcing)

    def coarser(a, d, w):
        return OC.stencil(a, d, dx * across / a, dy * down / d, w, np.full(d, beta), drag, None)

    return OC.levels(finest, dx, dy, True, coarser), across, down, dx, dy


def stommel_analytic(across, down, dx, dy, amplitude, x, y, drag=STOMMEL_DRAG, beta=STOMMEL_BETA):
    """Stommel's closed form (OceanCirculationTest.kt:68-79)."""
    lam = (across - 1) * dx
    b = down * dy
    alpha = beta / drag
    k = np.pi / b
    particular = amplitude / drag / (k * k)
    mp = (-alpha + np.sqrt(alpha * alpha + 4 * k * k)) / 2
    mm = (-alpha - np.sqrt(alpha * alpha + 4 * k * k)) / 2
    p = particular * (1 - np.exp(mm * lam)) / (np.exp(mp * lam) - np.exp(mm * lam))
    q = particular - p
    return np.sin(k * y) * (-particular + p * np.exp(mp * x) + q * np.exp(mm * x))


def gpu_stencil_basin(across, down, dx, dy):
    """GpuOceanTest.circulationStencilOf (GpuOceanTest.kt:158-162)."""
    from cartogenesis_amd.ocean import OceanCirculation as OC
    from cartogenesis_amd.problem import bottom_drag
    col = np.arange(across * down) % across
    water = ((col != 0) & (col != across - 1)).astype(np.uint8)
    row = np.arange(across * down) // across
    forcing = -1e-12 * np.sin(np.pi * (row + 0.5) / down)
    return OC.stencil(across, down, dx, dy, water, np.full(down, 6.61e-11), bottom_drag(), forcing)


def batch_bound(values, passes):
    """GpuOceanTest.batchBound (GpuOceanTest.kt:216-219)."""
    return passes * ROUNDED_OPERATIONS_PER_UPDATE * UNIT_ROUNDOFF * float(np.abs(values).max())
