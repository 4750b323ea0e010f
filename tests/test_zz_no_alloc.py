"""cg_ocean_relax_noalloc, the relaxation seam for callers that cannot allocate: device buffers in
and out, the library's static scratch in between. Held to cg_ocean_relax (the seam, itself
bit-identical to OceanCirculation.relax) to the bit on two calls in a row, the second on a smaller
grid, so the scratch's leftovers from the first call must not reach the second; a heat (per-cell
weights) stencil after them; grids beyond the scratch rejected before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

import cartogenesis_amd as cg
from cartogenesis_amd.ocean import OceanHeat
from cartogenesis_amd.problem import circulation_problem

pytestmark = pytest.mark.gpu


def noalloc(acc, st, start, passes, stream):
    n = st.cells
    f = torch.from_numpy(st.forcing.copy()).to("cuda")
    x = torch.from_numpy(np.ascontiguousarray(start, np.float32)).to("cuda")
    out = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = cg.lib().cg_ocean_relax_noalloc(acc.ctx, acc.holding(st), ctypes.c_void_p(f.data_ptr()),
                                         ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(passes),
                                         ctypes.c_void_p(stream))
    cg.check(rc, acc.ctx, "cg_ocean_relax_noalloc")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_relax_matches_the_seam_on_calls_of_shrinking_size(acc):
    stream = torch.cuda.current_stream().cuda_stream
    big = circulation_problem(1024, seed=3).levels[0]
    small = circulation_problem(256, seed=5, land_fraction=0.4).levels[0]
    rng = np.random.default_rng(0)
    for st, passes in ((big, 7), (small, 4), (small, 1)):
        start = np.where(st.is_water != 0, rng.normal(size=st.cells), 0.0).astype(np.float32)
        print(f"OCEAN relax_noalloc {st.cells_across}x{st.cells_down} passes {passes}", flush=True)
        want = acc.solve(st, start, passes)
        got = noalloc(acc, st, start, passes, stream)
        assert want is not None
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the context's own stream (NULL) as well, and zero passes (a copy)
    start = np.where(small.is_water != 0, 1.0, 0.0).astype(np.float32)
    assert np.array_equal(noalloc(acc, small, start, 3, None).view(np.uint32),
                          acc.solve(small, start, 3).view(np.uint32))
    assert np.array_equal(noalloc(acc, small, start, 0, stream), start)


def test_noalloc_relax_on_per_cell_weights(acc):
    p = circulation_problem(256, seed=1)
    st0 = p.levels[0]
    stream = np.where(st0.is_water != 0, np.random.default_rng(2).normal(size=st0.cells) * 1e4, 0.0).astype(np.float32)
    target = np.where(st0.is_water != 0, 15.0, 0.0).astype(np.float32)
    heat = OceanHeat.stencil(st0.cells_across, st0.cells_down, 40075e3 / st0.cells_across,
                             20037.5e3 / st0.cells_down, st0.is_water, stream, target, 30 * 86400.0, True, 6.371e6)
    start = np.where(heat.is_water != 0, 10.0, 0.0).astype(np.float32)
    want = acc.solve(heat, start, 5)
    got = noalloc(acc, heat, start, 5, torch.cuda.current_stream().cuda_stream)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_relax_rejects_grids_beyond_its_scratch(acc):
    st = circulation_problem(2048, seed=8).levels[0]
    handle = acc.holding(st)
    dummy = torch.zeros(8, dtype=torch.float32, device="cuda")
    p = ctypes.c_void_p(dummy.data_ptr())
    rc = cg.lib().cg_ocean_relax_noalloc(acc.ctx, handle, p, p, p, 2, None)
    assert rc == cg.CG_ERR_ARG
    assert b"scratch" in cg.lib().cg_ocean_last_error(acc.ctx)
