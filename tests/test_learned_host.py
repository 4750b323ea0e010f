"""Host-side checks of the learned preconditioner (no GPU): the network's shape table, the loss
gradient, the weight-gradient slicing and the convolution entry points' argument checks (which
return before anything is launched)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cartogenesis_amd import CG_ERR_ARG, PRECOND_FN, lib  # noqa: E402
from cartogenesis_amd.learned import layer_table, loss_and_grad  # noqa: E402


def test_layer_table_is_a_four_level_unet():
    table = layer_table(32)
    names = [t[0] for t in table]
    assert len(table) == 15 and names[0] == "enc0a" and names[-1] == "out"
    spec = {name: (c1, c2, co) for name, c1, c2, co, _ in table}
    # each layer reads what the previous one wrote; decoders also read their level's encoder output
    assert spec["enc0a"] == (16, 0, 32) and spec["mid_b"] == (256, 0, 256)
    assert spec["dec2a"] == (256, 128, 128) and spec["dec1a"] == (128, 64, 64) and spec["dec0a"] == (64, 32, 32)
    assert spec["out"] == (32, 0, 16) and [t[4] for t in table].count(0) == 1
    for name, c1, c2, co, _ in table:
        assert c1 % 16 == 0 and c2 % 16 == 0 and co % 16 == 0
    # 6.6e11 flops per application at 1024 x 2048 (two per multiply-add)
    level = {"enc0": 0, "enc1": 1, "enc2": 2, "mid_": 3, "dec2": 2, "dec1": 1, "dec0": 0, "out": 0}
    flops = sum(2 * 9 * (c1 + c2) * co * (1024 * 2048 >> (2 * level[n[:4]])) for n, c1, c2, co, _ in table)
    assert 6.5e11 < flops < 7.3e11


def test_loss_gradient_matches_autograd():
    g = torch.Generator().manual_seed(0)
    out = torch.randn(2, 8, 16, 16, generator=g, requires_grad=True)
    target = torch.randn(2, 8, 16, generator=g)
    water = (torch.rand(2, 8, 16, generator=g) > 0.3).float()
    loss, dout = loss_and_grad(out.detach(), target, water)
    ref = (((out[..., 0] - target) * water) ** 2).sum() / water.sum()
    ref.backward()
    assert torch.allclose(loss, ref.detach(), rtol=1e-6)
    assert torch.allclose(dout, out.grad, rtol=1e-6, atol=1e-9)


def test_wgrad_slices_are_multiples_of_four_within_bounds():
    for n, d, a in [(1, 8, 16), (4, 256, 512), (1, 1024, 2048), (2, 32, 64)]:
        s = lib().cg_conv3x3_wgrad_slices(n, d, a)
        assert s % 4 == 0 and 4 <= s <= 256


def test_convolution_entry_points_refuse_bad_shapes():
    L = lib()
    p = ctypes.c_void_p(16)  # never dereferenced: each call fails its checks first
    # channels not multiples of 16, across not a multiple of 16, missing second input, bad split
    assert L.cg_conv3x3(p, 24, None, 0, 1, 8, 32, p, 32, 1, p, 32, None, None, None) == CG_ERR_ARG
    assert L.cg_conv3x3(p, 32, None, 0, 1, 8, 40, p, 32, 1, p, 32, None, None, None) == CG_ERR_ARG
    assert L.cg_conv3x3(p, 32, None, 16, 1, 8, 32, p, 32, 1, p, 32, None, None, None) == CG_ERR_ARG
    assert L.cg_conv3x3(p, 32, None, 0, 1, 8, 32, p, 32, 1, p, 16, None, None, None) == CG_ERR_ARG
    assert L.cg_conv3x3_adjoint(p, 32, 1, 8, 32, p, 48, p, 8, p, None, None) == CG_ERR_ARG
    assert L.cg_conv3x3_wgrad(p, 32, None, 0, p, 30, 1, 8, 32, p, 1 << 20, p, None) == CG_ERR_ARG
    assert L.cg_pool2(p, 16, 1, 7, 32, 1.0, p, None) == CG_ERR_ARG
    assert L.cg_unet_pack_input(None, p, 1.0, 16, p, None) == CG_ERR_ARG
    assert L.cg_ocean_plan_set_preconditioner(None, PRECOND_FN(), None) == CG_ERR_ARG
    assert np.int32(L.cg_relu_mask(p, p, 0, None)) == 0  # nothing to do, nothing launched
