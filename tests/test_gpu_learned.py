"""The learned preconditioner's kernels against plain PyTorch fp32 on the CPU, the whole network's
gradients against CPU autograd, and the trained network inside FGMRES against the method without it.

The convolution kernels read bfloat16 and accumulate in binary32, so the CPU reference gets the
same bfloat16-rounded inputs in float32: the binary32 outputs then differ only by summation order
(tolerance 1e-5 of the largest magnitude). The network test rounds every activation to bfloat16 on
the device and not on the CPU, so its gradients are held to a cosine of 0.99 and norms within 5%.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
F = torch.nn.functional

pytestmark = pytest.mark.gpu

from cartogenesis_amd import CG_METHOD_FGMRES, CG_METHOD_LEARNED_FGMRES, check, lib  # noqa: E402


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def cpu_conv(x_nhwc, w9, relu=False):
    """y = conv3x3 with rows zero-padded and columns wrapping, NHWC in and out, weights [9][co][ci]."""
    co, ci = w9.shape[1], w9.shape[2]
    x = F.pad(x_nhwc.permute(0, 3, 1, 2), (1, 1, 0, 0), mode="circular")
    y = F.conv2d(x, w9.reshape(3, 3, co, ci).permute(2, 3, 0, 1), padding=(1, 0))
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 1)


def rand_bf16(*shape, seed=0, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(*shape, generator=g) * scale).to(torch.bfloat16)


@pytest.mark.parametrize("c1,c2,cout,relu,a,d", [(32, 16, 64, 1, 48, 8), (16, 0, 32, 0, 48, 8), (64, 32, 16, 1, 48, 8),
                                                 (32, 64, 96, 0, 48, 8), (32, 32, 64, 1, 128, 8),
                                                 (64, 0, 32, 0, 192, 8), (32, 32, 64, 1, 128, 6)])
def test_conv_matches_cpu(c1, c2, cout, relu, a, d):
    """Channel counts of 16 take the K = 16 kernel; multiples of 32 the K = 32 one reading L1/L2 (one
    16-pixel tile per wavefront below 64 columns, four from 64): these images have too few 4 x 64
    tiles for the LDS-staged kernel, which the next test covers."""
    n = 2
    x1, x2 = rand_bf16(n, d, a, c1, seed=1), rand_bf16(n, d, a, c2, seed=2) if c2 else None
    w = rand_bf16(9, cout, c1 + c2, seed=3, scale=0.1)
    # device copies held in names: a temporary's memory could be handed to the next one
    x1d, x2d, wd = x1.cuda(), x2.cuda() if c2 else None, w.cuda()
    y = torch.empty(n, d, a, cout, device="cuda")
    check(lib().cg_conv3x3(_p(x1d), c1, _p(x2d), c2, n, d, a, _p(wd), cout, relu, None, 0, None, _p(y), None),
          what="cg_conv3x3")
    xin = torch.cat([x1, x2], -1) if c2 else x1
    ref = cpu_conv(xin.float(), w.float(), bool(relu))
    got = y.cpu()
    assert (got - ref).abs().max() <= 1e-5 * ref.abs().max()
    # bfloat16 output, split across two tensors, rounds the same values
    o1 = 16
    ya = torch.empty(n, d, a, o1, dtype=torch.bfloat16, device="cuda")
    yb = torch.empty(n, d, a, cout - o1, dtype=torch.bfloat16, device="cuda") if cout > o1 else None
    check(lib().cg_conv3x3(_p(x1d), c1, _p(x2d), c2, n, d, a, _p(wd), cout, relu, _p(ya), o1, _p(yb), None, None),
          what="cg_conv3x3 bf16")
    both = torch.cat([ya.cpu(), yb.cpu()], -1) if yb is not None else ya.cpu()
    assert torch.equal(both, got.to(torch.bfloat16))


@pytest.mark.parametrize("c1,c2,cout,up1,n", [(64, 32, 32, 0, 4), (32, 0, 16, 0, 4), (64, 32, 128, 1, 2),
                                               (32, 0, 64, 0, 2)])
def test_staged_conv_walks_many_tiles_per_workgroup(c1, c2, cout, up1, n):
    """Layers with at least as many 4 x 64 tiles (times output-channel blocks) as the device has CUs
    take the LDS-staged kernel, whose grid holds only the workgroups that stay resident: at 64 x 1024
    with n images each workgroup walks several tiles, several 32-channel chunks per tile (96 input
    channels: the first input's two chunks, then the skip input's), or one chunk whose weights it
    stages once (32 input channels), at 16, 32, 64 and 128 output channels (one or two channel
    blocks), and with up1 reads the first input through nearest upsampling. Against the CPU
    convolution; the bfloat16 outputs split over two tensors round the binary32 ones."""
    d, a = 64, 1024
    x1 = rand_bf16(n, d >> up1, a >> up1, c1, seed=21)
    x2 = rand_bf16(n, d, a, c2, seed=22) if c2 else None
    w = rand_bf16(9, cout, c1 + c2, seed=23, scale=0.1)
    x1d, x2d, wd = x1.cuda(), x2.cuda() if c2 else None, w.cuda()
    conv = lib().cg_conv3x3_up if up1 else lib().cg_conv3x3
    y = torch.empty(n, d, a, cout, device="cuda")
    print(f"OCEAN staged conv {c1}+{c2}->{cout} up1={up1} n={n}", flush=True)
    check(conv(_p(x1d), c1, _p(x2d), c2, n, d, a, _p(wd), cout, 1, None, 0, None, _p(y), None), what="staged conv")
    x1f = x1.float().repeat_interleave(2, 1).repeat_interleave(2, 2) if up1 else x1.float()
    xin = torch.cat([x1f, x2.float()], -1) if c2 else x1f
    ref = cpu_conv(xin, w.float(), True)
    got = y.cpu()
    assert (got - ref).abs().max() <= 1e-5 * ref.abs().max()
    o1 = 16
    ya = torch.empty(n, d, a, o1, dtype=torch.bfloat16, device="cuda")
    yb = torch.empty(n, d, a, cout - o1, dtype=torch.bfloat16, device="cuda") if cout > o1 else None
    check(conv(_p(x1d), c1, _p(x2d), c2, n, d, a, _p(wd), cout, 1, _p(ya), o1, _p(yb), None, None),
          what="staged conv bf16")
    both = torch.cat([ya.cpu(), yb.cpu()], -1) if yb is not None else ya.cpu()
    assert torch.equal(both, got.to(torch.bfloat16))


@pytest.mark.parametrize("cout,relu,junk", [(32, 1, False), (16, 0, False), (24, 1, False), (32, 1, True)])
def test_first_layer_kernel_matches_cpu(cout, relu, junk):
    """cg_conv3x3_in3 (channels 0..2 of a 16-channel input) against the CPU convolution: binary32
    sums of 27 products, rounded once to bfloat16. Output channel counts that are multiples of 16
    take the matrix-core kernel (16 pixels a wavefront; 40 columns, so tiles cross rows), 24 the
    vector one. With junk, channels 3..15 hold large values and NaN, which must not reach the output."""
    n, d, a = 2, 8, 40
    x = torch.zeros(n, d, a, 16)
    if junk:
        x[..., 3:] = 1e30
        x[..., 3] = float("nan")
    x[..., :3] = torch.randn(n, d, a, 3, generator=torch.Generator().manual_seed(15))
    x = x.to(torch.bfloat16)
    w = rand_bf16(9, cout, 16, seed=16, scale=0.2)
    xd, wd = x.cuda(), w.cuda()
    y = torch.empty(n, d, a, cout, dtype=torch.bfloat16, device="cuda")
    check(lib().cg_conv3x3_in3(_p(xd), n, d, a, _p(wd), cout, relu, _p(y), None), what="cg_conv3x3_in3")
    xr = x.float()
    xr[..., 3:] = 0.0
    ref = cpu_conv(xr, w.float(), bool(relu))
    got = y.cpu().float()
    assert ((got - ref).abs() <= 2.0 ** -8 * ref.abs() + 1e-6 * ref.abs().max()).all()
    p = ctypes.c_void_p(16)
    assert lib().cg_conv3x3_in3(p, 1, 8, 32, p, 48, 1, p, None) < 0 and lib().cg_conv3x3_in3(p, 1, 8, 32, p, 12, 1, p,
                                                                                               None) < 0


def test_conv_adjoint_and_weight_gradient_match_cpu_autograd():
    n, d, a, c1, c2, cout = 2, 8, 32, 32, 16, 64
    x = rand_bf16(n, d, a, c1 + c2, seed=4)
    w = rand_bf16(9, cout, c1 + c2, seed=5, scale=0.1)
    dy = rand_bf16(n, d, a, cout, seed=6)
    xr = x.float().requires_grad_(True)
    wr = w.float().requires_grad_(True)
    (cpu_conv(xr, wr) * dy.float()).sum().backward()
    xd, wd, dyd = x.cuda(), w.cuda(), dy.cuda()
    dx = torch.empty(n, d, a, c1 + c2, device="cuda")
    check(lib().cg_conv3x3_adjoint(_p(dyd), cout, n, d, a, _p(wd), c1 + c2, None, 0, None, _p(dx), None),
          what="cg_conv3x3_adjoint")
    assert (dx.cpu() - xr.grad).abs().max() <= 1e-5 * xr.grad.abs().max()
    # split bfloat16 output, as the U-Net's skip concatenation uses it
    g1 = torch.empty(n, d, a, c1, dtype=torch.bfloat16, device="cuda")
    g2 = torch.empty(n, d, a, c2, dtype=torch.bfloat16, device="cuda")
    check(lib().cg_conv3x3_adjoint(_p(dyd), cout, n, d, a, _p(wd), c1 + c2, _p(g1), c1, _p(g2), None, None),
          what="cg_conv3x3_adjoint split")
    assert torch.equal(torch.cat([g1, g2], -1).cpu(), dx.cpu().to(torch.bfloat16))
    # the same gradient from the forward kernels with the mirrored, transposed weights
    wt = torch.empty(9, c1 + c2, cout, dtype=torch.bfloat16, device="cuda")
    check(lib().cg_conv3x3_flip_weights(_p(wd), cout, c1 + c2, _p(wt), None), what="cg_conv3x3_flip_weights")
    assert torch.equal(wt.cpu(), w.reshape(3, 3, cout, c1 + c2).flip(0, 1).reshape(9, cout, c1 + c2).transpose(1, 2))
    for across in (a, 128):  # the K = 32 kernel with one 16-pixel tile a wavefront, then the staged one
        dyw = rand_bf16(n, d, across, cout, seed=9).cuda()
        xw = rand_bf16(n, d, across, c1 + c2, seed=10).float().requires_grad_(True)
        (cpu_conv(xw, w.float()) * dyw.cpu().float()).sum().backward()
        dxf = torch.empty(n, d, across, c1 + c2, device="cuda")
        check(lib().cg_conv3x3(_p(dyw), cout, None, 0, n, d, across, _p(wt), c1 + c2, 0, None, 0, None, _p(dxf), None),
              what="cg_conv3x3 flipped")
        assert (dxf.cpu() - xw.grad).abs().max() <= 1e-5 * xw.grad.abs().max()
    # weight gradient, with the input split over two tensors
    slices = lib().cg_conv3x3_wgrad_slices(n, d, a)
    work = torch.empty(slices * 9 * cout * (c1 + c2), device="cuda")
    dw = torch.empty(9, cout, c1 + c2, device="cuda")
    x1d, x2d = xd[..., :c1].contiguous(), xd[..., c1:].contiguous()
    check(lib().cg_conv3x3_wgrad(_p(x1d), c1, _p(x2d), c2, _p(dyd), cout, n, d, a, _p(work), work.numel(), _p(dw),
                                 None), what="cg_conv3x3_wgrad")
    assert (dw.cpu() - wr.grad).abs().max() <= 1e-5 * wr.grad.abs().max()
    # rows that are not a multiple of 32 wide take the two-byte-gather kernel: the same sums
    a48 = 48
    x48 = rand_bf16(n, d, a48, c1 + c2, seed=13)
    dy48 = rand_bf16(n, d, a48, cout, seed=14)
    w48 = w.float().requires_grad_(True)
    (cpu_conv(x48.float(), w48) * dy48.float()).sum().backward()
    x48d, dy48d = x48.cuda(), dy48.cuda()
    x481, x482 = x48d[..., :c1].contiguous(), x48d[..., c1:].contiguous()
    s48 = lib().cg_conv3x3_wgrad_slices(n, d, a48)
    work48 = torch.empty(s48 * 9 * cout * (c1 + c2), device="cuda")
    dw48 = torch.empty(9, cout, c1 + c2, device="cuda")
    check(lib().cg_conv3x3_wgrad(_p(x481), c1, _p(x482), c2, _p(dy48d), cout, n, d, a48, _p(work48), work48.numel(),
                                 _p(dw48), None), what="cg_conv3x3_wgrad 48")
    assert (dw48.cpu() - w48.grad).abs().max() <= 1e-5 * w48.grad.abs().max()
    # too small a workspace is refused before anything runs
    assert lib().cg_conv3x3_wgrad(_p(x1d), c1, _p(x2d), c2, _p(dyd), cout, n, d, a, _p(work), 16, _p(dw), None) < 0


def test_pool_upsample_relu_add_are_exact():
    n, d, a, ch = 2, 8, 32, 16
    x = rand_bf16(n, d, a, ch, seed=7)
    xd = x.cuda()
    y = torch.empty(n, d // 2, a // 2, ch, dtype=torch.bfloat16, device="cuda")
    check(lib().cg_pool2(_p(xd), ch, n, d, a, 1.0, _p(y), None), what="cg_pool2")
    ref = F.avg_pool2d(x.float().permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
    assert ((y.cpu().float() - ref).abs() <= 2.0 ** -8 * ref.abs()).all()  # one rounding to bfloat16
    u = torch.empty(n, d, a, ch, dtype=torch.bfloat16, device="cuda")
    check(lib().cg_upsample2(_p(y), ch, n, d // 2, a // 2, 1.0, _p(u), None), what="cg_upsample2")
    assert torch.equal(u.cpu(), y.cpu().repeat_interleave(2, 1).repeat_interleave(2, 2))
    g = rand_bf16(n, d, a, ch, seed=8).cuda()
    g0 = g.clone()
    check(lib().cg_relu_mask(_p(g), _p(xd), g.numel(), None), what="cg_relu_mask")
    assert torch.equal(g.cpu(), torch.where(x.float() > 0, g0.cpu(), torch.zeros_like(g0.cpu())))
    s = torch.empty_like(g)
    check(lib().cg_add_bf16(_p(g), _p(xd), g.numel(), _p(s), None), what="cg_add_bf16")
    assert torch.equal(s.cpu(), (g.cpu().float() + x.float()).to(torch.bfloat16))


def cpu_unet(net, x):
    """The network of cartogenesis_amd.learned in float32 on the CPU, weights rounded to bfloat16."""
    w = {k: v.detach().cpu().to(torch.bfloat16).float().requires_grad_(True) for k, v in net.w32.items()}
    pool = lambda t: F.avg_pool2d(t.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)  # noqa: E731
    up = lambda t: t.repeat_interleave(2, 1).repeat_interleave(2, 2)  # noqa: E731
    c = lambda name, t, r=True: cpu_conv(t, w[name], r)  # noqa: E731
    e0 = c("enc0b", c("enc0a", x))
    e1 = c("enc1b", c("enc1a", pool(e0)))
    e2 = c("enc2b", c("enc2a", pool(e1)))
    m = c("mid_b", c("mid_a", pool(e2)))
    d2 = c("dec2b", c("dec2a", torch.cat([up(m), e2], -1)))
    d1 = c("dec1b", c("dec1a", torch.cat([up(d2), e1], -1)))
    d0 = c("dec0b", c("dec0a", torch.cat([up(d1), e0], -1)))
    return c("out", d0, False), w


def test_unet_gradients_match_cpu_autograd():
    from cartogenesis_amd.learned import UNet, loss_and_grad

    net = UNet(base=16, seed=3)
    n, d, a = 2, 16, 128
    g = torch.Generator().manual_seed(11)
    x = torch.zeros(n, d, a, 16)
    x[..., 0] = (torch.rand(n, d, a, generator=g) > 0.3).float()
    x[..., 1] = torch.rand(n, d, a, generator=g) * x[..., 0]
    x[..., 2] = torch.randn(n, d, a, generator=g) * x[..., 0]
    target = torch.randn(n, d, a, generator=g)
    water = x[..., 0].clone()
    xb = x.to(torch.bfloat16)
    out, acts = net.forward(xb.cuda(), keep=True)
    loss, dout = loss_and_grad(out, target.cuda(), water.cuda())
    work = torch.empty(net.workspace_floats(n, d, a), device="cuda")
    grads = net.backward(dout, acts, work)
    torch.cuda.synchronize()
    ref_out, w = cpu_unet(net, xb.float())
    ref_loss, _ = loss_and_grad(ref_out, target, water)
    ref_loss.backward()
    assert abs(float(loss) - float(ref_loss.detach())) <= 0.02 * float(ref_loss.detach())
    for name, *_ in net.layers:
        got, ref = grads[name].cpu().flatten(), w[name].grad.flatten()
        cos = float(torch.dot(got, ref) / (got.norm() * ref.norm()))
        assert cos > 0.99, (name, cos)
        assert abs(float(got.norm() / ref.norm()) - 1.0) < 0.05, name


def test_forward_without_keep_reads_the_decoder_input_through_the_upsampling():
    """The inference forward (the decoder convolutions read their coarse input through nearest
    upsampling) gives the training forward's output to the bit; cg_conv3x3_up refuses shapes its
    staged kernel does not serve."""
    from cartogenesis_amd.learned import UNet

    net = UNet(base=32, seed=5)
    x = rand_bf16(2, 16, 512, 16, seed=12).cuda()
    fused = net.forward(x)
    kept, acts = net.forward(x, keep=True)
    assert "u0" in acts  # training forms the upsampled tensor for the weight gradient
    assert torch.equal(fused.cpu(), kept.cpu())
    p = ctypes.c_void_p(16)
    assert lib().cg_conv3x3_up(p, 32, p, 16, 1, 16, 512, p, 32, 1, p, 32, None, None, None) < 0  # c2 % 32
    assert lib().cg_conv3x3_up(p, 32, p, 32, 1, 16, 96, p, 32, 1, p, 32, None, None, None) < 0  # across % 64


def test_learned_preconditioner_inside_fgmres():
    """Trains the network briefly on this library's solves of sixteen worlds (256 x 512) and solves
    another with FGMRES, with and without it: both reach 1e-8, and the same answer; prints the gain."""
    from cartogenesis_amd.learned import LearnedPreconditioner, UNet, train, training_pairs
    from cartogenesis_amd.ocean import HipOcean, Plan
    from cartogenesis_amd.problem import circulation_problem

    acc = HipOcean.create_or_null(0).accelerator
    try:
        data = training_pairs(acc, 256, seeds=range(100, 116))
        net = UNet(base=32, seed=0)
        losses = train(net, data, steps=300, batch=4, lr=1e-3)
        assert np.isfinite(losses).all() and np.mean(losses[-20:]) < 0.5 * np.mean(losses[:5]), losses[::10]
        prob = circulation_problem(256, seed=150)
        plan = Plan(acc, prob.levels)
        try:
            plain = plan.solve(1e-8, CG_METHOD_FGMRES)
            x_plain, _ = plan.read()
            pre = LearnedPreconditioner(net, prob.levels[0], prob.depth)
            pre.attach(plan)
            learned = plan.solve(1e-8, CG_METHOD_LEARNED_FGMRES)
            assert pre.error is None, pre.error
            x_learned, _ = plan.read()
            LearnedPreconditioner.detach(plan)
        finally:
            plan.close()
        print(f"\nlearned: loss {losses[0]:.3g} -> {np.mean(losses[-20:]):.3g}; FGMRES iterations plain "
              f"{plain.cycles}, with the network {learned.cycles} ({pre.applications} network applications)")
        assert plain.relative_residual < 1e-8 and learned.relative_residual < 1e-8
        assert pre.applications > 0
        rel = np.linalg.norm(x_learned - x_plain) / np.linalg.norm(x_plain)
        assert rel < 1e-6
    finally:
        acc.close()


def test_learned_method_errors_are_reported():
    """CG_METHOD_LEARNED_FGMRES without a preconditioner is refused; a preconditioner that returns an error
    fails the solve with it (both forced from the host); the plan then solves as before."""
    from cartogenesis_amd import CG_METHOD_REFINED, PRECOND_FN, CgError
    from cartogenesis_amd.ocean import HipOcean, Plan
    from cartogenesis_amd.problem import circulation_problem

    acc = HipOcean.create_or_null(0).accelerator
    try:
        prob = circulation_problem(256, seed=151)
        plan = Plan(acc, prob.levels)
        try:
            with pytest.raises(CgError, match="needs cg_ocean_plan_set_preconditioner"):
                plan.solve(1e-8, CG_METHOD_LEARNED_FGMRES)
            calls = []

            def failing(user, v, z, stream):
                calls.append(1)
                return -1

            cb = PRECOND_FN(failing)
            check(lib().cg_ocean_plan_set_preconditioner(plan.handle, cb, None), acc.ctx, "set_preconditioner")
            with pytest.raises(CgError, match="learned preconditioner returned -1"):
                plan.solve(1e-8, CG_METHOD_LEARNED_FGMRES)
            assert calls == [1]
            check(lib().cg_ocean_plan_set_preconditioner(plan.handle, PRECOND_FN(), None), acc.ctx,
                  "set_preconditioner")
            info = plan.solve(1e-8, CG_METHOD_REFINED)
            assert info.relative_residual < 1e-8
        finally:
            plan.close()
    finally:
        acc.close()
