"""The learned part of the solve: a 4-level U-Net of 3x3 convolutions (32 base channels) that maps
(mask, depth, residual) to a correction. Two uses: the learned coarse-grid correction of
CG_METHOD_LEARNED (the network on a coarse grid of the refined method's cycle, in place of the
grids below it; trained here on the forcings that grid sees in plain solves and their exact
solutions; run natively inside the plan's step graphs), and a preconditioner inside flexible GMRES
(CG_METHOD_LEARNED_FGMRES, a Python callback on the finest grid).

Every convolution, forward and backward, and every pool, upsample and ReLU gradient runs in the
library's kernels (cg_conv3x3, cg_conv3x3_adjoint, cg_conv3x3_wgrad, ...): bfloat16 activations,
binary32 accumulation on the matrix cores. torch holds the device memory, the loss and the Adam
update. Not a reference component: the reference's solve has no learned part; its training pairs
come from this library's own multigrid solves of the synthetic worlds.
"""
from __future__ import annotations

import ctypes
import math
import sys
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import CG_METHOD_REFINED, PRECOND_FN, check, lib

BF16 = torch.bfloat16
DEPTH_SCALE = 6000.0  # metres: depth enters the network as H / 6000


def _p(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def layer_table(base: int = 32) -> List[Tuple[str, int, int, int, int]]:
    """(name, channels from the first input, channels from the skip input, output channels, relu)."""
    b1, b2, b4, b8 = base, 2 * base, 4 * base, 8 * base
    return [
        ("enc0a", 16, 0, b1, 1), ("enc0b", b1, 0, b1, 1),
        ("enc1a", b1, 0, b2, 1), ("enc1b", b2, 0, b2, 1),
        ("enc2a", b2, 0, b4, 1), ("enc2b", b4, 0, b4, 1),
        ("mid_a", b4, 0, b8, 1), ("mid_b", b8, 0, b8, 1),
        ("dec2a", b8, b4, b4, 1), ("dec2b", b4, 0, b4, 1),
        ("dec1a", b4, b2, b2, 1), ("dec1b", b2, 0, b2, 1),
        ("dec0a", b2, b1, b1, 1), ("dec0b", b1, 0, b1, 1),
        ("out", b1, 0, 16, 0),
    ]


class UNet:
    """Weights in binary32 (the trained copy) and bfloat16 (what the kernels read), layout
    [9][cout][cin]. Inputs are NHWC bfloat16 with 16 channels; the output is NHWC binary32 with 16
    channels of which channel 0 is the correction."""

    def __init__(self, base: int = 32, seed: int = 0, device: str = "cuda"):
        self.base = base
        self.device = device
        self.layers = layer_table(base)
        self.spec = {name: (c1, c2, co, relu) for name, c1, c2, co, relu in self.layers}
        g = torch.Generator().manual_seed(seed)
        self.w32: Dict[str, torch.Tensor] = {}
        for name, c1, c2, co, _ in self.layers:
            cin = c1 + c2
            std = math.sqrt((1.0 if name == "out" else 2.0) / (9 * cin))
            self.w32[name] = (torch.randn(9, co, cin, generator=g) * std).to(device).requires_grad_(False)
        self.wb = {k: v.to(BF16).contiguous() for k, v in self.w32.items()}

    def refresh(self) -> None:
        """Rounds the binary32 weights into the bfloat16 copies the kernels read."""
        for k, v in self.w32.items():
            self.wb[k].copy_(v)

    def state(self) -> Dict[str, np.ndarray]:
        return {k: v.detach().cpu().numpy() for k, v in self.w32.items()}

    def load(self, state: Dict[str, np.ndarray]) -> None:
        for k, v in state.items():
            self.w32[k].copy_(torch.from_numpy(np.asarray(v, np.float32)))
        self.refresh()

    # ---- forward -------------------------------------------------------------------------------
    def _conv(self, name, a1, a2, n, d, a, stream, out32=False):
        c1, c2, co, relu = self.spec[name]
        if out32:
            y = torch.empty((n, d, a, co), dtype=torch.float32, device=self.device)
            check(lib().cg_conv3x3(_p(a1), c1, _p(a2), c2, n, d, a, _p(self.wb[name]), co, relu, None, 0, None,
                                   _p(y), stream), what=f"cg_conv3x3 {name}")
        else:
            y = torch.empty((n, d, a, co), dtype=BF16, device=self.device)
            check(lib().cg_conv3x3(_p(a1), c1, _p(a2), c2, n, d, a, _p(self.wb[name]), co, relu, _p(y), co, None,
                                   None, stream), what=f"cg_conv3x3 {name}")
        return y

    def _pool(self, x, n, d, a, stream, scale=1.0):
        ch = x.shape[-1]
        y = torch.empty((n, d // 2, a // 2, ch), dtype=BF16, device=self.device)
        check(lib().cg_pool2(_p(x), ch, n, d, a, float(scale), _p(y), stream), what="cg_pool2")
        return y

    def _up(self, x, n, d, a, stream, scale=1.0):
        """x is d x a; the result 2d x 2a."""
        ch = x.shape[-1]
        y = torch.empty((n, 2 * d, 2 * a, ch), dtype=BF16, device=self.device)
        check(lib().cg_upsample2(_p(x), ch, n, d, a, float(scale), _p(y), stream), what="cg_upsample2")
        return y

    def forward(self, x: torch.Tensor, stream=None, keep: bool = False):
        """x: (n, down, across, 16) bfloat16; down a multiple of 8 and across of 128; channels 3..15
        must be zero (the first layer reads channels 0..2 only: mask, depth, vector, as pack_input
        writes them). Returns the (n, down, across, 16) binary32 output, and with keep the
        activations for backward()."""
        n, d, a, _ = x.shape
        if d % 8 or a % 128:
            raise ValueError("the U-Net needs down % 8 == 0 and across % 128 == 0")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        t = {"x": x}
        dims = [(d >> k, a >> k) for k in range(4)]
        c1, _, co0, relu0 = self.spec["enc0a"]
        if c1 == 16 and co0 <= 32 and co0 % 8 == 0:  # the input's channels 3..15 are zero by construction
            t["enc0a"] = torch.empty((n, d, a, co0), dtype=BF16, device=self.device)
            check(lib().cg_conv3x3_in3(_p(x), n, d, a, _p(self.wb["enc0a"]), co0, relu0, _p(t["enc0a"]), s),
                  what="cg_conv3x3_in3 enc0a")
        else:
            t["enc0a"] = self._conv("enc0a", x, None, n, *dims[0], s)
        t["enc0b"] = self._conv("enc0b", t["enc0a"], None, n, *dims[0], s)
        t["p1"] = self._pool(t["enc0b"], n, *dims[0], s)
        t["enc1a"] = self._conv("enc1a", t["p1"], None, n, *dims[1], s)
        t["enc1b"] = self._conv("enc1b", t["enc1a"], None, n, *dims[1], s)
        t["p2"] = self._pool(t["enc1b"], n, *dims[1], s)
        t["enc2a"] = self._conv("enc2a", t["p2"], None, n, *dims[2], s)
        t["enc2b"] = self._conv("enc2b", t["enc2a"], None, n, *dims[2], s)
        t["p3"] = self._pool(t["enc2b"], n, *dims[2], s)
        t["mid_a"] = self._conv("mid_a", t["p3"], None, n, *dims[3], s)
        t["mid_b"] = self._conv("mid_b", t["mid_a"], None, n, *dims[3], s)
        def decoder(name, coarse, skip, up, lvl):
            """The decoder convolution of level lvl on (upsampled coarse, skip). Without keep it reads
            the coarse tensor through the upsampling itself where the staged kernel serves the shape
            and the image has at least 65536 pixels (below, that kernel's few workgroups are slower
            than upsampling first); backward needs the upsampled tensor, so with keep it is formed."""
            c1, c2, co, relu = self.spec[name]
            d_, a_ = dims[lvl]
            if not keep and c1 % 32 == 0 and c2 % 32 == 0 and a_ % 64 == 0 and d_ % 4 == 0 and n * d_ * a_ >= 65536:
                y = torch.empty((n, d_, a_, co), dtype=BF16, device=self.device)
                check(lib().cg_conv3x3_up(_p(coarse), c1, _p(skip), c2, n, d_, a_, _p(self.wb[name]), co, relu, _p(y),
                                          co, None, None, s), what=f"cg_conv3x3_up {name}")
                return y
            t[up] = self._up(coarse, n, *dims[lvl + 1], s)
            return self._conv(name, t[up], skip, n, d_, a_, s)

        t["dec2a"] = decoder("dec2a", t["mid_b"], t["enc2b"], "u2", 2)
        t["dec2b"] = self._conv("dec2b", t["dec2a"], None, n, *dims[2], s)
        t["dec1a"] = decoder("dec1a", t["dec2b"], t["enc1b"], "u1", 1)
        t["dec1b"] = self._conv("dec1b", t["dec1a"], None, n, *dims[1], s)
        t["dec0a"] = decoder("dec0a", t["dec1b"], t["enc0b"], "u0", 0)
        t["dec0b"] = self._conv("dec0b", t["dec0a"], None, n, *dims[0], s)
        out = self._conv("out", t["dec0b"], None, n, *dims[0], s, out32=True)
        return (out, t) if keep else out

    # ---- backward ------------------------------------------------------------------------------
    def workspace_floats(self, n: int, down: int, across: int) -> int:
        most = 0
        for name, c1, c2, co, _ in self.layers:
            lvl = {"enc0": 0, "enc1": 1, "enc2": 2, "mid_": 3, "dec2": 2, "dec1": 1, "dec0": 0, "out": 0}[name[:4]]
            d, a = down >> lvl, across >> lvl
            most = max(most, lib().cg_conv3x3_wgrad_slices(n, d, a) * 9 * co * (c1 + c2))
        return most

    def backward(self, dout: torch.Tensor, t: Dict[str, torch.Tensor], work: torch.Tensor,
                 stream=None) -> Dict[str, torch.Tensor]:
        """Weight gradients (binary32, [9][cout][cin]) of sum(dout * out) for the activations t of
        forward(keep=True); dout is the (n, down, across, 16) binary32 gradient of the output."""
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        n, d, a, _ = dout.shape
        dims = [(d >> k, a >> k) for k in range(4)]
        grads: Dict[str, torch.Tensor] = {}
        keep: List[torch.Tensor] = []  # flipped weights stay allocated until the enqueued work is done
        L = lib()

        def wgrad(name, a1, a2, g, lvl):
            c1, c2, co, _ = self.spec[name]
            dw = torch.empty((9, co, c1 + c2), dtype=torch.float32, device=self.device)
            check(L.cg_conv3x3_wgrad(_p(a1), c1, _p(a2), c2, _p(g), co, n, *dims[lvl], _p(work), work.numel(),
                                     _p(dw), s), what=f"cg_conv3x3_wgrad {name}")
            grads[name] = dw

        def adjoint(name, g, lvl):
            c1, c2, co, _ = self.spec[name]
            g1 = torch.empty((n, *dims[lvl], c1), dtype=BF16, device=self.device)
            g2 = torch.empty((n, *dims[lvl], c2), dtype=BF16, device=self.device) if c2 else None
            if co % 32 == 0:
                # the forward kernels (K = 32, staged in LDS) with the mirrored, transposed weights
                wt = torch.empty((9, c1 + c2, co), dtype=BF16, device=self.device)
                check(L.cg_conv3x3_flip_weights(_p(self.wb[name]), co, c1 + c2, _p(wt), s),
                      what=f"cg_conv3x3_flip_weights {name}")
                check(L.cg_conv3x3(_p(g), co, None, 0, n, *dims[lvl], _p(wt), c1 + c2, 0, _p(g1), c1, _p(g2), None, s),
                      what=f"cg_conv3x3 (data gradient) {name}")
                keep.append(wt)
            else:
                check(L.cg_conv3x3_adjoint(_p(g), co, n, *dims[lvl], _p(self.wb[name]), c1 + c2, _p(g1), c1,
                                           _p(g2), None, s), what=f"cg_conv3x3_adjoint {name}")
            return g1, g2

        def relu_back(g, y):
            check(L.cg_relu_mask(_p(g), _p(y), g.numel(), s), what="cg_relu_mask")
            return g

        def add(x, y):
            check(L.cg_add_bf16(_p(x), _p(y), x.numel(), _p(x), s), what="cg_add_bf16")
            return x

        g = dout.to(BF16).contiguous()
        wgrad("out", t["dec0b"], None, g, 0)
        g, _ = adjoint("out", g, 0)
        # decoder, level 0
        g = relu_back(g, t["dec0b"])
        wgrad("dec0b", t["dec0a"], None, g, 0)
        g, _ = adjoint("dec0b", g, 0)
        g = relu_back(g, t["dec0a"])
        wgrad("dec0a", t["u0"], t["enc0b"], g, 0)
        g_u0, g_e0 = adjoint("dec0a", g, 0)
        g = self._pool(g_u0, n, *dims[0], s, scale=4.0)  # the adjoint of nearest upsampling: 2x2 sums
        # level 1
        g = relu_back(g, t["dec1b"])
        wgrad("dec1b", t["dec1a"], None, g, 1)
        g, _ = adjoint("dec1b", g, 1)
        g = relu_back(g, t["dec1a"])
        wgrad("dec1a", t["u1"], t["enc1b"], g, 1)
        g_u1, g_e1 = adjoint("dec1a", g, 1)
        g = self._pool(g_u1, n, *dims[1], s, scale=4.0)
        # level 2
        g = relu_back(g, t["dec2b"])
        wgrad("dec2b", t["dec2a"], None, g, 2)
        g, _ = adjoint("dec2b", g, 2)
        g = relu_back(g, t["dec2a"])
        wgrad("dec2a", t["u2"], t["enc2b"], g, 2)
        g_u2, g_e2 = adjoint("dec2a", g, 2)
        g = self._pool(g_u2, n, *dims[2], s, scale=4.0)
        # bottleneck
        g = relu_back(g, t["mid_b"])
        wgrad("mid_b", t["mid_a"], None, g, 3)
        g, _ = adjoint("mid_b", g, 3)
        g = relu_back(g, t["mid_a"])
        wgrad("mid_a", t["p3"], None, g, 3)
        g, _ = adjoint("mid_a", g, 3)
        # encoder: the pool's adjoint spreads a quarter to each cell of the block
        g = add(self._up(g, n, *dims[3], s, scale=0.25), g_e2)
        g = relu_back(g, t["enc2b"])
        wgrad("enc2b", t["enc2a"], None, g, 2)
        g, _ = adjoint("enc2b", g, 2)
        g = relu_back(g, t["enc2a"])
        wgrad("enc2a", t["p2"], None, g, 2)
        g, _ = adjoint("enc2a", g, 2)
        g = add(self._up(g, n, *dims[2], s, scale=0.25), g_e1)
        g = relu_back(g, t["enc1b"])
        wgrad("enc1b", t["enc1a"], None, g, 1)
        g, _ = adjoint("enc1b", g, 1)
        g = relu_back(g, t["enc1a"])
        wgrad("enc1a", t["p1"], None, g, 1)
        g, _ = adjoint("enc1a", g, 1)
        g = add(self._up(g, n, *dims[1], s, scale=0.25), g_e0)
        g = relu_back(g, t["enc0b"])
        wgrad("enc0b", t["enc0a"], None, g, 0)
        g, _ = adjoint("enc0b", g, 0)
        g = relu_back(g, t["enc0a"])
        wgrad("enc0a", t["x"], None, g, 0)
        return grads


# ---- training pairs and training -----------------------------------------------------------------

def fixed_channels(water: np.ndarray, depth: np.ndarray) -> np.ndarray:
    """(cells, 2) binary32: the mask and the scaled depth, the network's two constant inputs."""
    return np.stack([(np.asarray(water) != 0).astype(np.float32),
                     np.asarray(depth, np.float32) / DEPTH_SCALE], axis=1)


def centre_scale(stencil) -> float:
    """The mean diagonal weight over water: the factor that makes the network's target O(1)."""
    w = np.asarray(stencil.is_water) != 0
    c = (np.asarray(stencil.east_weight, np.float64) + stencil.west_weight + stencil.north_weight
         + stencil.south_weight)
    return float(c[w].mean()) if w.any() else 1.0


def training_pairs(acc, rows: int, seeds: Sequence[int], months: Sequence[int] = (-1,), tol: float = 1e-8):
    """(input (P, rows, 2 rows, 16) bfloat16, target (P, rows, 2 rows) binary32, water (P, rows, 2 rows)
    binary32) on the device, from this library's refined solves of the synthetic worlds: the input
    holds the mask, the depth and the forcing scaled to unit root mean square; the target is the
    solution scaled to unit root mean square (flexible GMRES weighs each direction it is given, so
    only the shape of the correction matters)."""
    from .ocean import Plan
    from .problem import circulation_problem, monthly_forcings

    cols = 2 * rows
    xs, ts, ws = [], [], []
    for seed in seeds:
        prob = circulation_problem(rows, seed=int(seed), land_fraction=0.3)
        forcings = monthly_forcings(prob, months) if tuple(months) != (-1,) else [prob.levels[0].forcing]
        plan = Plan(acc, prob.levels)
        fixed = fixed_channels(prob.water, prob.depth)
        try:
            for f in forcings:
                f = np.asarray(f, np.float32)
                plan.set_forcing(f)
                info = plan.solve(tol, CG_METHOD_REFINED)
                if info.relative_residual > tol:
                    raise RuntimeError(f"training solve of seed {seed} stopped at {info.relative_residual:.2e}")
                psi, _ = plan.read()
                scale = math.sqrt(f.size) / max(float(np.linalg.norm(f.astype(np.float64))), 1e-300)
                x = np.zeros((rows * cols, 16), np.float32)
                x[:, :2] = fixed
                x[:, 2] = f * scale
                xs.append(x.reshape(rows, cols, 16))
                rms = max(float(np.linalg.norm(psi)) / math.sqrt(psi.size), 1e-300)
                ts.append((psi / rms).astype(np.float32).reshape(rows, cols))
                ws.append((prob.water != 0).astype(np.float32).reshape(rows, cols))
        finally:
            plan.close()
    dev = "cuda"
    return (torch.from_numpy(np.stack(xs)).to(dev).to(BF16), torch.from_numpy(np.stack(ts)).to(dev),
            torch.from_numpy(np.stack(ws)).to(dev))


def loss_and_grad(out: torch.Tensor, target: torch.Tensor, water: torch.Tensor):
    """Mean squared error of channel 0 over water cells and its gradient with respect to out."""
    count = water.sum().clamp_min(1.0)
    diff = (out[..., 0] - target) * water
    loss = (diff * diff).sum() / count
    dout = torch.zeros_like(out)
    dout[..., 0] = 2.0 * diff / count
    return loss, dout


def train(net: UNet, data, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 0,
          decay: bool = True) -> List[float]:
    """Adam on the mean squared error, the step size decaying from lr to 0 along a cosine over the
    steps (decay) or held at lr; returns the loss of every step."""
    x_all, t_all, w_all = data
    P = x_all.shape[0]
    params = [net.w32[k].requires_grad_(True) for k, *_ in net.layers]
    opt = torch.optim.Adam(params, lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=max(1, steps)) if decay else None
    g = torch.Generator().manual_seed(seed)
    n = min(batch, P)
    work = torch.empty(net.workspace_floats(n, x_all.shape[1], x_all.shape[2]), dtype=torch.float32,
                       device=net.device)
    losses = []
    for _ in range(steps):
        idx = torch.randperm(P, generator=g)[:n].to(x_all.device)
        x, t, w = x_all[idx].contiguous(), t_all[idx].contiguous(), w_all[idx].contiguous()
        out, acts = net.forward(x, keep=True)
        loss, dout = loss_and_grad(out, t, w)
        grads = net.backward(dout, acts, work)
        for k, *_ in net.layers:
            net.w32[k].grad = grads[k]
        opt.step()
        if sched:
            sched.step()
        with torch.no_grad():
            net.refresh()
        losses.append(float(loss))
    for p in params:
        p.requires_grad_(False)
        p.grad = None
    return losses


class LearnedPreconditioner:
    """The trained network as a plan's preconditioner (cg_ocean_plan_set_preconditioner): z = the
    network's correction for v scaled to unit root mean square, scaled back, zero on land. Its
    buffers are allocated here once; each application only enqueues kernels on the plan's stream."""

    def __init__(self, net: UNet, stencil, depth):
        self.net = net
        self.down, self.across = stencil.cells_down, stencil.cells_across
        cells = self.down * self.across
        self.cells = cells
        dev = net.device
        self.fixed = torch.from_numpy(fixed_channels(stencil.is_water, depth)).to(dev).to(BF16).contiguous()
        self.water = torch.from_numpy((np.asarray(stencil.is_water) != 0).astype(np.uint8)).to(dev)
        self.x = torch.zeros((1, self.down, self.across, 16), dtype=BF16, device=dev)
        self.scale_in = math.sqrt(cells)  # FGMRES hands unit-norm vectors: this makes them unit RMS
        self.scale_out = 1.0 / (centre_scale(stencil) * self.scale_in)  # any scale serves FGMRES
        self.applications = 0
        self.error: Optional[BaseException] = None
        self._streams: Dict[int, torch.cuda.ExternalStream] = {}
        torch.cuda.synchronize()
        self._cb = PRECOND_FN(self._apply)

    def _apply(self, user, v, z, stream) -> int:
        try:
            L = lib()
            key = int(stream or 0)
            if key not in self._streams:
                self._streams[key] = torch.cuda.ExternalStream(key) if key else torch.cuda.default_stream()
            # The network's activations are allocated with the plan's stream current, so torch's
            # caching allocator hands their memory on only in that stream's order.
            with torch.cuda.stream(self._streams[key]):
                s = torch.cuda.current_stream().cuda_stream
                check(L.cg_unet_pack_input(_p(self.fixed), ctypes.c_void_p(v), self.scale_in, self.cells, _p(self.x),
                                           s), what="cg_unet_pack_input")
                out = self.net.forward(self.x)
                check(L.cg_unet_unpack_output(_p(out), _p(self.water), self.scale_out, self.cells, ctypes.c_void_p(z),
                                              s), what="cg_unet_unpack_output")
            self.applications += 1
            return 0
        except BaseException as e:  # an exception must not cross the C boundary
            self.error = e
            return -4

    def attach(self, plan) -> None:
        check(lib().cg_ocean_plan_set_preconditioner(plan.handle, self._cb, None), plan.acc.ctx,
              "cg_ocean_plan_set_preconditioner")

    @staticmethod
    def detach(plan) -> None:
        check(lib().cg_ocean_plan_set_preconditioner(plan.handle, PRECOND_FN(), None), plan.acc.ctx,
              "cg_ocean_plan_set_preconditioner")


# ---- the learned coarse-grid correction (CG_METHOD_LEARNED) ---------------------------------------

COARSE_MOST_CELLS = 64 * 128  # the network serves the finest grid of at most this many cells


def coarse_level(levels, most_cells: int = COARSE_MOST_CELLS) -> int:
    """The grid of a ladder the coarse network serves: the finest of at most most_cells cells whose
    shape the network takes (rows a multiple of 8, columns of 128), not the finest grid; -1 if none."""
    for k, st in enumerate(levels):
        if k == 0 or st.cells > most_cells:
            continue
        if st.cells_down % 8 == 0 and st.cells_across % 128 == 0:
            return k
        return -1
    return -1


def coarse_depth(prob, level: int) -> np.ndarray:
    """The depth on a coarse grid of the problem's ladder: block means of the finest grid's (0 on land)."""
    from .ocean import coarse_mean

    st = prob.levels[level]
    return coarse_mean(np.asarray(prob.depth, np.float32), prob.cols, prob.rows, st.cells_across, st.cells_down)


def _bf16_bits(t: torch.Tensor) -> np.ndarray:
    return t.detach().to(BF16).contiguous().view(torch.int16).cpu().numpy().view(np.uint16)


def coarse_pairs(acc, seeds: Sequence[int], rows: int = 1024, land_fraction: float = 0.3, tol: float = 1e-8,
                 most_steps: int = 40, threads: int = 16):
    """Training pairs of the coarse network from plain multigrid solves of the synthetic worlds of
    rows x 2 rows (seeds given, annual forcing): every step's forcing of the grid the network serves
    (cg_ocean_plan_capture_coarse) with the exact solution of that grid's problem (the refined method
    on the ladder below it, to 1e-11), the forcing scaled to unit rms and the solution by the same rms
    and the output scale (the median ratio of the solutions' rms to the forcings'). Returns a dict:
    inputs (P, down, across, 16) bfloat16, targets and water (P, down, across) binary32 on the device,
    out_scale, level, the grid's shape, and the counts of worlds and solves behind them."""
    from concurrent.futures import ThreadPoolExecutor

    from .ocean import Plan
    from .problem import circulation_problem

    seeds = [int(v) for v in seeds]
    # Per pair only the scaled forcing, the target and the pair's world; per world its mask and depth
    # (the network's fixed input channels). The 16-channel inputs are assembled on the device.
    fins, fs, world_of, ratios = [], [], [], []
    fixed_w, water_w = [], []
    level, shape = -1, None
    solves = 0

    def make(seed):
        return circulation_problem(rows, seed=seed, land_fraction=land_fraction)

    def generated():
        # At most 2 x threads worlds (a few tens of MB each at 1024 x 2048) are generated ahead of the
        # device, however many seeds there are; they come back in seed order.
        ahead = 2 * max(1, threads)
        with ThreadPoolExecutor(max_workers=max(1, threads)) as pool:
            pending = [pool.submit(make, sd) for sd in seeds[:ahead]]
            nxt = len(pending)
            done = 0
            while pending:
                fut = pending.pop(0)
                if nxt < len(seeds):
                    pending.append(pool.submit(make, seeds[nxt]))
                    nxt += 1
                yield fut.result()
                done += 1
                if len(seeds) > 256 and done % 250 == 0:
                    print(f"coarse pairs: {done} of {len(seeds)} worlds", file=sys.stderr, flush=True)

    for prob in generated():
        L = coarse_level(prob.levels)
        if L < 0:
            raise ValueError(f"no grid of the {rows}-row ladder suits the coarse network")
        st = prob.levels[L]
        if shape is None:
            level, shape = L, (st.cells_down, st.cells_across)
        elif (st.cells_down, st.cells_across) != shape:
            raise ValueError("the worlds' coarse grids differ in shape")
        cells = st.cells
        plan = Plan(acc, prob.levels)
        sub = Plan(acc, prob.levels[L:])
        try:
            out = np.empty((most_steps, cells), np.float32)
            n = check(lib().cg_ocean_plan_capture_coarse(plan.handle, L, float(tol), most_steps,
                                                         out.ctypes.data_as(ctypes.c_void_p)),
                      acc.ctx, "cg_ocean_plan_capture_coarse")
            solves += 1
            depth = coarse_depth(prob, L)
            water = (np.asarray(st.is_water) != 0)
            w = len(fixed_w)
            fixed_w.append(np.stack([water.astype(np.float32), depth / DEPTH_SCALE], axis=1))
            water_w.append(water.astype(np.float32))
            for f in out[:n]:
                rms_f = float(np.sqrt(np.mean(f.astype(np.float64) ** 2)))
                if rms_f == 0.0:
                    continue
                sub.set_forcing(f)
                info = sub.solve(1e-11, CG_METHOD_REFINED)
                if not info.relative_residual < 1e-11:
                    raise RuntimeError(f"coarse target solve stopped at {info.relative_residual:.2e}")
                x64, _ = sub.read()
                fins.append(f / np.float32(rms_f))
                fs.append((x64 / rms_f).astype(np.float32))
                world_of.append(w)
                ratios.append(float(np.sqrt(np.mean(x64 ** 2))) / rms_f)
        finally:
            sub.close()
            plan.close()
    if not fins:
        raise RuntimeError("no training pairs")
    out_scale = float(np.median(ratios))
    dev = "cuda"
    P = len(fins)
    idx = torch.as_tensor(np.asarray(world_of, np.int64), device=dev)
    inputs = torch.zeros((P, shape[0] * shape[1], 16), dtype=BF16, device=dev)
    inputs[:, :, :2] = torch.from_numpy(np.stack(fixed_w)).to(dev)[idx].to(BF16)
    inputs[:, :, 2] = torch.from_numpy(np.stack(fins)).to(dev).to(BF16)
    del fins
    targets = torch.from_numpy(np.stack(fs)).to(dev).reshape(P, *shape) / np.float32(out_scale)
    water = torch.from_numpy(np.stack(water_w)).to(dev)[idx].reshape(P, *shape)
    return {"inputs": inputs.reshape(P, *shape, 16), "targets": targets, "water": water, "out_scale": out_scale,
            "level": level, "shape": shape, "worlds": len(seeds), "solves": solves, "pairs": P}


class CoarseNetwork:
    """A trained U-Net as a plan's learned coarse-grid correction (cg_ocean_plan_set_coarse_net):
    the library holds a copy of its bfloat16 weights and runs it inside the plan's step graphs."""

    def __init__(self, net: UNet, out_scale: float):
        self.net = net
        self.out_scale = float(out_scale)

    def weights(self) -> np.ndarray:
        return np.concatenate([_bf16_bits(self.net.wb[name]).reshape(-1) for name, *_ in self.net.layers])

    def attach(self, plan, prob, level: Optional[int] = None) -> int:
        """Sets the network on the plan's grid `level` (default coarse_level of the problem's ladder),
        with that grid's depth; returns the level."""
        L = coarse_level(prob.levels) if level is None else int(level)
        w = self.weights()
        depth = np.ascontiguousarray(coarse_depth(prob, L), np.float32)
        check(lib().cg_ocean_plan_set_coarse_net(plan.handle, L, self.net.base, w.ctypes.data_as(ctypes.c_void_p),
                                                 w.size, depth.ctypes.data_as(ctypes.c_void_p), self.out_scale),
              plan.acc.ctx, "cg_ocean_plan_set_coarse_net")
        return L

    @staticmethod
    def detach(plan) -> None:
        check(lib().cg_ocean_plan_set_coarse_net(plan.handle, 0, 32, None, 0, None, 1.0), plan.acc.ctx,
              "cg_ocean_plan_set_coarse_net")


def train_coarse(data, base: int = 32, steps: int = 1000, batch: int = 16, lr: float = 1e-3, seed: int = 0):
    """A U-Net trained on coarse_pairs' data (Adam on the mean squared error over water, the library's
    forward and backward kernels); returns (CoarseNetwork, losses)."""
    net = UNet(base, seed=seed)
    losses = train(net, (data["inputs"], data["targets"], data["water"]), steps=steps, batch=batch, lr=lr, seed=seed)
    return CoarseNetwork(net, data["out_scale"]), losses
