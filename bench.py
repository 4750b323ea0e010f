"""Benchmark: circulation solves per second at 1024 x 2048 to relative residual 1e-8 on one MI355X.

One step = one solve of the synthetic 1024 x 2048 world (seed 0, 30% land, annual forcing) from
psi = 0 to max|F - L psi| / max|F| < 1e-8, with the ladder, forcing and answer resident in HBM
(a device-resident plan). Prints one JSON line.

`value` is plain GPU multigrid (binary32 V(2,2) cycles of the reference's ladder inside binary64
iterative refinement). The learned leg runs in a child process and is reported under "learned" and
at the top level (learned_solves_per_s, learned_vs_plain_mg): the U-Net trained on the GPU as the
cycle's coarse-grid correction, the refined solve with it (one instance and the batched year) over
the plain multigrid's, the network's rate, and its largest convolution at the full grid against the
roofline that bounds it.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

METRIC = "solves/s at 1024×2048 to 1e-8 residual: learned-precond vs plain GPU MG"
HBM_PEAK_GBS = 8000.0
ROWS, COLS = 1024, 2048
# Algorithmic bytes per launch at 1024 x 2048 (N = 2,097,152 cells), each array read or written once:
#  - the binary64 outer step k_outer_residual: x64 read 8 + delta 4 + f 4 + mask 1, x64 written 8 +
#    correction forcing 4 = 29 B/cell (the per-cell FP64 residual of the byte model, 25 B, plus the
#    x += delta fused into it);
#  - the finest up leg k_relax_tile<4,8,false,2>: x 4 + f 4 + mask 1 + coarse x 1 read, x 4 written
#    = 14 B/cell (the 13 B two-pass smoothing figure plus the fused prolongation's coarse values).
#  - the finest up leg with the outer step fused in, k_relax_tile<4,8,false,3>: x 4 + correction
#    forcing 4 + mask 1 + coarse x 1 + x64 8 + f0 4 read, x64 8 + next forcing 4 written = 34 B/cell
#    (the 14 + 29 of the two launches less the correction written and read back and the mask read
#    twice).
CELLS = 1024 * 2048
OUTER_BYTES = 29 * CELLS
UP_BYTES = 14 * CELLS
FUSED_BYTES = 34 * CELLS
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE and WRITE_SIZE (scaled by a 1 GiB copy),
# for the default settings (two passes per grid), by the finest grid's tile rows; see
# profiles/hbm_traffic_*.md.
COPY_BYTES = 1 << 30
MEASURED_TRAFFIC = {"outer": 82793991, "up": {4: 30604038, 6: 30719584, 8: 30631686},
                    "fused": {4: 73863905, 6: 73804771, 8: 73819458}}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """(nproc, lscpu model name, threads used): the threads are the OpenMP share the host allots to
    one GPU (OMP_NUM_THREADS) when it sets one, else every core."""
    nproc = os.cpu_count() or 1
    model = None
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return nproc, model, (min(share, nproc) if share > 0 else nproc)


def _timed_solves(fn, budget_s: float, most: int):
    times, last = [], None
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        last = fn()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all > budget_s or len(times) >= most:
            break
    return len(times) / sum(times), len(times), last


def cpu_baseline(levels, levels256, budget_s: float, threads: int, nproc: int, model):
    """The oracle (the reference algorithm restated in C: OceanCirculation's levels, V(2,2), row-band
    threads like parallelChunks) on the host. The headline figure: its binary64 twin solving the
    benchmark's 1024 x 2048 problem to 1e-8 with `threads` OpenMP threads, repeated until budget_s.
    Beside it, bounded samples of: the same on one thread; the reference method itself (binary32
    values, binary64 residual) to the reference's own tolerance 1e-3 at 1024 x 2048, on all threads
    and one; the binary64 twin to 1e-10 at 256 x 512 (seed 0), on all threads and one."""
    import oracle

    def ladder_of(lv):
        return [oracle.Stencil(s.cells_across, s.cells_down, s.is_water, s.east_weight, s.west_weight, s.north_weight,
                               s.south_weight, s.forcing, s.centre_weight) for s in lv]

    big, small = ladder_of(levels), ladder_of(levels256)
    zero = np.zeros(big[0].cells, np.float32)
    legs = {}
    oracle.set_threads(threads)
    value, count, (_, cycles, _) = _timed_solves(lambda: oracle.solve64(big, 1e-8), budget_s, 20)
    for nt in (threads, 1):
        oracle.set_threads(nt)
        tag = "all" if nt == threads else "one_thread"
        if nt == 1:
            v1, c1, _ = _timed_solves(lambda: oracle.solve64(big, 1e-8), budget_s / 2, 3)
            legs["binary64_1e-8_1024x2048_one_thread"] = {"solves_per_s": round(v1, 4), "solves": c1}
        vr, cr, (_, cyc_r, rr) = _timed_solves(lambda: oracle.solve(big, zero, 1e-3), budget_s / 4, 5)
        legs[f"reference_method_1e-3_1024x2048_{tag}"] = {"solves_per_s": round(vr, 4), "solves": cr,
                                                          "cycles": cyc_r, "relative_residual": rr}
        vs, cs_, (_, cyc_s, _) = _timed_solves(lambda: oracle.solve64(small, 1e-10), budget_s / 4, 20)
        legs[f"binary64_1e-10_256x512_{tag}"] = {"solves_per_s": round(vs, 4), "solves": cs_, "cycles": cyc_s}
    oracle.set_threads(threads)
    return {"value": round(value, 4), "unit": "solves/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"{count} full 1024x2048 solves to 1e-8 ({cycles} V(2,2) cycles each) by the oracle's "
                      f"binary64 C restatement of OceanCirculation.solve, {threads} OpenMP threads (the host's "
                      f"per-GPU share of its {nproc} cores)",
            "legs": legs}


def batched_year(acc, prob, tol: float, reps: int, warmup: int, autotune: bool = True):
    """The 12 monthly forcings of the benchmark world solved together in one plan (one launch per
    kernel for all twelve), from zero to tol each: solves/s = 12 x reps / wall time."""
    import cartogenesis_amd as cg
    from cartogenesis_amd.ocean import BatchPlan
    from cartogenesis_amd.problem import monthly_forcings

    log("bench: batched year (12 monthly forcings of the 1024x2048 world)")
    forcings = monthly_forcings(prob)
    plan = BatchPlan(acc, prob.levels, len(forcings))
    try:
        for i, f in enumerate(forcings):
            plan.set_forcing(i, f)
        tuned = plan.autotune(5) if autotune else None
        for _ in range(max(1, warmup // 2)):
            infos = plan.solve_all(tol)
        t0 = time.perf_counter()
        for _ in range(reps):
            infos = plan.solve_all(tol)
        elapsed = time.perf_counter() - t0
        # The roofline kernel as the batched steps launch it: the finest up leg with the outer step
        # fused in, all twelve instances in one launch, 200 launches between HIP events on the
        # library's stream.
        us_batched = cg.lib().cg_ocean_plan_time_kernel(plan.handle, 6, 200)
    finally:
        plan.close()
    value = len(forcings) * reps / elapsed
    log(f"bench: batched year {value:.1f} solves/s, {1000 * elapsed / reps:.2f} ms per year, fused kernel "
        f"{us_batched:.2f} us for {len(forcings)} instances")
    rows0 = tuned[0] if tuned and tuned[0] else 8
    roof = None
    if us_batched > 0:
        gbs = len(forcings) * FUSED_BYTES / (us_batched * 1e-6) / 1e9
        roof = {"bound": "hbm",
                "kernel": f"k_relax_tile<4,{rows0},false,3> over {len(forcings)} instances (grid y), as the batched "
                          "steps launch it",
                "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "avg_launch_us": round(us_batched, 3),
                "algorithmic_bytes_per_launch": len(forcings) * FUSED_BYTES}
    return {
        "workload": "12 monthly forcings (band amplitude 0.1 cos(2 pi m / 12), shifted 10 deg sin(2 pi m / 12)) of "
                    "the 1024x2048 world, solved together in one plan from zero to tol each",
        "instances": len(forcings), "reps": reps,
        "solves_per_s": round(value, 3),
        "ms_per_batch": round(1000.0 * elapsed / reps, 4),
        "cycles": [int(i.cycles) for i in infos],
        "max_relative_residual": max(float(i.relative_residual) for i in infos),
        "autotuned_tile_rows": tuned,
        "roofline": roof,
    }


def uneven_batches(acc, specs, instances: int, tol: float, autotune: bool = True):
    """Batched plans of several worlds of uneven size, each plan holding `instances` monthly
    forcings of its world, timed two ways: the plans one after another on one context, and each
    plan on a context (a HIP stream) of its own, solved by its own host thread, so that the worlds'
    latency-bound launches overlap on the GPU. Returns (sequential, concurrent) solves/s, per-plan
    milliseconds of the sequential run, and the largest residual."""
    import threading
    from cartogenesis_amd.ocean import BatchPlan, HipOcean, PlanGroup
    from cartogenesis_amd.problem import circulation_problem, monthly_forcings

    worlds = []
    for rows, seed, land in specs:
        prob = circulation_problem(rows, seed=seed, land_fraction=land)
        forcings = monthly_forcings(prob)
        worlds.append((rows, prob, [forcings[i % len(forcings)] for i in range(instances)]))

    def make(ctx_acc, prob, forcings):
        plan = BatchPlan(ctx_acc, prob.levels, instances)
        for i, f in enumerate(forcings):
            plan.set_forcing(i, f)
        if autotune:
            plan.autotune(3)
        plan.solve_all(tol)  # warm
        return plan

    total = instances * len(worlds)
    plans = [make(acc, prob, forcings) for _, prob, forcings in worlds]
    group = None
    try:
        t0 = time.perf_counter()
        per, cycles, worst = {}, {}, 0.0
        for k, ((rows, _, _), plan) in enumerate(zip(worlds, plans)):
            t1 = time.perf_counter()
            infos = plan.solve_all(tol)
            per[f"{k}:{rows}x{2 * rows}"] = round(1000.0 * (time.perf_counter() - t1), 3)
            worst = worst_of(worst, infos)
            cycles[f"{k}:{rows}x{2 * rows}"] = [min(i.cycles for i in infos), max(i.cycles for i in infos)]
        sequential = total / (time.perf_counter() - t0)
        # The same plans as one group: one host loop; the merged step (one graph a step, each tile
        # launch covering one depth of every plan's ladder), then each plan's own step graph on a
        # stream of the group's.
        import cartogenesis_amd as cg

        group = PlanGroup(acc, plans)
        reps = 3
        timed = {}
        for merged in (1, 0):
            acc.set_option(cg.CG_OPT_GROUP_MERGED, merged)
            group.solve(tol)  # warm: the graphs
            t0 = time.perf_counter()
            for _ in range(reps):
                for infos in group.solve(tol):
                    worst = worst_of(worst, infos)
            timed[merged] = reps * total / (time.perf_counter() - t0)
        acc.set_option(cg.CG_OPT_GROUP_MERGED, 1)
        grouped = (timed[1], timed[0])
    finally:
        if group:
            group.close()
        for plan in plans:
            plan.close()
    accs = [HipOcean.create_or_null(0).accelerator for _ in worlds]
    plans = [make(a, prob, forcings) for a, (_, prob, forcings) in zip(accs, worlds)]
    try:
        results = [None] * len(plans)

        def run(k):
            results[k] = plans[k].solve_all(tol)

        threads = [threading.Thread(target=run, args=(k,)) for k in range(len(plans))]
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        concurrent = total / (time.perf_counter() - t0)
        for infos in results:
            worst = worst_of(worst, infos)
    finally:
        for plan in plans:
            plan.close()
        for a in accs:
            a.close()
    return sequential, concurrent, grouped, per, cycles, worst


def worst_of(worst: float, infos) -> float:
    """The largest relative residual so far; NaN if any instance's residual is NaN."""
    for i in infos:
        r = float(i.relative_residual)
        if r != r or worst != worst:
            return float("nan")
        worst = max(worst, r)
    return worst


def batched_worlds(acc, tol: float, autotune: bool = True):
    """Four worlds of uneven size, twelve months each (one batched plan of twelve per world)."""
    log("bench: four-world year")
    specs = [(512, 1, 0.20), (768, 2, 0.27), (1024, 3, 0.33), (1024, 4, 0.40)]
    seq, conc, grouped, per, cycles, worst = uneven_batches(acc, specs, 12, tol, autotune)
    log(f"bench: four-world year {seq:.1f} solves/s one world after another, {conc:.1f} with the worlds concurrent, "
        f"{grouped[0]:.1f} as one group (merged step), {grouped[1]:.1f} as one group (a stream per world)")
    return {"workload": "12 months x worlds 512x1024, 768x1536, 1024x2048, 1024x2048 (seeds 1-4, land 20/27/33/40%)",
            "solves_per_s": round(grouped[0], 3), "method": "one group (cg_ocean_group_solve), merged step: one host "
            "loop, one hipGraph launch a refinement step for the 48 instances, each tile launch covering one depth of "
            "every world's ladder and all its months",
            "solves_per_s_group_streams": round(grouped[1], 3),
            "solves_per_s_concurrent_streams": round(conc, 3), "solves_per_s_one_world_at_a_time": round(seq, 3),
            "ms_per_world_one_at_a_time": per, "cycles_min_max": cycles, "max_relative_residual": worst}


def batched_combined(acc, tol: float, autotune: bool = True):
    """64 instances of uneven size: 16 each at 256x512, 512x1024, 1024x2048 and 2048x4096 (months
    0..11 then 0..3), one batched plan of 16 per size."""
    log("bench: combined 64 instances")
    specs = [(256, 5, 0.3), (512, 6, 0.3), (1024, 7, 0.3), (2048, 8, 0.3)]
    seq, conc, grouped, per, cycles, worst = uneven_batches(acc, specs, 16, tol, autotune)
    log(f"bench: combined 64 instances {seq:.1f} solves/s one size after another, {conc:.1f} concurrent, "
        f"{grouped[0]:.1f} as one group (merged step), {grouped[1]:.1f} as one group (a stream per size)")
    return {"workload": "64 instances: 16 monthly forcings each at 256x512, 512x1024, 1024x2048, 2048x4096",
            "solves_per_s": round(grouped[0], 3), "method": "one group (cg_ocean_group_solve), merged step: one host "
            "loop, one hipGraph launch a refinement step for the 64 instances, each tile launch covering one depth of "
            "every size's ladder and all its instances",
            "solves_per_s_group_streams": round(grouped[1], 3),
            "solves_per_s_concurrent_streams": round(conc, 3), "solves_per_s_one_size_at_a_time": round(seq, 3),
            "ms_per_size_one_at_a_time": per, "cycles_min_max": cycles, "max_relative_residual": worst}


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bfloat16 matrix-core peak


def unet_flops(rows: int, cols: int, base: int = 32) -> dict:
    """Multiply-add flops (2 per product) of each U-Net convolution on a rows x cols grid; the first
    layer's input counts its 3 data channels (mask, depth, vector), which are all it multiplies."""
    from cartogenesis_amd.learned import layer_table

    level = {"enc0": 0, "enc1": 1, "enc2": 2, "mid_": 3, "dec2": 2, "dec1": 1, "dec0": 0, "out": 0}
    return {name: 2 * 9 * (3 if name == "enc0a" else c1 + c2) * co * (rows >> level[name[:4]]) * (cols >> level[name[:4]])
            for name, c1, c2, co, _ in layer_table(base)}


def coarse_net_flops(down: int, across: int, base: int = 32) -> int:
    """Multiply-add flops (2 per product) of one coarse-network application on a down x across grid;
    the first layer counts its 3 data channels."""
    from cartogenesis_amd.learned import layer_table

    level = {"enc0": 0, "enc1": 1, "enc2": 2, "mid_": 3, "dec2": 2, "dec1": 1, "dec0": 0, "out": 0}
    return sum(2 * 9 * (3 if name == "enc0a" else c1 + c2) * co * (down >> level[name[:4]]) * (across >> level[name[:4]])
               for name, c1, c2, co, _ in layer_table(base))


def learned_leg(acc, prob, tol: float, worlds: int, steps: int, reps: int, batch_reps: int, rows=None,
                autotune: bool = True):
    """The learned coarse-grid correction (CG_METHOD_LEARNED): a U-Net trained on the GPU (the library's
    kernels) on the coarse grid's forcings of plain solves of `worlds` synthetic 1024x2048 worlds (seeds
    100, 101, ..., config[1]'s distributions) and their exact solutions, then set on the benchmark's
    plan: solves/s with and without it, one instance and the batched year, the steps, the error against
    the binary64 oracle, and the network's matrix-core rate."""
    import torch
    import cartogenesis_amd as cg
    import oracle
    from cartogenesis_amd.learned import coarse_level, coarse_pairs, train_coarse
    from cartogenesis_amd.ocean import BatchPlan, Plan
    from cartogenesis_amd.problem import monthly_forcings

    log(f"bench: learned: coarse pairs from {worlds} worlds of 1024x2048, {steps} training steps")
    t0 = time.perf_counter()
    data = coarse_pairs(acc, range(100, 100 + worlds), rows=ROWS, threads=int(os.environ.get("OMP_NUM_THREADS", "8")))
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    cn, losses = train_coarse(data, base=32, steps=steps, batch=16)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    level, shape = data["level"], data["shape"]
    out = {"design": "coarse-grid correction: the U-Net on the cycle's %dx%d grid (level %d) in place of the grids "
                     "below it, inside the graph-captured refined step" % (shape[1], shape[0], level),
           "training": {"worlds": worlds, "solves": data["solves"], "pairs": data["pairs"], "grid": f"{shape[1]}x{shape[0]}",
                        "seeds": f"100-{99 + worlds}", "steps": steps, "batch": 16, "data_s": round(t_data, 2),
                        "train_s": round(t_train, 2), "loss_first": round(losses[0], 4),
                        "loss_last20": round(float(np.mean(losses[-20:])), 4), "out_scale": round(data["out_scale"], 6)}}
    del data
    plan = Plan(acc, prob.levels)
    try:
        if rows:
            plan.set_tile_rows(rows + [0] * (len(prob.levels) - len(rows)))
        elif autotune:
            plan.autotune(20)

        def timed(method, n):
            info = plan.solve(tol, method)
            t = time.perf_counter()
            for _ in range(n):
                info = plan.solve(tol, method)
            return n / (time.perf_counter() - t), info

        plain_value, plain = timed(cg.CG_METHOD_REFINED, reps)
        assert coarse_level(prob.levels) == level
        cn.attach(plan, prob)
        ln_value, ln = timed(cg.CG_METHOD_LEARNED, reps)
        fell = cg.lib().cg_ocean_plan_net_fallbacks(plan.handle)
        x64, _ = plan.read()
        ladder = [oracle.Stencil(s.cells_across, s.cells_down, s.is_water, s.east_weight, s.west_weight,
                                 s.north_weight, s.south_weight, s.forcing, s.centre_weight) for s in prob.levels]
        oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "8")))
        ref, _, _ = oracle.solve64(ladder, 1e-12)
        err = float(np.linalg.norm(x64 - ref) / np.linalg.norm(ref))
        out |= {"solves_per_s": round(ln_value, 3), "steps": ln.cycles, "relative_residual": ln.relative_residual,
                "rel_l2_vs_fp64_oracle": err, "fallbacks": fell,
                "plain_mg_solves_per_s": round(plain_value, 3), "plain_mg_steps": plain.cycles,
                "learned_vs_plain_mg": round(ln_value / plain_value, 4) if plain_value > 0 else None,
                "ms_per_step_learned": round(1000.0 / ln_value / max(ln.cycles, 1), 4),
                "ms_per_step_plain": round(1000.0 / plain_value / max(plain.cycles, 1), 4)}
        log(f"bench: learned {ln_value:.1f} solves/s ({ln.cycles} steps, {fell} fallbacks, rel L2 {err:.2e}) against "
            f"plain MG {plain_value:.1f} ({plain.cycles} steps) in the same process")
        # The network alone on its grid: one application's launches, timed with HIP events on the
        # stream they are enqueued on (apply_coarse_net synchronises its own copies).
        L = prob.levels[level]
        f = np.where(L.is_water != 0, 1.0, 0.0).astype(np.float32)
        z = np.empty(L.cells, np.float32)
        cg.lib().cg_ocean_plan_apply_coarse_net(plan.handle, f, z)
        t = time.perf_counter()
        for _ in range(20):
            cg.lib().cg_ocean_plan_apply_coarse_net(plan.handle, f, z)
        apply_us = (time.perf_counter() - t) / 20 * 1e6
        fl = coarse_net_flops(L.cells_down, L.cells_across)
        out["mfma_roofline_network"] = {
            "bound": "mfma (latency-bound at this size)", "flops_per_application": fl,
            "application_us_with_copies": round(apply_us, 2),
            "achieved": round(fl / (apply_us * 1e-6) / 1e12, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(fl / (apply_us * 1e-6) / 1e12 / BF16_PEAK_TFLOPS, 6)}
    finally:
        plan.close()
    if batch_reps > 0:
        forcings = monthly_forcings(prob)
        bp = BatchPlan(acc, prob.levels, len(forcings))
        try:
            for i, f in enumerate(forcings):
                bp.set_forcing(i, f)
            bp.autotune(5)
            cn.attach(bp, prob)
            res = {}
            for name, m in (("plain", cg.CG_METHOD_REFINED), ("learned", cg.CG_METHOD_LEARNED)):
                infos = (cg.CgSolution * len(forcings))()
                cg.lib().cg_ocean_plan_solve_all(bp.handle, tol, m, infos)
                t = time.perf_counter()
                for _ in range(batch_reps):
                    cg.check(cg.lib().cg_ocean_plan_solve_all(bp.handle, tol, m, infos), acc.ctx, "solve_all")
                res[name] = (len(forcings) * batch_reps / (time.perf_counter() - t), [int(i.cycles) for i in infos])
            out["batched_year"] = {"plain_solves_per_s": round(res["plain"][0], 3), "plain_steps": res["plain"][1],
                                   "learned_solves_per_s": round(res["learned"][0], 3),
                                   "learned_steps": res["learned"][1],
                                   "fallbacks": cg.lib().cg_ocean_plan_net_fallbacks(bp.handle),
                                   "learned_vs_plain_mg": round(res["learned"][0] / res["plain"][0], 4)}
            log(f"bench: learned batched year {res['learned'][0]:.1f} solves/s against plain {res['plain'][0]:.1f}")
        finally:
            bp.close()
        out["combined"] = learned_combined(acc, cn, tol, batch_reps)
    # The largest U-Net convolution at the full 1024x2048 grid (dec0a, 96 -> 32 channels), the matrix-core
    # kernel's rate where its size is not latency-bound.
    out["mfma_roofline"] = conv_roofline(cn.net)
    return out


def learned_combined(acc, cn, tol: float, reps: int) -> dict:
    """The 64 uneven instances (16 monthly forcings each at 256x512, 512x1024, 1024x2048 and 2048x4096)
    as one group, every plan with the network on its 64x128 grid (the same globe grid in every
    ladder), by the learned method and plainly; each size's first instance against the binary64
    oracle."""
    import cartogenesis_amd as cg
    import oracle
    from cartogenesis_amd.ocean import BatchPlan, PlanGroup
    from cartogenesis_amd.problem import circulation_problem, monthly_forcings

    log("bench: learned combined 64 instances")
    specs = [(256, 5, 0.3), (512, 6, 0.3), (1024, 7, 0.3), (2048, 8, 0.3)]
    probs, plans, group = [], [], None
    try:
        for rows, seed, land in specs:
            prob = circulation_problem(rows, seed=seed, land_fraction=land)
            forcings = monthly_forcings(prob)
            plan = BatchPlan(acc, prob.levels, 16)
            for i in range(16):
                plan.set_forcing(i, forcings[i % 12])
            plan.autotune(3)
            cn.attach(plan, prob)
            probs.append((prob, forcings[0]))
            plans.append(plan)
        group = PlanGroup(acc, plans)
        res = {}
        for name, m in (("plain", cg.CG_METHOD_REFINED), ("learned", cg.CG_METHOD_LEARNED)):
            group.solve(tol, m)
            t = time.perf_counter()
            for _ in range(max(reps, 1)):
                infos = group.solve(tol, m)
            res[name] = (64 * max(reps, 1) / (time.perf_counter() - t), infos)
        oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "8")))
        errs, steps = {}, {}
        for (rows, _, _), (prob, f0), plan, infos in zip(specs, probs, plans, res["learned"][1]):
            ladder = [oracle.Stencil(s.cells_across, s.cells_down, s.is_water, s.east_weight, s.west_weight,
                                     s.north_weight, s.south_weight, s.forcing, s.centre_weight) for s in prob.levels]
            ladder[0] = ladder[0].with_forcing(f0)
            ref = oracle.solve64(ladder, 1e-12)[0]
            errs[f"{rows}x{2 * rows}"] = float(np.linalg.norm(plan.read(0)[0] - ref) / np.linalg.norm(ref))
            steps[f"{rows}x{2 * rows}"] = [min(i.cycles for i in infos), max(i.cycles for i in infos)]
        fallbacks = sum(cg.lib().cg_ocean_plan_net_fallbacks(p.handle) for p in plans)
        worst = max(max(i.relative_residual for i in infos) for infos in res["learned"][1])
    finally:
        if group:
            group.close()
        for plan in plans:
            plan.close()
    log(f"bench: learned combined {res['learned'][0]:.1f} solves/s against plain {res['plain'][0]:.1f}")
    return {"workload": "64 instances: 16 monthly forcings each at 256x512, 512x1024, 1024x2048, 2048x4096, one "
                        "group (cg_ocean_group_solve), the network on every ladder's 64x128 grid",
            "learned_solves_per_s": round(res["learned"][0], 3), "plain_solves_per_s": round(res["plain"][0], 3),
            "learned_vs_plain_mg": round(res["learned"][0] / res["plain"][0], 4), "learned_steps_min_max": steps,
            "fallbacks": fallbacks, "max_relative_residual": worst,
            "rel_l2_vs_fp64_oracle_month0": errs}


def conv_roofline(net) -> dict:
    import torch

    flops = unet_flops(ROWS, COLS)["dec0a"]
    a1 = torch.randn((1, ROWS, COLS, 64), device="cuda").to(torch.bfloat16)
    a2 = torch.randn((1, ROWS, COLS, 32), device="cuda").to(torch.bfloat16)
    st = torch.cuda.current_stream().cuda_stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    net._conv("dec0a", a1, a2, 1, ROWS, COLS, st)
    e0.record()
    for _ in range(20):
        net._conv("dec0a", a1, a2, 1, ROWS, COLS, st)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 20 * 1000.0
    tf = flops / (us * 1e-6) / 1e12
    algo = 2 * ROWS * COLS * (64 + 32 + 32) + 2 * 9 * 32 * 96
    # the memory roof at this intensity: flops per algorithmic byte x 8 TB/s
    mem_roof = flops / algo * HBM_PEAK_GBS * 1e9 / 1e12
    return {"bound": "hbm at this intensity (%.0f flop/B, under the %.0f flop/B ridge of 2.5 PFLOP/s over 8 TB/s)"
                     % (flops / algo, BF16_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)),
            "kernel": "k_conv3x3_tile<2> (dec0a: 96 -> 32 channels, 1024x2048, bf16 in, fp32 accumulate)",
            "achieved": round(tf, 2), "peak": round(min(BF16_PEAK_TFLOPS, mem_roof), 1), "unit": "TFLOP/s",
            "frac": round(tf / min(BF16_PEAK_TFLOPS, mem_roof), 4), "frac_of_mfma_peak": round(tf / BF16_PEAK_TFLOPS, 4),
            "avg_launch_us": round(us, 2), "flops_per_launch": flops, "algorithmic_bytes_per_launch": algo}


def traffic_probe(rows: str) -> int:
    """The short program the live traffic reading profiles: three 1 GiB device-to-device copies (the
    reference of known size) and, on the 1024x2048 plan with the given tile heights after one solve,
    six launches of the fused finest up leg with the outer step (kind 3 of the kernel timer)."""
    import ctypes

    import cartogenesis_amd as cg
    from cartogenesis_amd.ocean import HipOcean, Plan
    from cartogenesis_amd.problem import circulation_problem

    acc = HipOcean.create_or_null(0).accelerator
    if acc is None:
        return 1
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = COPY_BYTES
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        return 1
    hip.hipMemset(a, 0, ctypes.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    prob = circulation_problem(ROWS, COLS, seed=0, land_fraction=0.3)
    plan = Plan(acc, prob.levels)
    tuned = [int(v) for v in rows.split(",")]
    plan.set_tile_rows(tuned + [0] * (len(prob.levels) - len(tuned)))
    plan.solve(1e-8)
    us = cg.lib().cg_ocean_plan_time_kernel(plan.handle, 3, 5)
    plan.close()
    acc.close()
    return 0 if us > 0 else 1


def pmc_dispatches(csv_dir: str, counter: str):
    """{kernel name: [counter value of each dispatch]} from rocprofv3's counter CSV under csv_dir."""
    import csv
    import glob

    per = {}
    for path in glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                per.setdefault(r["Kernel_Name"], {}).setdefault(r["Dispatch_Id"], 0.0)
                per[r["Kernel_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in per.items()}


def live_traffic(rows0: int, rows: str):
    """HBM bytes per launch of the roofline kernel, read in this run: two rocprofv3 passes over
    `--traffic-probe` (FETCH_SIZE, then WRITE_SIZE: one counter a pass), each counter scaled by what it
    reads for a 1 GiB copy in the same pass. Returns (read, written, note); None where it could not be
    read (no profiler, a pass that failed or timed out)."""
    import shutil
    import signal
    import statistics
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, None, "rocprofv3 not found"
    name = f"k_relax_tile<4, {rows0}, false, 3>"
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="cg_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--traffic-probe", "--rows", rows]
        log(f"bench: traffic pass {counter}")
        proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd="/tmp", env=env,
                                start_new_session=True)
        try:
            rc = proc.wait(timeout=240)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            shutil.rmtree(d, ignore_errors=True)
            return None, None, f"the {counter} pass timed out"
        disp = pmc_dispatches(d, counter)
        shutil.rmtree(d, ignore_errors=True)
        # The three 1 GiB copies are the largest copy dispatches (the plan's uploads copy small
        # buffers through the same kernels); the fused kernel's dispatches all move the same bytes.
        copies = sorted((v for k, vals in disp.items() if "copyBuffer" in k for v in vals), reverse=True)[:3]
        fused = [v for k, vals in disp.items() if name in k for v in vals]
        if rc != 0 or len(copies) < 3 or not fused or copies[-1] <= 0:
            return None, None, f"the {counter} pass gave no reading (status {rc})"
        out[counter] = statistics.median(fused) / statistics.median(copies) * COPY_BYTES
    return out["FETCH_SIZE"], out["WRITE_SIZE"], ("rocprofv3 FETCH_SIZE and WRITE_SIZE of this run's probe, each "
                                                  "scaled by its reading for a 1 GiB device-to-device copy")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--tile-rows", type=int, default=8)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile heights instead of timing each grid's candidates at plan creation")
    ap.add_argument("--rows", default=None,
                    help="comma-separated tile heights per grid (0: default rule) instead of autotuning, e.g. 4,4,2,2,2")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--method", choices=["refined", "fgmres"], default="refined")
    ap.add_argument("--restart", type=int, default=8, help="FGMRES restart length")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the short program the live traffic reading profiles (three 1 GiB copies, six "
                         "launches of the fused finest kernel at the --rows tile heights) and exit")
    ap.add_argument("--no-traffic", action="store_true",
                    help="report the committed traffic reading instead of profiling this run's kernel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy six times (for scaling HBM counters) and exit")
    ap.add_argument("--level-passes", default="22224442",
                    help="passes per level of the binary32 cycle, one hex digit per level, finest last")
    ap.add_argument("--batch-reps", type=int, default=5,
                    help="timed repetitions of the batched year (12 monthly forcings solved together); 0 skips it")
    ap.add_argument("--combined", action=argparse.BooleanOptionalAction, default=True,
                    help="time 64 instances: 16 monthly forcings each at 256x512, 512x1024, 1024x2048 and "
                         "2048x4096, one plan of 16 per size (default on)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned preconditioner's leg")
    ap.add_argument("--learned-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--conv-only", action="store_true",
                    help="only the U-Net's largest convolution (dec0a at 1024x2048, random weights), 5 launches: "
                         "the short program the convolution's counters are read from")
    ap.add_argument("--unet-forward", action="store_true",
                    help="only the U-Net's forward at 1024x2048 (random weights, batch 1), one warm-up and five "
                         "timed: the short program the per-layer kernel trace is read from")
    ap.add_argument("--plain-value", type=float, default=0.0, help=argparse.SUPPRESS)
    ap.add_argument("--learned-worlds", type=int, default=64,
                    help="1024x2048 training worlds of the coarse network (seeds 100, 101, ...)")
    ap.add_argument("--learned-steps", type=int, default=3000, help="training steps (batch 16)")
    ap.add_argument("--worlds", action=argparse.BooleanOptionalAction, default=True,
                    help="time the four-world year: 12 months on each of 512x1024, 768x1536, 1024x2048 and "
                         "1024x2048 (seeds 1-4, land 20/27/33/40%%) (default on)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the circulation solve is a single-GPU path here "
              "(instances would shard one per GPU with no exchange); run with --gpus 1", file=sys.stderr)
        return 2

    if args.traffic_probe:
        return traffic_probe(args.rows or "6,6,4,4,4")

    if args.copy_calibration:
        # A plain copy of known size, far beyond L2 and the infinity cache: the reference point
        # that scales rocprofv3's FETCH_SIZE and WRITE_SIZE to bytes.
        import torch

        n = 256 * 1024 * 1024
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        c = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
            c.copy_(b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": n * 4, "copies": 6}))
        return 0

    if args.conv_only:
        import torch

        torch.cuda.init()
        from cartogenesis_amd.learned import UNet

        net = UNet(32, seed=0)
        a1 = torch.randn((1, ROWS, COLS, 64), device="cuda").to(torch.bfloat16)
        a2 = torch.randn((1, ROWS, COLS, 32), device="cuda").to(torch.bfloat16)
        st = torch.cuda.current_stream().cuda_stream
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        net._conv("dec0a", a1, a2, 1, ROWS, COLS, st)
        e0.record()
        for _ in range(5):
            net._conv("dec0a", a1, a2, 1, ROWS, COLS, st)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / 5 * 1000.0
        fl = unet_flops(ROWS, COLS)["dec0a"]
        algo = 2 * ROWS * COLS * (64 + 32 + 32) + 2 * 9 * 32 * 96  # inputs and output once, the weights
        print(json.dumps({"kernel": "k_conv3x3_tile<2>", "avg_launch_us": round(us, 2),
                          "tflops": round(fl / (us * 1e-6) / 1e12, 1), "algorithmic_bytes": algo}), flush=True)
        return 0

    if args.unet_forward:
        import torch

        torch.cuda.init()
        from cartogenesis_amd.learned import UNet

        net = UNet(32, seed=0)
        x = torch.zeros((1, ROWS, COLS, 16), device="cuda")
        x[..., 0] = 1.0
        x[..., 1] = torch.rand((1, ROWS, COLS), device="cuda")
        x[..., 2] = torch.randn((1, ROWS, COLS), device="cuda")
        x = x.to(torch.bfloat16)
        st = torch.cuda.current_stream().cuda_stream
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        net.forward(x, st)
        e0.record()
        for _ in range(5):
            net.forward(x, st)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / 5 * 1000.0
        fl = sum(unet_flops(ROWS, COLS).values())
        print(json.dumps({"unet_forward_us": round(us, 1), "flops": fl, "tflops": round(fl / (us * 1e-6) / 1e12, 1),
                          "layer_flops": unet_flops(ROWS, COLS)}), flush=True)
        return 0

    if args.learned_only:
        import torch  # before the library: its HIP runtime is the process's (see the learned leg in main)

        torch.cuda.init()
        from cartogenesis_amd.ocean import HipOcean
        from cartogenesis_amd.problem import circulation_problem

        prob = circulation_problem(ROWS, COLS, seed=0, land_fraction=0.3)
        acc = HipOcean.create_or_null(0).accelerator
        try:
            rows = [int(v) for v in args.rows.split(",")] if args.rows else None
            out = learned_leg(acc, prob, args.tol, args.learned_worlds, args.learned_steps, reps=args.steps,
                              batch_reps=args.batch_reps, rows=rows, autotune=not args.no_autotune)
        finally:
            acc.close()
        print(json.dumps(out), flush=True)
        return 0

    import cartogenesis_amd as cg
    from cartogenesis_amd.ocean import HipOcean, Plan
    from cartogenesis_amd.problem import circulation_problem

    log("bench: building the 1024x2048 problem")
    prob = circulation_problem(ROWS, COLS, seed=0, land_fraction=0.3)
    res = HipOcean.create_or_null(0)
    if res.accelerator is None:
        raise RuntimeError(f"no MI355X accelerator: {res.unavailable_because}")
    acc = res.accelerator
    acc.set_option(cg.CG_OPT_TILE_ROWS, args.tile_rows)
    acc.set_option(cg.CG_OPT_USE_GRAPH, 0 if args.no_graph else 1)
    acc.set_option(cg.CG_OPT_FGMRES_RESTART, args.restart)
    acc.set_option(cg.CG_OPT_LEVEL_PASSES, int(args.level_passes, 16))
    method = cg.CG_METHOD_FGMRES if args.method == "fgmres" else cg.CG_METHOD_REFINED
    plan = Plan(acc, prob.levels)
    lib = cg.lib()
    tuned = None
    if args.rows:
        tuned = [int(v) for v in args.rows.split(",")]
        tuned += [0] * (len(prob.levels) - len(tuned))
        plan.set_tile_rows(tuned)
        log(f"bench: tile rows per grid {tuned}")
    elif not args.no_autotune:
        tuned = plan.autotune(20)
        log(f"bench: autotuned tile rows per grid {tuned}")

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        info = plan.solve(args.tol, method)
    log(f"bench: timing {args.steps} solves")
    cycles = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info = plan.solve(args.tol, method)
        cycles.append(info.cycles)
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    rel_res = info.relative_residual
    x64, _ = plan.read()
    log(f"bench: {value:.1f} solves/s, {np.mean(cycles):.1f} cycles, residual {rel_res:.3e}")

    # Kernel durations with HIP events on the library's stream: 200 back-to-back launches of each
    # kernel on the plan's buffers as the timed solves left them, the average per launch.
    reps = 200
    us_outer = lib.cg_ocean_plan_time_kernel(plan.handle, 2, reps)
    us_up = lib.cg_ocean_plan_time_kernel(plan.handle, 1, reps)
    us_down = lib.cg_ocean_plan_time_kernel(plan.handle, 0, reps)
    us_fused = lib.cg_ocean_plan_time_kernel(plan.handle, 3, reps)
    us_bottom = lib.cg_ocean_plan_time_kernel(plan.handle, 4, reps)
    us_fused_sums = lib.cg_ocean_plan_time_kernel(plan.handle, 5, reps)
    fused = us_fused > 0
    # The committed traffic reading holds for the default tile rows and two passes on the finest grid
    # (the roofline kernel does not depend on the coarser grids' passes).
    defaults = args.tile_rows == 8 and args.level_passes.endswith("2")
    rows0 = tuned[0] if tuned and tuned[0] else 8  # the finest grid's tile rows
    outer_gbs = OUTER_BYTES / (us_outer * 1e-6) / 1e9 if us_outer > 0 else None
    up_gbs = UP_BYTES / (us_up * 1e-6) / 1e9 if us_up > 0 else None
    fused_gbs = FUSED_BYTES / (us_fused * 1e-6) / 1e9 if fused else None
    log(f"bench: outer residual {us_outer:.2f} us ({outer_gbs:.0f} GB/s algorithmic), finest up leg {us_up:.2f} us "
        f"({up_gbs:.0f} GB/s), finest down leg {us_down:.2f} us" +
        (f", up leg with the outer step fused {us_fused:.2f} us ({fused_gbs:.0f} GB/s)" if fused else "") +
        f", bottom {us_bottom:.2f} us")

    # The roofline kernel's HBM traffic, read live by two short profiled runs of the same kernel at the
    # same tile heights (the committed reading where the profiler cannot run).
    t_read = t_written = None
    t_note = "--no-traffic"
    if fused and not args.no_traffic:
        t_read, t_written, t_note = live_traffic(rows0, ",".join(str(v) for v in tuned) if tuned else "0")
        log(f"bench: traffic read {t_read} written {t_written} ({t_note})")

    # Beside the headline: the reference method itself (binary32 values, binary64 residual, the
    # coarsest relaxed) on the device to the reference's own tolerance 1e-3, and the 256 x 512
    # problem (seed 0) to 1e-10 by the refined method, for the host legs of the same names.
    prob256 = circulation_problem(256, 512, seed=0, land_fraction=0.3)
    side = {}
    for name, pr, tol, m in (("reference_method_1e-3_1024x2048", prob, 1e-3, cg.CG_METHOD_REFERENCE),
                             ("refined_1e-10_256x512", prob256, 1e-10, cg.CG_METHOD_REFINED)):
        sp = Plan(acc, pr.levels)
        try:
            sp.solve(tol, m)
            t1 = time.perf_counter()
            for _ in range(20):
                si = sp.solve(tol, m)
            side[name] = {"solves_per_s": round(20 / (time.perf_counter() - t1), 2), "cycles": si.cycles,
                          "relative_residual": si.relative_residual}
        finally:
            sp.close()
    log(f"bench: device side legs {side}")

    batched = None
    if args.batch_reps > 0:
        batched = batched_year(acc, prob, args.tol, args.batch_reps, args.warmup, not args.no_autotune)
    learned = None
    if not args.no_learned:
        # In a child process: torch's bundled HIP runtime has the soname the library links, and a
        # process holds one runtime (the first loaded), so the child imports torch before the library.
        log("bench: learned preconditioner leg (child process)")
        cmd = [sys.executable, os.path.abspath(__file__), "--learned-only", "--tol", str(args.tol),
               "--steps", str(args.steps), "--batch-reps", str(args.batch_reps),
               "--learned-worlds", str(args.learned_worlds), "--learned-steps", str(args.learned_steps)]
        # The main plan's tile heights (autotuned or given): tuning again in the child, after the
        # network's training, has picked tiles several times slower.
        if tuned:
            cmd += ["--rows", ",".join(str(v) for v in tuned)]
        elif args.no_autotune:
            cmd += ["--no-autotune"]
        proc = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=600)
        if proc.returncode != 0:
            raise RuntimeError(f"the learned leg failed with status {proc.returncode}")
        learned = json.loads(proc.stdout.strip().splitlines()[-1])
    worlds = batched_worlds(acc, args.tol, not args.no_autotune) if args.worlds else None
    combined = batched_combined(acc, args.tol, not args.no_autotune) if args.combined else None

    err = None
    vs_ref = None
    cpu = None
    if not args.no_oracle_check or not args.no_cpu:
        import oracle
        nproc, model, threads = host_cores()
        oracle.set_threads(threads)
        ladder = [oracle.Stencil(s.cells_across, s.cells_down, s.is_water, s.east_weight, s.west_weight,
                                 s.north_weight, s.south_weight, s.forcing, s.centre_weight) for s in prob.levels]
        if not args.no_oracle_check:
            log("bench: binary64 oracle to 1e-12 for the error report")
            ref, _, _ = oracle.solve64(ladder, 1e-12)
            err = float(np.linalg.norm(x64 - ref) / np.linalg.norm(ref))
            # The reference method itself (OceanCirculation.solve: binary32 values, binary64 residual)
            # on the CPU from zero to the same tolerance, capped at its 200 cycles: the answer a
            # caller of the reference gets on these inputs, and how far this solve is from it.
            log("bench: the reference method on the CPU to the same tolerance")
            t1 = time.perf_counter()
            refm, rcyc, rres = oracle.solve(ladder, np.zeros(ladder[0].cells, np.float32), args.tol)
            vs_ref = {"rel_l2": float(np.linalg.norm(x64 - refm) / np.linalg.norm(refm)),
                      "reference_cycles": rcyc, "reference_relative_residual": rres,
                      "reference_rel_l2_vs_fp64_oracle": float(np.linalg.norm(refm - ref) / np.linalg.norm(ref)),
                      "seconds": round(time.perf_counter() - t1, 2)}
            log(f"bench: vs the reference method {vs_ref}")
        if not args.no_cpu:
            log(f"bench: CPU baseline ({threads} of {nproc} cores, {model})")
            cpu = cpu_baseline(prob.levels, prob256.levels, args.cpu_budget, threads, nproc, model)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "circulation 1024x2048 (seed 0, 30% land, annual forcing) from zero to residual 1e-8 and "
                        "estimated error 1e-10: `value` is the plain GPU multigrid, one instance (binary32 V(2,2) "
                        "red-black cycles of the reference ladder inside binary64 iterative refinement); the learned "
                        "preconditioner's solves/s and its ratio to `value` are learned_solves_per_s and "
                        "learned_vs_plain_mg (details under \"learned\")",
            "rows": ROWS, "cols": COLS, "tol": args.tol, "method": args.method,
            "graph": not args.no_graph, "tile_rows": args.tile_rows, "level_passes": args.level_passes,
            "autotuned_tile_rows": tuned,
            "cycles_per_solve": float(np.mean(cycles)),
            "relative_residual": rel_res,
            "rel_l2_vs_fp64_oracle": err,
            "vs_cpu_reference_method": vs_ref,
        },
        "roofline": ({
            "bound": "hbm",
            "kernel": f"k_relax_tile<4,{rows0},false,3> (the finest up leg with the binary64 outer step fused in: "
                      "prolongation, 2 red-black passes, x64 += delta, binary64 residual, next forcing, max), the "
                      "largest HBM-bound kernel of a step",
            "achieved": round(fused_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(fused_gbs / HBM_PEAK_GBS, 4),
            "traffic": (round(t_read + t_written) if t_read is not None else
                        MEASURED_TRAFFIC["fused"].get(rows0) if defaults else None),
            "traffic_read": None if t_read is None else round(t_read),
            "traffic_written": None if t_written is None else round(t_written),
            "traffic_source": (t_note if t_read is not None else
                               f"committed reading (profiles/hbm_traffic_refined_1024x2048.md): {t_note}"),
            "tile_rows": rows0,
            "avg_launch_us": round(us_fused, 3),
            "avg_launch_us_with_error_sums": round(us_fused_sums, 3) if us_fused_sums > 0 else None,
            "algorithmic_bytes_per_launch": FUSED_BYTES,
            "separate_outer_residual": {"kernel": "k_outer_residual", "avg_launch_us": round(us_outer, 3),
                                        "achieved": None if outer_gbs is None else round(outer_gbs, 1),
                                        "algorithmic_bytes_per_launch": OUTER_BYTES,
                                        "traffic": MEASURED_TRAFFIC["outer"] if defaults else None},
        } if fused else {
            "bound": "hbm",
            "kernel": "k_outer_residual (binary64 x += delta, residual, correction forcing, max), the largest "
                      "HBM-bound kernel by trace time; the one-workgroup bottom of the cycle, equal in time, is "
                      "latency-bound",
            "achieved": None if outer_gbs is None else round(outer_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if outer_gbs is None else round(outer_gbs / HBM_PEAK_GBS, 4),
            "traffic": MEASURED_TRAFFIC["outer"] if defaults else None,
            "avg_launch_us": round(us_outer, 3),
            "algorithmic_bytes_per_launch": OUTER_BYTES,
        }) | {
            "smoother": {
                "kernel": f"k_relax_tile<4,{rows0},false,2> (finest up leg: prolongation + 2 fused red-black passes)",
                "achieved": None if up_gbs is None else round(up_gbs, 1),
                "frac": None if up_gbs is None else round(up_gbs / HBM_PEAK_GBS, 4),
                "avg_launch_us": round(us_up, 3),
                "algorithmic_bytes_per_launch": UP_BYTES,
                "traffic": MEASURED_TRAFFIC["up"].get(rows0) if defaults else None,
                "down_leg_avg_launch_us": round(us_down, 3),
            },
            "bottom": {"kernel": "k_direct_bottom (the cycle's 32x64 grid by its binary64 dense inverse)",
                       "avg_launch_us": round(us_bottom, 3)},
        },
        "cpu_baseline": cpu,
        "device_side_legs": side,
    }
    if learned:
        out["learned"] = learned
        out["learned_solves_per_s"] = learned.get("solves_per_s")
        out["learned_vs_plain_mg"] = learned.get("learned_vs_plain_mg")
    if batched:
        out["batched"] = batched
        # The same roofline kernel at the batched year's shape (twelve instances a launch), beside the
        # single-instance figure above.
        if batched.get("roofline"):
            out["roofline_batched"] = batched["roofline"]
    if worlds:
        out["batched_worlds"] = worlds
    if combined:
        out["batched_combined"] = combined
    print(json.dumps(out), flush=True)
    plan.close()
    acc.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
