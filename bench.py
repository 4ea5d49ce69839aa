"""Benchmark: MLS-MPM substeps of the 4.2M-particle spheres-and-slab scene on one GPU.

A step is one substep (sort, P2G, grid update, G2P) over all particles.  The
scene: three Neo-Hookean spheres and a weakly compressible fluid slab, 256^3
grid with 4^3 blocks, h = 1/256, dt = 5e-5, box walls, f32.  Prints one JSON
line.  The CPU baseline is the oracle (the C++ restatement of the reference
CPU step, OpenMP) timed on a bounded sample of the same scene.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per unit of work of each kernel (the dataflow with the
# stress evaluated after the gather and the fused affine momentum A formed on
# chip):
#   P2G  reads x, v, A, m                        64 B / particle; writes (m, p) 16 B / active node
#   grid reads and writes (m, p | v)             32 B / active node
#   G2P  reads x, F, m, V0, mu, lambda           64 B / particle; writes x, v, F, A 96 B / particle;
#        reads v                                 16 B / active node
#   substep total: 224 B / particle + 64 B / active node.
# The fused gather-scatter kernel (g2p2g) does a G2P and the next P2G.
BYTES = {
    "p2g": (64, 16),
    "grid_update": (0, 32),
    "g2p": (160, 16),
    "g2p2g": (224, 32),
}
# Stages with HIP events inside the timed region: the three modes of the
# transfer kernel, one of which dominates.
TIMED_STAGES = ("p2g", "g2p2g", "g2p")


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, substeps: int, threads: int) -> dict:
    from oracle import pyoracle
    from squishy_volumes_amd import scenes

    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(scene.consts, scene.time_step, scene.io_state, scene.options, 4)
    o.step_to(scenes.advance_time(o.time, scene.time_step, 2), scene.frame_input)  # 2 warm-up substeps
    t0 = time.perf_counter()
    o.step_to(scenes.advance_time(o.time, scene.time_step, substeps), scene.frame_input)
    dt = time.perf_counter() - t0
    n = scene.io_state.particles.count
    o.close()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": n * substeps / dt,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": model,
        "nproc": os.cpu_count(),
        "sample": f"{scene.name}: all {n} particles, {substeps} substeps after 2 warm-up substeps, "
        f"float oracle (C++ restatement of the reference CPU step), {threads} OpenMP threads, {dt:.2f} s",
    }


def copy_calibration(mib: int):
    """Three launches of the engine's plain copy kernel over `mib` MiB."""
    import ctypes as C

    import torch

    from squishy_volumes_amd import abi

    lib = abi.load_library()
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        rc = lib.mpm_dev_copy(abi.ptr(src), abi.ptr(dst), src.numel() * 4, C.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"mpm_dev_copy failed: {rc}")
    torch.cuda.synchronize()
    log(f"bench: copy calibration, 3 x {mib} MiB")


def copy_bandwidth(mib: int = 2048, launches: int = 5) -> dict:
    """The on-box copy bandwidth over `mib` MiB, read + written bytes over the
    average HIP-event time (after one untimed launch), of the engine's plain
    copy kernel and of torch's tensor copy; `achieved` is the larger."""
    import ctypes as C

    import torch

    from squishy_volumes_amd import abi

    lib = abi.load_library()
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    copy = lambda: lib.mpm_dev_copy(abi.ptr(src), abi.ptr(dst), src.numel() * 4, C.c_void_p(stream.cuda_stream))
    copy()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(launches):
        if copy() != 0:
            raise RuntimeError("mpm_dev_copy failed")
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / launches
    dst.copy_(src)
    a.record(stream)
    for _ in range(launches):
        dst.copy_(src)
    b.record(stream)
    torch.cuda.synchronize()
    ms_torch = a.elapsed_time(b) / launches
    del src, dst
    torch.cuda.empty_cache()
    rate = lambda t: 2.0 * (mib << 20) / (t * 1e-3) / 1e9
    gbs = max(rate(ms), rate(ms_torch))
    return {"bytes_per_launch": 2 * (mib << 20), "engine_copy_ms": ms, "torch_copy_ms": ms_torch,
            "achieved": gbs, "unit": "GB/s", "frac_of_peak": gbs / HBM_PEAK_GBS}


def traffic_for(kernel: str, workload: str):
    f = ROOT / "profiles" / f"traffic_{kernel}_{workload}.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def learned_form(args) -> dict:
    """The learned closure on the 1M-particle snow block impacting a sphere:
    pairs from the plain (analytic) run, training on the GPU, then the same
    window of substeps timed with the analytic and with the learned stress."""
    import numpy as np
    import torch

    from squishy_volumes_amd import closure, scenes
    from squishy_volumes_amd.gpu_state import GpuState

    sc = scenes.snow_block_on_sphere()
    n = sc.io_state.particles.count
    hard = sc.options.snow_hardening
    adv = lambda g, k: g.step_to(scenes.advance_time(g.time, sc.time_step, k), sc.frame_input)
    tuning = {"resort_interval": args.resort_interval} if args.resort_interval else {}

    # Pairs from the plain run: every `every` substeps over `gen` substeps,
    # about args.pairs in total.
    gen, every = args.closure_substeps, 50
    plain = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options, tuning)
    rng = np.random.default_rng(0)
    chunks = []
    per = max(1, args.pairs // (gen // every))
    for _ in range(gen // every):
        adv(plain, every)
        p = closure.pairs_from_particles(plain.to_io_state().particles, hard)
        chunks.append(p.subset(rng.choice(len(p), size=min(per, len(p)), replace=False)))
    data = closure.Pairs.concat(chunks)
    log(f"closure: {len(data)} pairs from {gen} plain substeps")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net = closure.Mlp(seed=0)
    x = torch.from_numpy(data.x).cuda()
    t = torch.from_numpy(data.t).cuda()
    tlog = net.fit(x, t, epochs=args.epochs, batch=8192, lr=3e-3, lr_final=3e-5)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    model = net.model(data.x)
    log(f"closure: trained {args.epochs} epochs in {train_s:.1f} s, mse {tlog.epochs[0]:.3e} -> {tlog.epochs[-1]:.3e}")

    def timed(learned: bool, fused: bool = False):
        tun = dict(tuning)
        if learned and fused:
            tun["learned_fused"] = 1  # the network inside the fused gather-scatter
        g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options, tun)
        if learned:
            g.set_closure(model, judge_stride=64)
        adv(g, args.closure_start)           # up to the impact
        if learned:
            g.closure_stats(reset=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        adv(g, args.closure_window)
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        stats = g.closure_stats(reset=True) if learned else None
        parts = g.to_io_state().particles
        # Per-stage breakdown over 50 further substeps (untimed).
        g.kernel_times(reset=True)
        g.set_profiling(True)
        adv(g, 50)
        g.set_profiling(False)
        stages = {k: round(v[0] / 50, 4) for k, v in g.kernel_times(reset=True).items() if v[1]}
        g.close()
        return el, stats, parts, stages

    mlp_rates = mlp_kernel_rates(net)
    log(f"closure: MLP kernels {mlp_rates}")
    el_plain, _, parts_plain, stages_plain = timed(False)
    el_learned, stats, parts_learned, stages_learned = timed(True)
    el_fused, _, _, stages_fused = timed(True, fused=True)

    # Offline: the network against the analytic stress in FP64 on the final
    # learned state (the judge above compares with the FP32 analytic stress).
    fin = closure.pairs_from_particles(parts_learned, hard)
    inside = np.all((fin.x >= model.hull_lo) & (fin.x <= model.hull_hi), axis=1)
    y = net.forward(torch.from_numpy(fin.x[inside]).cuda()).cpu().numpy().astype(np.float64)
    tau = closure.stress_from_outputs(y, fin.s[inside], fin.E[inside], fin.nu[inside], fin.jp[inside], hard)
    tau64 = closure.analytic_principal_stress(fin.s[inside], fin.E[inside], fin.nu[inside], fin.jp[inside], hard)
    err64 = float(np.linalg.norm(tau - tau64) / np.linalg.norm(tau64))
    ex = float(np.linalg.norm(parts_learned.positions - parts_plain.positions) / np.linalg.norm(parts_plain.positions))
    w = args.closure_window
    return {
        "metric": "particle-substeps/s (learned closure vs analytic stress)",
        "learned_path": "the network in its own pass between the gather and a scatter from the state",
        "value": n * w / el_learned,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": w,
        "ms_per_step": el_learned / w * 1e3,
        "higher_is_better": True,
        "dtype": "f32 state, bf16 MLP with f32 accumulation",
        "data": "synthetic (seeded snow block)",
        "config": {"workload": "snow_block_on_sphere_1M", "particles": n, "h": sc.consts.grid_node_size,
                   "dt": sc.time_step, "window": [args.closure_start, args.closure_start + w]},
        "plain": {"value": n * w / el_plain, "ms_per_step": el_plain / w * 1e3,
                  "stage_ms_per_substep": stages_plain},
        "learned_stage_ms_per_substep": stages_learned,
        "learned_vs_plain": el_plain / el_learned,
        "learned_in_fused_kernel": {"value": n * w / el_fused, "ms_per_step": el_fused / w * 1e3,
                                    "vs_plain": el_plain / el_fused, "stage_ms_per_substep": stages_fused},
        "stress_error_vs_analytic_f32": stats["relative_stress_error"],
        "stress_error_vs_analytic_f64": err64,
        "network_evaluations": stats["evaluated"],
        "hull_fallbacks": stats["fallbacks"],
        "judged": stats["judged"],
        "position_error_vs_plain": ex,
        "training": {"pairs": len(data), "epochs": args.epochs, "seconds": train_s,
                     "mse_first_epoch": tlog.epochs[0], "mse_last_epoch": tlog.epochs[-1]},
        "mlp_kernels": mlp_rates,
    }


# Algorithmic FLOP per sample of the 6-64-64-3 network (the MFMA tiles pad the
# input and output widths to 16; the padding is not counted).
MLP_FORWARD_FLOP = 2 * (6 * 64 + 64 * 64 + 64 * 3)
# Forward, weight gradients of all three layers, input gradients of layers 2 and 3.
MLP_GRAD_FLOP = 2 * MLP_FORWARD_FLOP + 2 * (64 * 64 + 64 * 3)
BF16_DENSE_PEAK_TFLOPS = 2500.0


def mlp_kernel_rates(net, samples: int = 1 << 22, reps: int = 20) -> dict:
    """The network kernels alone on one large batch, against the BF16 dense
    MFMA peak: the forward pass and the gradient pass (forward and backward),
    timed with events on the stream they run on."""
    import torch

    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.rand((samples, 6), generator=g, device="cuda") * 2.0 - 1.0
    t = torch.rand((samples, 3), generator=g, device="cuda")
    out = {}
    for name, flop, run in (
        ("forward", MLP_FORWARD_FLOP, lambda: net.forward(x)),
        ("grad", MLP_GRAD_FLOP, lambda: net.accumulate(x, t, 1.0 / samples)),
    ):
        run()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run()
        e.record()
        e.synchronize()
        ms = s.elapsed_time(e) / reps
        tflops = flop * samples / (ms * 1e-3) / 1e12
        out[name] = {"samples": samples, "flop_per_sample": flop, "avg_launch_ms": ms, "achieved": tflops,
                     "peak": BF16_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / BF16_DENSE_PEAK_TFLOPS}
    net.grad.zero_()
    net.loss.zero_()
    return out


def batched_form(args) -> dict:
    """64 uneven random-object scenes (5k..500k particles each, log-uniform,
    scaled to 8M in all; h = 1/128, dt = 5e-5) advanced together by one
    engine with autotuned launch knobs, against the plain form: each scene on
    its own engine, one after the other, with the default knobs."""
    import torch

    from squishy_volumes_amd import autotune, scenes
    from squishy_volumes_amd.batch import Batch
    from squishy_volumes_amd.gpu_state import GpuState

    scs = scenes.uneven_scene_set(args.scenes, seed=0, total=args.batch_particles)
    b = Batch(scs)
    sc = b.scene
    n = b.count
    steps = args.batch_substeps
    log(f"batched: {len(scs)} scenes, {n} particles, repeat {b.repeat}")
    if args.no_autotune:
        tuning = {"resort_interval": args.resort_interval or 8, "persistent_blocks": args.persistent_blocks}
        tune = None
    else:
        tune = autotune.autotune(sc, log=log)
        tuning = tune.best
    adv = lambda g, k: g.step_to(scenes.advance_time(g.time, sc.time_step, k), b.frame_input)
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options, tuning)
    if args.warmup:
        adv(g, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    status = adv(g, steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert status.result == 0 and status.substeps == steps, (status.result, status.substeps)
    nodes = int(status.active_nodes)
    g.kernel_times(reset=True)
    g.set_profiling(True)
    adv(g, 50)
    g.set_profiling(False)
    breakdown = g.kernel_times(reset=True)
    g.close()
    log(f"batched: {steps} substeps in {el:.3f} s")

    # The plain form: every scene alone, the same substeps, default knobs.
    el_plain = 0.0
    for s1 in scs:
        g1 = GpuState.from_io_state(0, s1.consts, s1.time_step, s1.io_state, s1.options)
        if args.warmup:
            g1.step_to(scenes.advance_time(g1.time, s1.time_step, args.warmup), s1.frame_input)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st1 = g1.step_to(scenes.advance_time(g1.time, s1.time_step, steps), s1.frame_input)
        torch.cuda.synchronize()
        el_plain += time.perf_counter() - t1
        assert st1.result == 0
        g1.close()
    log(f"plain per-scene form: {steps} substeps of every scene in {el_plain:.3f} s")
    substep_bytes = 224 * n + 64 * nodes
    counts = [s1.io_state.particles.count for s1 in scs]
    return {
        "metric": "particle-substeps/s (batched scenes vs one engine per scene)",
        "value": n * steps / el,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": el / steps * 1e3,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (seeded random cubes and spheres)",
        "config": {"workload": f"uneven_scenes_{len(scs)}", "scenes": len(scs), "particles": n,
                   "particles_per_scene": {"min": min(counts), "max": max(counts)},
                   "h": sc.consts.grid_node_size, "dt": sc.time_step, "scene_repeat": list(b.repeat),
                   "active_nodes": nodes, "active_blocks": int(status.active_blocks)},
        "tuning": tuning,
        "autotune": None if tune is None else {"ms_per_substep": tune.ms_per_substep, "trials": tune.trials},
        "plain": {"value": n * steps / el_plain, "ms_per_step": el_plain / steps * 1e3},
        "batched_vs_plain": el_plain / el,
        "substep_roofline": {"bytes_per_substep": substep_bytes,
                             "achieved": substep_bytes * steps / el / 1e9,
                             "frac": substep_bytes * steps / el / 1e9 / HBM_PEAK_GBS},
        "stage_ms_per_substep": {k: v[0] / 50 for k, v in breakdown.items() if v[1]},
    }


def combined_form(args) -> dict:
    """128 batched scenes of elastic, snow and fluid objects (10k..1M
    particles each, log-uniform, scaled to 16M in all; h = 1/128, dt = 5e-5),
    300 substeps: the plain batched form against the batched form with the
    learned closure shared by all scenes, pre-trained on pairs from the first
    100 plain substeps and fine-tuned on the GPU every 50 substeps on pairs
    from the current state (the fine-tuning inside the timed region)."""
    import numpy as np
    import torch

    from squishy_volumes_amd import abi, closure, scenes
    from squishy_volumes_amd.batch import Batch
    from squishy_volumes_amd.gpu_state import GpuState

    scs = scenes.mixed_scene_set(args.combined_scenes, total=args.combined_particles)
    b = Batch(scs)
    sc = b.scene
    n = b.count
    flags = sc.io_state.particles.flags
    n_snow = int(((flags & abi.SNOW) != 0).sum())
    steps, every = args.combined_substeps, 50
    tuning = {"resort_interval": args.resort_interval or 16, "persistent_blocks": args.persistent_blocks or 2048}
    adv = lambda g, k: g.step_to(scenes.advance_time(g.time, sc.time_step, k), b.frame_input)
    make = lambda: GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options, tuning)
    log(f"combined: {len(scs)} scenes, {n} particles ({n_snow} snow), repeat {b.repeat}")

    # Pre-training: pairs from the first 100 plain substeps.
    stride = max(1, (4 * n_snow) // args.pairs)
    g0 = make()
    xs, ts = [], []
    for k in range(4):
        adv(g0, 25)
        x, t = g0.closure_pairs(stride, k % stride)
        xs.append(x)
        ts.append(t)
    g0.close()
    x, t = torch.cat(xs), torch.cat(ts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net = closure.Mlp(seed=0)
    tlog = net.fit(x, t, epochs=args.combined_epochs, batch=8192, lr=3e-3, lr_final=1e-4)
    torch.cuda.synchronize()
    pre_s = time.perf_counter() - t0
    hull = closure.ClosureModel.hull_of(x)
    log(f"combined: pre-trained on {x.shape[0]} pairs in {pre_s:.1f} s, mse {tlog.epochs[0]:.2e} -> {tlog.epochs[-1]:.2e}")

    # The plain batched form.
    gp = make()
    adv(gp, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = adv(gp, steps)
    torch.cuda.synchronize()
    el_plain = time.perf_counter() - t0
    assert st.result == 0 and st.substeps == steps
    nodes = int(st.active_nodes)
    gp.close()
    log(f"combined: plain batched {steps} substeps in {el_plain:.3f} s")

    # The learned form, fine-tuned online.
    ft_stride = max(1, n_snow // args.finetune_pairs)
    gl = make()
    gl.set_closure(net.device_model(hull), judge_stride=64)
    adv(gl, args.warmup)
    gl.closure_stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ft_s, ft_pairs, ft_log = 0.0, 0, []
    for c in range(steps // every):
        st = adv(gl, every)
        assert st.result == 0
        if c == steps // every - 1:
            break
        t1 = time.perf_counter()
        x, t = gl.closure_pairs(ft_stride, c % ft_stride)
        flog = net.fit(x, t, epochs=args.finetune_epochs, batch=8192, lr=3e-4, lr_final=1e-4, seed=c + 1)
        hull = closure.ClosureModel.hull_union(hull, closure.ClosureModel.hull_of(x))
        gl.set_closure(net.device_model(hull), judge_stride=64)
        ft_s += time.perf_counter() - t1
        ft_pairs += int(x.shape[0])
        ft_log.append(flog.epochs[-1])
    torch.cuda.synchronize()
    el_learned = time.perf_counter() - t0
    stats = gl.closure_stats(reset=True)
    gl.close()
    log(f"combined: learned batched {steps} substeps in {el_learned:.3f} s ({ft_s:.3f} s fine-tuning), "
        f"stress error {stats['relative_stress_error']:.2e}")
    return {
        "metric": "particle-substeps/s (combined: batched scenes with the learned closure, fine-tuned online)",
        "value": n * steps / el_learned,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": el_learned / steps * 1e3,
        "higher_is_better": True,
        "dtype": "f32 state, bf16 MLP with f32 accumulation",
        "data": "synthetic (seeded elastic, snow and fluid objects)",
        "config": {"workload": f"mixed_scenes_{len(scs)}", "scenes": len(scs), "particles": n, "snow_particles": n_snow,
                   "h": sc.consts.grid_node_size, "dt": sc.time_step, "scene_repeat": list(b.repeat),
                   "active_nodes": nodes, "tuning": tuning},
        "plain_batched": {"value": n * steps / el_plain, "ms_per_step": el_plain / steps * 1e3},
        "learned_vs_plain": el_plain / el_learned,
        "stress_error_vs_analytic_f32": stats["relative_stress_error"],
        "network_evaluations": stats["evaluated"],
        "hull_fallbacks": stats["fallbacks"],
        "judged": stats["judged"],
        "pretraining": {"pairs": int(sum(v.shape[0] for v in xs)), "epochs": args.combined_epochs, "seconds": pre_s,
                        "mse_first_epoch": tlog.epochs[0], "mse_last_epoch": tlog.epochs[-1]},
        "finetuning": {"every_substeps": every, "rounds": len(ft_log), "pairs": ft_pairs,
                       "epochs_per_round": args.finetune_epochs, "seconds_in_timed_region": ft_s,
                       "mse_last_epoch_per_round": ft_log},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-substeps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--resort-interval", type=int, default=0, help="substeps between particle sorts (0: engine default)")
    ap.add_argument("--persistent-blocks", type=int, default=0)
    ap.add_argument(
        "--copy-mib",
        type=int,
        default=0,
        help="first launch the plain copy kernel (k_copy_bytes) three times over a buffer of this many MiB, "
        "outside the timed region, so that hardware counters can be scaled by a copy of known size",
    )
    ap.add_argument("--form", choices=("plain", "learned", "batched", "combined"), default="plain",
                    help="plain: the headline 4.2M-particle substep; learned: the learned closure on the snow "
                    "scene; batched: 64 uneven scenes in one engine against one engine per scene; combined: 128 "
                    "mixed-material scenes with the learned closure fine-tuned online against the plain batched form")
    ap.add_argument("--combined-scenes", type=int, default=128)
    ap.add_argument("--combined-particles", type=int, default=16_000_000)
    ap.add_argument("--combined-substeps", type=int, default=300)
    ap.add_argument("--combined-epochs", type=int, default=150)
    ap.add_argument("--finetune-pairs", type=int, default=200_000)
    ap.add_argument("--finetune-epochs", type=int, default=3)
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--batch-substeps", type=int, default=200)
    ap.add_argument("--batch-particles", type=int, default=8_000_000)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--closure-substeps", type=int, default=1000)
    ap.add_argument("--closure-start", type=int, default=300)
    ap.add_argument("--closure-window", type=int, default=200)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py runs on exactly one GPU: the engine is single-device (one simulation per GPU, "
            "no collectives); multi-GPU sharding is a design note only.",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch

    from squishy_volumes_amd import scenes
    from squishy_volumes_amd.gpu_state import GpuState

    if args.copy_mib:
        copy_calibration(args.copy_mib)
    if args.form == "learned":
        print(json.dumps(learned_form(args)))
        return
    if args.form == "batched":
        print(json.dumps(batched_form(args)))
        return
    if args.form == "combined":
        print(json.dumps(combined_form(args)))
        return

    scene = scenes.spheres_and_slab()
    n = scene.io_state.particles.count
    log(f"bench: {scene.name}, {n} particles, warmup {args.warmup}, steps {args.steps}")
    tuning = {}
    if args.resort_interval:
        tuning["resort_interval"] = args.resort_interval
    if args.persistent_blocks:
        tuning["persistent_blocks"] = args.persistent_blocks
    gpu = GpuState.from_io_state(0, scene.consts, scene.time_step, scene.io_state, scene.options, tuning)
    fi = scene.frame_input

    t = gpu.time
    if args.warmup:
        gpu.step_to(scenes.advance_time(t, scene.time_step, args.warmup), fi)
    log("bench: warm-up done")

    # Timed region: exactly `steps` substeps on the engine's default path (the
    # re-sort cycles replayed as captured HIP graphs), no event inside it.
    target = scenes.advance_time(gpu.time, scene.time_step, args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    status = gpu.step_to(target, fi)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert status.substeps == args.steps, (status.substeps, args.steps)
    assert status.result == 0, status.result
    graph_launches = int(status.graph_launches)
    log(f"bench: {args.steps} substeps in {elapsed:.3f} s ({graph_launches} graph launches)")
    nodes = int(status.active_nodes)
    active_blocks = int(status.active_blocks)
    gpu.close()
    # Kernel times from a separate run over the same substeps: a second engine
    # from the same state, the same warm-up, then the timed region's substeps
    # again with HIP events on the engine's stream around the transfer kernels
    # (the dominant kernel is one of them); then around every stage.
    gpu = GpuState.from_io_state(0, scene.consts, scene.time_step, scene.io_state, scene.options, tuning)
    if args.warmup:
        gpu.step_to(scenes.advance_time(gpu.time, scene.time_step, args.warmup), fi)
    kernel_steps = args.steps
    gpu.kernel_times(reset=True)
    gpu.set_profiling(True, stages=TIMED_STAGES)
    gpu.step_to(scenes.advance_time(gpu.time, scene.time_step, kernel_steps), fi)
    gpu.set_profiling(False)
    stages = gpu.kernel_times(reset=True)
    breakdown_steps = min(args.steps, 50)
    gpu.set_profiling(True)
    gpu.step_to(scenes.advance_time(gpu.time, scene.time_step, breakdown_steps), fi)
    gpu.set_profiling(False)
    breakdown = gpu.kernel_times(reset=True)
    # (not in counter runs: their copy-kernel dispatches are the calibration's alone)
    copy_bw = None if args.copy_mib else copy_bandwidth()
    if copy_bw:
        log(f"bench: copy kernel {copy_bw['achieved']:.0f} GB/s")

    value = n * args.steps / elapsed
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in stages.items()}
    dominant = max(TIMED_STAGES, key=lambda k: stages.get(k, (0, 1))[0])
    per_particle, per_node = BYTES[dominant]
    bytes_per_launch = per_particle * n + per_node * nodes
    avg_s = stage_ms[dominant] / 1e3
    achieved = bytes_per_launch / avg_s / 1e9
    substep_bytes = 224 * n + 64 * nodes
    result = {
        "metric": "particle-substeps/s",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spheres-and-slab scene)",
        "config": {
            "workload": "spheres_and_slab_4.2M",
            "particles": n,
            "grid": "256^3 dense block grid, 4^3 blocks, h=1/256",
            "dt": scene.time_step,
            "materials": "3 Neo-Hookean spheres (E log-uniform 1e4..1e6, nu 0.3) + inviscid fluid slab (K 1e5, gamma 7)",
            "boundary": "box walls (separating slip)",
            "active_nodes": nodes,
            "active_blocks": active_blocks,
        },
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_for(dominant, "spheres_and_slab"),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": stage_ms[dominant],
            "frac_of_copy": achieved / copy_bw["achieved"] if copy_bw else None,
        },
        "substep_roofline": {
            "bytes_per_substep": substep_bytes,
            "achieved": substep_bytes * args.steps / elapsed / 1e9,
            "frac": substep_bytes * args.steps / elapsed / 1e9 / HBM_PEAK_GBS,
            "frac_of_copy": substep_bytes * args.steps / elapsed / 1e9 / copy_bw["achieved"] if copy_bw else None,
        },
        "copy_bandwidth": copy_bw,
        "stage_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in breakdown.items() if v[1]},
        "stage_launches_per_substep": {k: v[1] / breakdown_steps for k, v in breakdown.items() if v[1]},
        "timed_path": f"default: {graph_launches} captured re-sort cycles replayed as HIP graphs, no events",
        "stage_breakdown": f"roofline.avg_launch_ms: the timed region's {kernel_steps} substeps run again by a second "
                           "engine from the same state after the same warm-up, with events around the transfer "
                           f"kernels; stage_ms_per_launch: {breakdown_steps} further substeps with an event pair "
                           "around every launch",
        "svd_fp64_fallbacks": int(status.svd_fp64_fallbacks),
    }
    gpu.close()
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        log(f"bench: CPU baseline, {args.cpu_substeps} substeps on {threads} threads")
        result["cpu_baseline"] = cpu_baseline(scene, args.cpu_substeps, threads)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
