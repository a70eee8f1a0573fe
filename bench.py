#!/usr/bin/env python3
"""Throughput of the explicit MPS step on one MI355X.

Workload: the 3-D dam break with 300,000 fluid particles in FP32 (tank
1.0 x 0.6 x 0.3 m, l0 = 5 mm, 3 wall layers, dt = 2.5e-4 s).  A step is one
pass of the ClcMPS loop body over the resident state (inputs are in HBM
before the timed region).  Prints one JSON line:

  value         live fluid particles x steps / second (particle-steps/s)
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its average launch time (HIP events on the handle's stream),
                against 8 TB/s; "valu" gives the same kernel's candidate-test
                flops against the FP32 VALU peak
  cpu_baseline  the FP32 CPU oracle on this host's cores, on a bounded
                sample of the same workload (first steps of the same scene)
  batched       512 uneven 2-D dam breaks (1.95k-8.2k fluid each) in one
                handle, launch shape autotuned, steps replayed from a
                hipGraph, against the same scenes as single-scene handles
                stepped one after another with plain launches ("speedup")
  adjoint       the 300k scene for 200 steps with checkpoints every 25:
                (forward with checkpoints + reverse sweep) / plain forward
  tensor_gradient  the 300k scene with the tensor-corrected gradient: step rate
                and the gradient kernel's average duration

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batch]
                       [--no-adjoint] [--no-tensor]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_CANDIDATE = 9  # 3 sub, 3 mul, 2 add, 1 compare per candidate pair

METRIC = "particle-steps/s, 3D dam break 300k FP32"


def algorithmic_bytes(kernel: str, nf: int, nw: int, m: int) -> float:
    """Compulsory HBM bytes of one launch: every array the phase reads or
    writes, counted once, at 12 B per FP32 3-vector, 4 B per scalar."""
    table = {
        "vsc_up1": 49 * nf + 12 * m,       # pos, vel in; pos, vel out; typ
        "col": 37 * nf + 13 * nw + 12 * m,
        "prs": 17 * nf + 12 * m,
        "col_prs": 42 * nf + 13 * nw + 12 * m,  # col + prs, positions and mirrors read once
        "grad_up2": 61 * nf + 16 * m,
        "mirror_gen": 13 * nf + 26 * nw + 29 * m,
    }
    return float(table.get(kernel, 0))


KERNEL_SYMBOL = {  # timing name -> kernel symbol in the rocprofv3 summaries
    "vsc_up1": "mps::k_vsc<float, 3, true>",
    "col": "mps::k_col_list<float, 3, false>",
    "col_prs": "mps::k_colprs_list<float, 3, true>",  # the step's form (it also builds the list)
    "prs": "mps::k_prs_list<float, 3>",
    "grad_up2": "mps::k_grad_list<float, 3, true, false>",
    "mirror_gen": "mps::k_mirror_group<float, 3>",
    "nlist": "mps::k_nlist_build<float, 3>",
}


def measured_traffic(kernel: str):
    """HBM-side bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "cfg1_hbm_traffic.md")
    sym = KERNEL_SYMBOL.get(kernel)
    if not sym or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"| `{sym}` |"):
            return float(line.strip().strip("|").split("|")[-1])
    return None


def candidates_per_pass(h) -> int:
    """Fluid candidates the 27-cell stencil visits, summed over live fluid
    particles (one pass of a fluid neighbour loop)."""
    c = h.constants()
    cs, ids = h.buckets(0)
    nb = c["nb"]
    occ = np.diff(cs).reshape(nb[2], nb[1], nb[0]).astype(np.int64)
    p = np.pad(occ, 1)
    box = sum(p[1 + dz:1 + dz + nb[2], 1 + dy:1 + dy + nb[1], 1 + dx:1 + dx + nb[0]]
              for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    return int((occ * box).sum())


def cpu_baseline(seconds: float) -> dict:
    from mps_cuda_amd import configs
    from oracle import pyoracle
    threads = pyoracle.set_threads(0)
    scene, params = configs.cfg1("fp32")
    o = pyoracle.Oracle(params, scene)
    nf = int((scene.typ == 0).sum())
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter() - t0
    n = max(1, int(seconds / max(t1, 1e-6)))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    live = int((o.state()["typ"] == 0).sum())
    return {"value": live * n / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP32 CPU oracle (OpenMP), same 300k scene, steps 2-{n + 1} ({n} steps, "
                      f"{dt:.1f} s, {nf} fluid at start)"}


def batched_leg(steps: int, seq_steps: int, log) -> dict:
    """cfg3: 512 uneven 2-D dam breaks stepped as one batch (autotuned launch
    shape, hipGraph replay) against the same scenes stepped one after another
    with plain eager launches of single-scene handles."""
    from mps_cuda_amd import autotune, configs, ffi

    scenes, params = configs.cfg3("fp32")
    nf = [int((s.typ == ffi.FLD).sum()) for s in scenes]
    log(f"[bench] cfg3: {len(scenes)} scenes, {sum(nf)} fluid ({min(nf)}-{max(nf)} per scene)")
    tuned = autotune.tune(params, scenes, steps=16, use_cache=False, log=log)
    log(f"[bench] cfg3 autotune: {tuned}")
    h = ffi.Handle(params, scenes)
    autotune.apply(h, tuned)
    h.step(max(4, tuned["graph"]))
    h.sync()
    live0 = int((h.state()["typ"] == ffi.FLD).sum())
    t0 = time.perf_counter()
    h.step(steps)
    h.sync()
    el_b = time.perf_counter() - t0
    live1 = int((h.state()["typ"] == ffi.FLD).sum())
    h.close()
    rate_b = 0.5 * (live0 + live1) * steps / el_b
    log(f"[bench] cfg3 batched: {steps} steps in {el_b:.3f} s")

    singles = [ffi.Handle(p, s) for p, s in zip(params, scenes)]
    for g in singles:
        g.step(2)
    for g in singles:
        g.sync()
    live = 0
    t0 = time.perf_counter()
    for g in singles:
        g.step(seq_steps)
        g.sync()
    el_s = time.perf_counter() - t0
    for g in singles:
        live += int((g.state()["typ"] == ffi.FLD).sum())
        g.close()
    rate_s = live * seq_steps / el_s
    log(f"[bench] cfg3 sequential: {len(singles)} x {seq_steps} steps in {el_s:.3f} s")
    return {"workload": "cfg3: 512 independent 2-D dam breaks, FP32, l0 = 0.005, tank 1.0 x 0.6",
            "scenes": len(scenes), "fluid": int(sum(nf)), "steps": steps,
            "value": round(rate_b, 1), "unit": "particle-steps/s",
            "ms_per_step": round(1e3 * el_b / steps, 4),
            "sequential": {"value": round(rate_s, 1), "steps_per_scene": seq_steps,
                           "ms_per_scene_step": round(1e3 * el_s / (seq_steps * len(scenes)), 4),
                           "launch": "plain (eager) launches of single-scene handles, one scene after another"},
            "speedup": round(rate_b / rate_s, 2),
            "autotune": {k: tuned[k] for k in ("block", "wave_grid", "graph", "trials")}}


def adjoint_leg(steps: int, every: int, log) -> dict:
    """cfg4 timing case: cfg1 (300k fluid, FP32), `steps` steps with a
    checkpoint every `every` steps, forward + reverse sweep, against the
    same steps run plainly."""
    from mps_cuda_amd import configs, ffi

    scene, params = configs.cfg1("fp32")
    plain = ffi.Handle(params, scene)
    plain.step(2)
    plain.sync()
    plain.set_state(scene.typ, scene.pos, np.zeros((3, scene.n)), np.zeros(scene.n), np.zeros(scene.n))
    t0 = time.perf_counter()
    plain.step(steps)
    plain.sync()
    t_plain = time.perf_counter() - t0
    plain.close()

    h = ffi.Handle(params, scene)
    fl = scene.typ == ffi.FLD

    def sweep():
        t0 = time.perf_counter()
        h.adjoint_forward(steps, every)
        h.sync()
        t1 = time.perf_counter()
        st = h.state()  # loss gradient on the host (not timed)
        bar = np.zeros((3, h.n))
        bar[:, fl] = 2.0 * (st["pos"][:, fl] - scene.pos[:, fl]) / fl.sum()
        t2 = time.perf_counter()
        g = h.adjoint_backward(bar)
        t3 = time.perf_counter()
        return t1 - t0, t3 - t2, g

    sweep()  # first sweep allocates checkpoints and tapes
    h.set_state(scene.typ, scene.pos, np.zeros((3, scene.n)), np.zeros(scene.n), np.zeros(scene.n))
    t_f, t_b, g = sweep()
    h.close()
    log(f"[bench] adjoint: plain {t_plain:.3f} s, forward+checkpoints {t_f:.3f} s, reverse {t_b:.3f} s")
    return {"workload": f"cfg4 timing case: cfg1 3-D dam break, 300k fluid, FP32, {steps} steps, "
                        f"checkpoint every {every}",
            "ratio": round((t_f + t_b) / t_plain, 3), "unit": "adjoint time / forward time",
            "forward_s": round(t_plain, 4), "forward_checkpointed_s": round(t_f, 4),
            "reverse_s": round(t_b, 4),
            "grad": {"nu": g["nu"], "c": g["c"], "vel0_norm": float(np.linalg.norm(g["vel0"]))}}


def tensor_leg(steps: int, warmup: int, log) -> dict:
    """cfg1 with the tensor-corrected gradient (grad_model MPS_GRAD_TENSOR):
    step rate and the gradient kernel's average duration (HIP events)."""
    from mps_cuda_amd import configs, ffi

    scene, params = configs.cfg1("fp32", grad_model=ffi.GRAD_TENSOR)
    h = ffi.Handle(params, scene)
    h.step(warmup)
    h.sync()
    live0 = int((h.state()["typ"] == ffi.FLD).sum())
    t0 = time.perf_counter()
    h.step(steps)
    h.sync()
    elapsed = time.perf_counter() - t0
    live = 0.5 * (live0 + int((h.state()["typ"] == ffi.FLD).sum()))
    h.close()
    # event-timed replay of the same steps, as in the plain leg
    h = ffi.Handle(params, scene)
    h.timing(True)
    h.step(warmup)
    h.timing(True)  # resets the accumulators
    h.step(steps)
    kt = h.timing_read()
    h.close()
    g = kt["grad_up2"]
    log(f"[bench] tensor gradient: {steps} steps in {elapsed:.3f} s, grad_up2 {1e3 * g[0] / g[1]:.1f} us")
    return {"workload": "cfg1: 3-D dam break, 300k fluid, FP32, tensor-corrected gradient",
            "value": round(live * steps / elapsed, 1), "unit": "particle-steps/s",
            "ms_per_step": round(1e3 * elapsed / steps, 4), "grad_up2_ms": round(g[0] / g[1], 5)}


def plain_leg(args) -> dict:
    """cfg1: timed steps (event-free), then a deterministic replay with
    per-kernel HIP events for the roofline."""
    from mps_cuda_amd import configs, ffi

    scene, params = configs.cfg1("fp32")
    nf0 = int((scene.typ == ffi.FLD).sum())
    nw = int((scene.typ == ffi.WLL).sum())
    h = ffi.Handle(params, scene)
    print(f"[bench] cfg1: {nf0} fluid, {nw} walls; warmup {args.warmup}", file=sys.stderr, flush=True)
    h.step(args.warmup)
    h.sync()
    live0 = int((h.state()["typ"] == ffi.FLD).sum())
    t0 = time.perf_counter()
    h.step(args.steps)
    h.sync()
    elapsed = time.perf_counter() - t0
    live1 = int((h.state()["typ"] == ffi.FLD).sum())
    live = 0.5 * (live0 + live1)
    value = live * args.steps / elapsed
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    h.close()

    # Kernel durations over the same steps: the step is deterministic, so a
    # second handle replays the timed region exactly, this time with HIP
    # events carried in each launch's dispatch packet (hipExtLaunchKernelGGL)
    # on the handle's stream.  The warm-up fills the event pool.
    h = ffi.Handle(params, scene)
    h.timing(True)
    h.step(args.warmup)
    h.timing(True)  # resets the accumulators
    h.step(args.steps)
    kt = h.timing_read()
    h.timing(False)
    m = len(h.mirrors()[0])
    # the kernel with the longest launch (the mirror generator's two launches
    # add up to more GPU time, but one of them runs beside other kernels)
    dom = max((k for k in kt if algorithmic_bytes(k, 1, 1, 1) > 0 and kt[k][1] > 0),
              key=lambda k: kt[k][0] / kt[k][1])
    avg_ms = kt[dom][0] / kt[dom][1]
    nlive = int((h.state()["typ"] == ffi.FLD).sum())
    bytes_ = algorithmic_bytes(dom, nlive, nw, m)
    achieved = bytes_ / (avg_ms * 1e-3) / 1e9
    passes = {"vsc_up1": 1, "prs": 1, "grad_up2": 2, "col": 1, "col_prs": 1}.get(dom, 1)
    flops = FLOP_PER_CANDIDATE * candidates_per_pass(h) * passes
    tflops = flops / (avg_ms * 1e-3) / 1e12
    traffic = measured_traffic(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None,
                "bytes_per_launch": bytes_, "avg_ms": round(avg_ms, 5),
                "valu": {"achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "flops_per_launch": flops}}
    return dict(value=value, elapsed=elapsed, nf0=nf0, nw=nw, live=live, m=m, roofline=roofline, kt=kt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the cfg1 leg (profiling other legs)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--seq-steps", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-tensor", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=25)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    plain = None if args.no_plain else plain_leg(args)

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_seconds)

    batch = None
    if not args.no_batch:
        batch = batched_leg(args.batch_steps, args.seq_steps,
                            lambda msg: print(msg, file=sys.stderr, flush=True))

    adj = None
    if not args.no_adjoint:
        adj = adjoint_leg(args.adjoint_steps, args.adjoint_every,
                          lambda msg: print(msg, file=sys.stderr, flush=True))

    tensor = None
    if not args.no_tensor:
        tensor = tensor_leg(args.steps, args.warmup, lambda msg: print(msg, file=sys.stderr, flush=True))

    P = plain or dict(value=None, elapsed=None, nf0=None, nw=None, live=None, m=None, roofline=None, kt={})
    out = {
        "metric": METRIC, "value": None if plain is None else round(P["value"], 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None if plain is None else round(1e3 * P["elapsed"] / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cfg1: 3-D dam break, 300k fluid, FP32", "fluid": P["nf0"], "walls": P["nw"],
                   "live_fluid": P["live"], "mirrors": P["m"], "l0": 0.005, "dt": 2.5e-4,
                   "tank": [1.0, 0.6, 0.3], "column": [0.25, 0.5, 0.3]},
        "roofline": P["roofline"],
        "cpu_baseline": cpu,
        "batched": batch,
        "adjoint": adj,
        "tensor_gradient": tensor,
        "kernels_ms": {k: round(v[0] / v[1], 5) for k, v in sorted(P["kt"].items(), key=lambda kv: -kv[1][0])
                       if v[1] > 0},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
