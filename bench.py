#!/usr/bin/env python3
"""Benchmark of the photon-transport hot path on one MI355X.

Workload: the 5-layer slab (cfg1): h 0.3/0.5/0.2/0.3 cm + semi-infinite, 1e8 photons,
replay with per-photon pathlength dump (-R -P -B).  One step = one full hot-path pass on
device: persistent first-pass transport of all photons, replay of every detected photon,
and the pathlength dump; inputs are resident and nothing is written to disk.

value = photons / second of wall time over the timed steps (whole job, one GPU).
roofline = the first-pass transport kernel, FP32 VALU bound: algorithmic FLOP per photon
(62 per interaction step + 44 per boundary step, counted on the same run) divided by the
kernel's average duration from HIP events on the library's stream.
cpu_baseline = the CPU restatement of the same loop (oracle, OpenMP) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

FLOP_INTERACTION = 62  # step 2, boundary test 2, move 6, drop 3, HG 11, sin 3, psi 3, sqrt 3, rotate 20, renorm 9
FLOP_BOUNDARY = 44     # step, clip, move, Fresnel + Snell on transmission
PEAK_FP32_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--config", default="cfg1")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-photons", type=float, default=0, help="CPU baseline sample size")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--bpc", type=int, default=0)
    ap.add_argument("--bnd", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=float, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter scale)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        nbytes = int(args.calibrate_copy)
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "read_bytes": nbytes, "write_bytes": nbytes}))
        return

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(photon ranges would shard across GPUs with one final reduction).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import mcml_gpu_amd as M

    if args.config == "cfg3":
        return bench_batch(args)
    sim = M.load_config(args.config)
    sim.number_photons = int(args.photons)
    opts = M.make_opts(replay=True, pathlength=True, binary=True, block_size=args.block,
                       chunk=args.chunk, refill_min=args.refill, blocks_per_cu=args.bpc,
                       boundary_min=args.bnd)
    ctx = M.Context(0)
    for _ in range(args.warmup):
        ctx.run(sim, opts, seed=args.seed)
    kern_ms, rep_ms, dump_ms, walls = [], [], [], []
    res = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        res = ctx.run(sim, opts, seed=args.seed + i)
        walls.append(time.perf_counter() - ts)
        kern_ms.append(res.ms_first_pass_kernel)
        rep_ms.append(res.ms_replay_kernel)
        dump_ms.append(res.ms_dump_kernel)
    elapsed = time.perf_counter() - t0
    n = sim.number_photons
    value = n * args.steps / elapsed
    k_ms = float(np.mean(kern_ms))
    flop_per_photon = (FLOP_INTERACTION * res.steps_interaction + FLOP_BOUNDARY * res.steps_boundary) / n
    achieved = flop_per_photon * n / (k_ms * 1e-3) / 1e12
    detected = int(res.detected.sum())
    dump_bytes = 2 * detected * (sim.number_layers + 2) * 4  # SoA read + record write

    line = {
        "metric": "photons/sec/GPU (plain MC)",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s: 5-layer slab (h 0.3/0.5/0.2/0.3 cm + semi-inf, g 0.9, n 1.4), "
                        "%.0e photons, replay + pathlength dump (-R -P -B)" % (args.config, n),
            "photons": n,
            "layers": sim.number_layers,
            "sds": sim.num_sds,
            "block": res.block_size, "grid": res.grid_size, "chunk": res.chunk,
            "refill_min": res.refill_min, "boundary_min": res.boundary_min,
        },
        "kernel_photons_per_s": n / (k_ms * 1e-3),
        "first_pass_kernel_ms": k_ms,
        "replay_kernel_ms": float(np.mean(rep_ms)),
        "dump_kernel_ms": float(np.mean(dump_ms)),
        "steps_per_photon": res.steps / n,
        "interactions_per_photon": res.steps_interaction / n,
        "boundaries_per_photon": res.steps_boundary / n,
        "detected": detected,
        "roofline": {
            "bound": "valu",
            "kernel": "mcml_first_pass",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "flop_per_photon": flop_per_photon,
            "traffic": None,
        },
        "dump": {"bytes": dump_bytes,
                 "GB_per_s": (dump_bytes / (np.mean(dump_ms) * 1e-3) / 1e9) if np.mean(dump_ms) > 0 else None},
    }

    if not args.no_cpu:
        import oracle_ffi as O
        cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16)
        sample = int(args.cpu_photons) if args.cpu_photons else 150000 * cores
        d = O.derive(sim)
        tc = time.perf_counter()
        O.first_pass(d, args.seed, 0, sample, nthreads=cores)
        dt = time.perf_counter() - tc
        line["cpu_baseline"] = {
            "value": sample / dt, "unit": "photons/s", "cores": cores, "kind": "port",
            "sample": "%d photons of the same config, first pass (oracle, OpenMP), %.1f s" % (sample, dt),
        }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_batch(args):
    """cfg3: 512 three-layer sets with log-uniform budgets (about 1.1e9 photons) in one
    persistent launch; launch shape picked by mcml_autotune."""
    import numpy as np
    import mcml_gpu_amd as M
    from mcml_gpu_amd import batch

    sims, params, budgets = batch.cfg3_sets(512, seed=1)
    total = int(budgets.sum())
    ctx = M.Context(0)
    tuned = ctx.autotune(sims, trial_photons=2e7)
    opts = M.make_opts(block_size=tuned.block_size, chunk=tuned.chunk, refill_min=tuned.refill_min,
                       use_graph=True, skip_records=True)
    for _ in range(args.warmup):
        ctx.run_batch(sims, opts, seed=args.seed)
    kern, walls_fp = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = ctx.run_batch(sims, opts, seed=args.seed + i)
        kern.append(res.ms_first_pass_kernel)
        walls_fp.append(res.ms_first_pass_wall)
    elapsed = time.perf_counter() - t0
    k_ms = float(np.mean(kern))
    flop = (FLOP_INTERACTION * res.steps_interaction + FLOP_BOUNDARY * res.steps_boundary)
    achieved = flop / (k_ms * 1e-3) / 1e12
    line = {
        "metric": "photons/sec/GPU (plain MC, batched)", "value": total * args.steps / elapsed,
        "unit": "photons/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cfg3: 512 three-layer sets, budgets log-uniform 1e5-1e7, one persistent launch",
                   "photons": total, "sets": len(sims), "block": res.block_size, "chunk": res.chunk,
                   "refill_min": res.refill_min, "graph": True},
        "kernel_photons_per_s": total / (k_ms * 1e-3), "first_pass_kernel_ms": k_ms,
        "first_pass_wall_ms": float(np.mean(walls_fp)), "launches": res.first_pass_launches,
        "detected": int(res.detected.sum()), "steps_per_photon": res.steps / total,
        "roofline": {"bound": "valu", "kernel": "mcml_first_pass", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": None},
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
