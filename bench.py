#!/usr/bin/env python3
"""Benchmark of the photon-transport hot path on one MI355X.

Workload: the 5-layer slab (cfg1): h 0.3/0.5/0.2/0.3 cm + semi-infinite, 1e8 photons,
replay with per-photon pathlength dump (-R -P -B).  --config cfg0/cfg2 or the reference's
ring-detection examples (head_model, hw8a) run the same measurement on those inputs.  One step = one full hot-path pass on
device: persistent first-pass transport of all photons, replay of every detected photon,
and the pathlength dump; inputs are resident and nothing is written to disk.

value = photons / second of wall time over the timed steps (whole job, one GPU).
roofline = the first-pass transport kernel, FP32 VALU bound: algorithmic FLOP per photon
(62 per interaction step + 44 per boundary step, with the per-photon step counts the oracle
counts at seed 1 over 1e5 photons, committed in tests/golden/algorithmic_counts.json; the
run's own counters for a configuration not listed there, and as `flop_per_photon_live`)
times the photons, divided by the kernel's average duration from HIP events on the
library's stream.
cpu_baseline = the CPU restatement of the same loop (oracle, OpenMP) on a bounded sample.
reference_shape = one run of the same photon step in the reference's launch structure
(32-thread workgroups, one photon per thread, 50 000-step relaunches with per-launch host
read-back and reduction, then a separate replay), timed the same way, as the stand-in for
the reference's own photons/s; `speedup` = value / its photons/s.

--config cfg3: the 512-set batch in one persistent launch (photons/s).
--config cfg4: the reflectance surrogate (R evaluations/s of the trained network at its
measured error against high-photon plain MC).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons N] [--config cfgN] [--no-cpu]
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
# per-photon interaction and boundary counts of each configuration, counted by the oracle at
# seed 1 over photons [0, 1e5) (tests/golden/make_algorithmic_counts.py)
COUNTS_FILE = os.path.join(ROOT, "tests", "golden", "algorithmic_counts.json")
PEAK_FP32_TFLOPS = 157.3
EXAMPLES = {"head_model": "tests/golden/head_model", "hw8a": "tests/golden/HW8_A"}
# HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by what the
# counter reads for a plain 1 GiB device copy (profiles/transport_cfg1_1e8.md,
# profiles/surrogate_fused_cfg4.md): transport 5.00e7 B read + 9.10e7 B written per 1e8
# photons over the pass's six launches (ticket atomics, counters, and the photon states the
# main launch hands to the compacted drain launches); fused inference 8.55e6 + 1.10e6 B per
# 1e5 rows.
TRANSPORT_TRAFFIC_PER_PHOTON = (4.997e7 + 9.103e7) / 1e8
MLP_TRAFFIC_PER_SAMPLE = (8.551e6 + 1.100e6) / 1e5
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
    ap.add_argument("--no-refshape", action="store_true", help="skip the reference-launch-shape run")
    ap.add_argument("--refshape-block", type=int, default=32,
                    help="workgroup size of the reference-launch-shape run (the reference uses 32)")
    ap.add_argument("--cpu-photons", type=float, default=0, help="CPU baseline sample size")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--bpc", type=int, default=0)
    ap.add_argument("--bnd", type=int, default=0)
    ap.add_argument("--spill-below", type=int, default=0,
                    help="drain compaction threshold (0 = library default, -1 = off)")
    ap.add_argument("--separate-replay", action="store_true",
                    help="-A: replay in a pass of its own after the first pass instead of within it")
    ap.add_argument("--pathlength", action="store_true",
                    help="cfg3: also record per-layer pathlengths and dump the records (-R -P)")
    ap.add_argument("--absorbance", action="store_true",
                    help="replay into the A_rz/A0_z grids (-R -A -B) on any single-set config (cfg2 always does)")
    ap.add_argument("--train-sets", type=int, default=16384, help="cfg4: simulated training sets")
    ap.add_argument("--train-photons", type=float, default=1e6, help="cfg4: photons per training set")
    ap.add_argument("--white", type=int, default=13, help="cfg4: absorption vectors per set (White MC)")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--heldout-sets", type=int, default=1000)
    ap.add_argument("--heldout-photons", type=float, default=1e8)
    ap.add_argument("--infer-batch", type=int, default=100000)
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
    if args.config == "cfg4":
        return bench_surrogate(args)
    if args.config in EXAMPLES:  # the reference's own example inputs (ring detection)
        d = os.path.join(ROOT, EXAMPLES[args.config])
        sim = M.read_input(os.path.join(d, "sim_setup.json"), os.path.join(d, "GPUMC_input.txt"))[0]
    else:
        sim = M.load_config(args.config)
    sim.number_photons = int(args.photons)
    # cfg2 is the absorbance-map workload (-R -A -B: the first pass, then a replay of every
    # detected photon into the A_rz/A0_z grids); the others dump pathlengths (-R -P -B)
    absorb = args.config == "cfg2" or args.absorbance
    opts = M.make_opts(replay=True, pathlength=not absorb, absorbance=absorb, binary=True, block_size=args.block,
                       chunk=args.chunk, refill_min=args.refill, blocks_per_cu=args.bpc,
                       boundary_min=args.bnd, spill_below=args.spill_below, separate_replay=args.separate_replay)
    flags = "-R -A -B" if absorb else "-R -P -B"
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
    flop_live = (FLOP_INTERACTION * res.steps_interaction + FLOP_BOUNDARY * res.steps_boundary) / n
    flop_per_photon, flop_source = flop_live, "this run's step counters"
    with open(COUNTS_FILE) as f:
        counts = json.load(f)
    if args.config in counts["configs"]:
        c = counts["configs"][args.config]
        flop_per_photon = (FLOP_INTERACTION * c["interaction"] + FLOP_BOUNDARY * c["boundary"]) / counts["photons"]
        flop_source = "oracle, seed %d, %d photons" % (counts["seed"], counts["photons"])
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
            "workload": "%s: %d-layer slab, %d SDS (%s detection), %.0e photons, %s" % (
                args.config, sim.number_layers, sim.num_sds, "fiber" if sim.detector_probe_oblique else "ring", n,
                "replay into %d x %d absorbance grids (%s)" % (res.A0_z.shape[-1], res.A_rz.shape[-1], flags) if absorb
                else "replay + pathlength dump (%s)" % flags),
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
            "kernel": "mcml_first_pass" if absorb else "mcml_first_pass_pl",  # -R -P: inline pathlengths
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "flop_per_photon": flop_per_photon,
            "flop_source": flop_source,
            "flop_per_photon_live": flop_live,
            "traffic": TRANSPORT_TRAFFIC_PER_PHOTON * n,
        },
        "dump": {"bytes": dump_bytes,
                 "GB_per_s": (dump_bytes / (np.mean(dump_ms) * 1e-3) / 1e9) if np.mean(dump_ms) > 0 else None},
    }

    # where the lanes' time goes: one more run of the first pass's statistics variant (not
    # timed with the steps above): lane-blocks by state and the drain after the queue ran dry
    rs = ctx.run(sim, M.make_opts(replay=True, pathlength=not absorb, absorbance=absorb, binary=True,
                                  block_size=args.block, chunk=args.chunk, refill_min=args.refill,
                                  blocks_per_cu=args.bpc, boundary_min=args.bnd, spill_below=args.spill_below,
                                  lane_stats=True), seed=args.seed)
    lb = rs.lane_blocks.astype(np.float64)
    line["lanes"] = {
        "lane_blocks": [int(v) for v in rs.lane_blocks],
        "frac_stepping": lb[0] / lb.sum(), "frac_boundary_wait": lb[1] / lb.sum(),
        "frac_refill_wait": lb[2] / lb.sum(), "frac_dry": lb[3] / lb.sum(),
        "steps_per_stepping_lane_block": rs.steps / lb[0] if lb[0] else None,
        "ms_drain": rs.ms_drain, "ms_clock": rs.ms_clock, "kernel_ms": rs.ms_first_pass_kernel,
    }

    if not args.no_refshape:
        # the same photon step in the reference's launch structure (one photon per thread,
        # 32-thread workgroups, 50 000-step relaunches, per-launch host read-back and
        # reduction of per-thread detection buffers), then the replay as a separate pass:
        # the stated stand-in for the reference's own photons/s, which cannot run here
        ropts = M.make_opts(replay=True, pathlength=not absorb, absorbance=absorb, binary=True,
                            reference_shape=True, block_size=args.refshape_block)
        ts = time.perf_counter()
        rr = ctx.run(sim, ropts, seed=args.seed)
        rwall = time.perf_counter() - ts
        line["reference_shape"] = {
            "value": n / rwall, "unit": "photons/s",
            "kernel_photons_per_s": n / (rr.ms_first_pass_kernel * 1e-3),
            "first_pass_kernel_ms": rr.ms_first_pass_kernel,
            "first_pass_wall_ms": rr.ms_first_pass_wall,
            "replay_kernel_ms": rr.ms_replay_kernel,
            "wall_ms": rwall * 1e3, "launches": rr.first_pass_launches,
            "block": rr.block_size, "threads": rr.block_size * rr.grid_size,
            "speedup": value / (n / rwall),
            "kernel_speedup": rr.ms_first_pass_kernel / k_ms,
            "detected": int(rr.detected.sum()),
        }

    if not args.no_cpu:
        import oracle_ffi as O
        # the host share this process may use: its CPU affinity, and OMP_NUM_THREADS where set
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cores = min(usable, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or usable)
        sample = int(args.cpu_photons) if args.cpu_photons else 150000 * cores
        d = O.derive(sim)
        tc = time.perf_counter()
        O.first_pass(d, args.seed, 0, sample, nthreads=cores)
        dt = time.perf_counter() - tc
        line["cpu_baseline"] = {
            "value": sample / dt, "unit": "photons/s", "cores": cores, "cores_total": os.cpu_count(),
            "cores_usable": usable, "kind": "port",
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
                       boundary_min=tuned.boundary_min,
                       use_graph=True, skip_records=not args.pathlength, replay=args.pathlength,
                       pathlength=args.pathlength, binary=args.pathlength)
    for _ in range(args.warmup):
        ctx.run_batch(sims, opts, seed=args.seed)
    kern, walls_fp, dumps = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = ctx.run_batch(sims, opts, seed=args.seed + i)
        kern.append(res.ms_first_pass_kernel)
        walls_fp.append(res.ms_first_pass_wall)
        dumps.append(res.ms_dump_kernel)
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
                   "refill_min": res.refill_min, "boundary_min": res.boundary_min, "graph": True},
        "kernel_photons_per_s": total / (k_ms * 1e-3), "first_pass_kernel_ms": k_ms,
        "first_pass_wall_ms": float(np.mean(walls_fp)), "launches": res.first_pass_launches,
        "detected": int(res.detected.sum()), "steps_per_photon": res.steps / total,
        "roofline": {"bound": "valu", "kernel": "mcml_first_pass_pl" if args.pathlength else "mcml_first_pass",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "flop_per_photon": flop / total,
                     "flop_source": "this run's step counters (512 sets)",
                     "traffic": TRANSPORT_TRAFFIC_PER_PHOTON * total},
    }
    if args.pathlength:
        n_rec = int(res.detected.sum())
        stride = sims[0].number_layers + 2
        # per record: sorted key 8 + source index 4 + source row read + row write + weight 4 + public record 16
        dump_bytes = n_rec * (8 + 4 + 2 * 4 * stride + 4 + 16)
        d_ms = float(np.mean(dumps))
        line["config"]["workload"] += ", -R -P: per-layer pathlength records dumped in file order"
        line["dump"] = {"kernel": "mcml_dump", "records": n_rec, "bytes": dump_bytes, "kernel_ms": d_ms,
                        "GB_per_s": dump_bytes / (d_ms * 1e-3) / 1e9 if d_ms > 0 else None,
                        "frac_of_hbm_peak": dump_bytes / (d_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if d_ms > 0 else None}
    print(json.dumps(line), flush=True)
    ctx.close()


MLP_FLOP_PER_SAMPLE = 2 * (14 * 256 + 3 * 256 * 256 + 256 * 3)  # forward multiply-adds x 2
PEAK_BF16_TFLOPS = 2500.0


def bench_surrogate(args):
    """cfg4: train the 14 -> 4x256 GELU -> 3 network on White-MC-expanded batched MC, evaluate it
    on held-out sets simulated with plain MC, and time batched inference."""
    import numpy as np
    import mcml_gpu_amd as M
    from mcml_gpu_amd import batch
    from mcml_gpu_amd import surrogate as SG

    ctx = M.Context(0)
    t0 = time.perf_counter()
    params = batch.draw_parameters(args.train_sets, seed=1)
    xs, Rs, rvs = [], [], []
    per = 1024
    for i in range(0, args.train_sets, per):
        p = params[i:i + per]
        x, R, rv = SG.white_dataset(ctx, p, np.full(len(p), int(args.train_photons)), args.white, seed=1 + i)
        xs.append(x), Rs.append(R), rvs.append(rv)
    x, R, rv = np.concatenate(xs), np.concatenate(Rs), np.concatenate(rvs)
    t_data = time.perf_counter() - t0
    print("training samples %d from %d sets x %.3g photons (%.1f s)" % (len(x), args.train_sets, args.train_photons,
                                                                       t_data), flush=True)
    net, norm, hist = SG.fit(x, R, rv, epochs=args.epochs, batch_size=1024, lr=2e-3, seed=0)
    print("trained %d epochs in %.1f s (last-batch loss %.3g)" % (args.epochs, net.train_ms / 1e3, hist[-1]),
          flush=True)

    t1 = time.perf_counter()
    hp, hR, hrv = SG.reference_set(
        ctx, args.heldout_sets, args.heldout_photons, seed=9001,
        progress=lambda i, n: print("held-out MC %d/%d sets (%.0f s)" % (i, n, time.perf_counter() - t1), flush=True))
    t_ref = time.perf_counter() - t1
    # plain-MC throughput on these sets (first pass only, no replay), for the MC cost per evaluation
    n_pl = min(20, args.heldout_sets)
    _, res_pl, _ = SG.simulate(ctx, hp[:n_pl], np.full(n_pl, int(1e7)), seed=77, replay=False)
    tp = time.perf_counter()
    SG.simulate(ctx, hp[:n_pl], np.full(n_pl, int(1e7)), seed=78, replay=False)
    mc_pps = n_pl * 1e7 / (time.perf_counter() - tp)
    sur = SG.Surrogate(net, norm, bound=0.02)
    med = sur.validate(hp, hR)
    err = SG.relative_error(SG.predict_R(net, norm, hp), hR)

    q = SG.features(batch.draw_parameters(args.infer_batch, seed=4242))
    for _ in range(max(1, args.warmup)):
        net.predict(q)
    kms = []
    for _ in range(args.steps):
        net.predict(q)
        kms.append(net.last_ms)
    k_ms = float(np.mean(kms))
    evals_per_s = args.infer_batch / (k_ms * 1e-3)
    achieved = MLP_FLOP_PER_SAMPLE * args.infer_batch / (k_ms * 1e-3) / 1e12
    # plain MC cost of one evaluation at 2% relative standard error at every SDS
    photons_2pct = float(np.median(SG.photons_for_error(hrv * args.heldout_photons, 0.02)))
    mc_evals_per_s = mc_pps / photons_2pct
    line = {
        "metric": "surrogate R-evals/sec at <=2% median rel err vs 1e8-photon MC", "value": evals_per_s,
        "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": k_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (simulated by this library)",
        "config": {"workload": "cfg4: MLP 14-4x256 GELU-3, inference batch %d; trained on %d samples (%d sets "
                               "x %.3g photons, White MC x%d); evaluated on %d held-out sets x %.3g photons"
                               % (args.infer_batch, len(x), args.train_sets, args.train_photons, args.white,
                                  args.heldout_sets, args.heldout_photons),
                   "epochs": args.epochs},
        "median_rel_err": med, "median_rel_err_per_sds": [float(v) for v in np.median(err, axis=0)],
        "within_bound_per_sds": bool(np.all(np.median(err, axis=0) <= 0.02)),
        "reference_rel_se_per_sds": [float(v) for v in np.median(np.sqrt(hrv), axis=0)],
        "p90_rel_err": float(np.quantile(err, 0.9)), "within_bound": bool(sur.usable),
        "mc_photons_per_s": mc_pps, "mc_photons_for_2pct": photons_2pct, "mc_evals_per_s": mc_evals_per_s,
        "gain_vs_mc": evals_per_s / mc_evals_per_s, "train_s": net.train_ms / 1e3, "data_s": t_data,
        "heldout_s": t_ref,
        "roofline": {"bound": "mfma", "kernel": "mcml_mlp_fused", "achieved": achieved,
                     "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS,
                     "flop_per_sample": MLP_FLOP_PER_SAMPLE, "traffic": MLP_TRAFFIC_PER_SAMPLE * args.infer_batch},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import mlp_oracle as MO
        p = net.get_params()
        n_cpu = 20000
        tc = time.perf_counter()
        reps = 0
        while time.perf_counter() - tc < 10.0:
            MO.forward(p, SG.WIDTHS, q[:n_cpu], bf16=False)
            reps += 1
        dt = time.perf_counter() - tc
        line["cpu_baseline"] = {"value": reps * n_cpu / dt, "unit": "evals/s",
                                "cores": int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1),
                                "cores_total": os.cpu_count(),
                                "kind": "port", "sample": "%d x %d rows, numpy float64 forward" % (reps, n_cpu)}
    print(json.dumps(line), flush=True)
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
