"""Benchmark: 2-D radial x angular real-time split-operator propagation on one MI355X.

Workload (the headline configuration): r in [0.5, 50] on 2048 points, 128
Gauss-Legendre theta nodes, Legendre transform to l = 0..127, V(r, theta) =
C12/r^12 - C6/r^6 (1 + a P2(cos theta)) with seeded (C6, C12, a), m = 1,
Gaussian r0 = 8, sigma = 1 times P0, dt = 0.002, border damping on the last
10% of r with its loss checked each step. Complex FP32 storage, FP64 norms.

A step is one Propagation::step of that stack. The timed region is K steps
with psi resident in HBM, bracketed by device synchronisation. Per-kernel
durations come from HIP events on the library's stream (each pass of the
steady-state step launched on its own, back to back). The CPU baseline is
the f64 C oracle on the host cores for a bounded number of steps.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-point-steps/s; time-to-ground-state at 1e-8 E, learned vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
N_R, N_THETA, DT = 2048, 128, 0.002


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=1000, help="steps of the CPU oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=200)
    ap.add_argument("--c3-steps", type=int, default=100, help="steps of the 256^3 Cartesian workload (0: skip)")
    ap.add_argument("--c4-instances", type=int, default=64, help="batched ground-state instances (0: skip)")
    ap.add_argument("--c4-max-steps", type=int, default=40000)
    ap.add_argument("--c4-f32-steps", type=int, default=2000, help="steps of the FP32 batched throughput run")
    ap.add_argument("--c4-solo", action="store_true", help="also relax the instances one by one")
    ap.add_argument("--c5-train", type=int, default=1024, help="training potentials of the warm start (0: skip)")
    ap.add_argument("--c5-cold-steps", type=int, default=28672, help="steps of the cold-started training solutions")
    ap.add_argument("--c5-warm-steps", type=int, default=10240, help="steps of the warm-started training solutions")
    ap.add_argument("--c5-eval", type=int, default=64, help="held-out potentials of the warm-start evaluation")
    ap.add_argument("--c5-updates", type=int, default=4000, help="optimiser steps per fit")
    ap.add_argument("--no-autotune", action="store_true", help="tile sizes from the fixed rule instead of timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-MiB buffer 3 times (a plain copy kernel of known traffic for "
                         "scaling hardware byte counters), then run the benchmark")
    return ap.parse_args()


def pass_record(t):
    pts = t["points"]
    cb = t["complex_bytes"]
    # algorithmic bytes: psi read + written once, each full-grid operator table read once
    bytes_ = (2 + t["nd_tables"]) * pts * cb
    sec = t["avg_ms"] * 1e-3
    kind = {0: "elementwise", 1: "fft", 2: "legendre_gemm"}[t["xform_type"]]
    rec = {
        "kernel": kind, "transforms": t["nx"], "axis": t["axis"], "avg_us": t["avg_ms"] * 1e3,
        "workgroups": t["tiles"], "alg_bytes": bytes_, "gbs": bytes_ / sec / 1e9,
        "mat_flops": t["mat_flops"], "fft_flops": t["fft_flops"],
    }
    if t["mat_flops"] > 0:
        rec["tflops"] = t["mat_flops"] / sec / 1e12
    return rec


def cartesian_256(so, W, steps, kernel_iters, autotune=True):
    """The 256^3 Cartesian workload (complex FP32): throughput and per-pass rates."""
    import time

    p, _ = W.cartesian_3d(so, n=256, steps=steps, dtype="f32")
    if autotune:
        p.set_autotune(True)
    p.steps(5)
    so.synchronize()
    t0 = time.perf_counter()
    p.steps(steps)
    so.synchronize()
    wall = time.perf_counter() - t0
    pts = 256 ** 3
    passes = [pass_record(t) for t in p.time_passes(max(10, kernel_iters // 10))]
    step_bytes = sum(r["alg_bytes"] for r in passes)
    del p
    # accuracy on a 64^3 subsample run of the same physics (dt 0.005, 500 steps): FP32 and
    # FP32-storage/FP64-arithmetic against FP64 storage, and the norm drift of each
    import numpy as np

    res = {}
    for dt_ in ("f32", "f32x", "f64"):
        q, _ = W.cartesian_3d(so, n=64, steps=500, dtype=dt_)
        q.steps(500)
        res[dt_] = (q.wave_function().array, q.wave_function().norm())
        del q
    ref = res["f64"][0]
    acc = {k: {"rel_l2_vs_f64": float(np.linalg.norm(res[k][0] - ref) / np.linalg.norm(ref)),
               "norm_drift": float(res[k][1] - 1.0)} for k in res}
    gbs = step_bytes / (wall / steps) / 1e9
    # the step as a whole against HBM: every pass is an HBM-bound lane sweep, and the measured traffic per
    # launch equals the algorithmic bytes (counter summary committed with the repository)
    return {"value": pts * steps / wall, "unit": "grid-point-steps/s", "ms_per_step": wall / steps * 1e3,
            "steps": steps, "alg_bytes_per_step": step_bytes, "step_gbs": gbs,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                         "scope": "whole step (4 FFT passes)"},
            "passes": passes, "subsample_64_500_steps": acc}


def ground_states(so, W, count, max_steps, f32_steps, solo):
    """Batched imaginary-time ground states: `count` seeded radial x angular instances with uneven
    grids (N_r in {512, 768, 1024, 1536, 2048}, N_theta in {32, 64, 128}), each with its own
    (C6, C12, a), relaxed with FP64 storage and FP64 renormalisation every step until the relative
    change of the decay-estimator energy per step is below 1e-10 (imaginary dt = 0.05)."""
    dt, tol = 0.05, 1e-10

    def relax(batched):
        gs = W.GroundStateBatch(so, count=count, seed0=4000, dt=dt, max_steps=max_steps, tol=tol, dtype="f64",
                                batched=batched)
        so.synchronize()
        t0 = time.perf_counter()
        last = [0]

        def progress():
            if gs.steps_done - last[0] >= 4096:
                last[0] = gs.steps_done
                print(f"ground states: {gs.steps_done} steps, {len(gs.active)} active", file=sys.stderr, flush=True)

        while gs.active and gs.steps_done < gs.max_steps:
            gs.advance(min(256, gs.max_steps - gs.steps_done))
            gs.check()
            progress()
        so.synchronize()
        wall = time.perf_counter() - t0
        pt_steps = sum(it["n_r"] * it["n_theta"] * it["steps"] for it in gs.inst)
        out = {"time_to_ground_state_s": wall, "converged": sum(it["converged"] for it in gs.inst),
               "steps_max": max(it["steps"] for it in gs.inst), "steps_min": min(it["steps"] for it in gs.inst),
               "pt_steps_per_s": pt_steps / wall}
        return out, gs

    rec, gs = relax(True)
    rec.update({"instances": count, "dt": dt, "tol": tol, "storage": "complex f64",
                "energy_estimator": "decay: E = -ln(1 - LeakControl loss of the step) / (2 dt)",
                "points": gs.points(active_only=False),
                "energies": [it["energy"] for it in gs.inst],
                "sizes": [[it["n_r"], it["n_theta"]] for it in gs.inst]})
    energies64 = [it["energy"] for it in gs.inst]
    del gs
    if solo:
        srec, sgs = relax(False)
        rec["solo"] = srec
        rec["solo_max_energy_diff"] = max(abs(a["energy"] - b) for a, b in zip(sgs.inst, energies64))
        del sgs
    # FP32 storage: batched throughput over a fixed number of steps, and how far its decay-estimator
    # energies land from the same steps stored in FP64
    g32 = W.GroundStateBatch(so, count=count, seed0=4000, dt=dt, max_steps=f32_steps, tol=0.0, dtype="f32")
    g32.advance(16)
    so.synchronize()
    t0 = time.perf_counter()
    g32.advance(f32_steps - 16)
    so.synchronize()
    wall = time.perf_counter() - t0
    g32.check()
    e32 = [it["energy"] for it in g32.inst]
    del g32
    g64 = W.GroundStateBatch(so, count=count, seed0=4000, dt=dt, max_steps=f32_steps, tol=0.0, dtype="f64")
    g64.advance(f32_steps)
    g64.check()
    import numpy as np

    e64 = [it["energy"] for it in g64.inst]
    diffs = [abs(a - e) for a, e in zip(e32, e64)]
    rels = [d / abs(e) for d, e in zip(diffs, e64)]
    del g64
    rec["f32"] = {"steps": f32_steps - 16, "pt_steps_per_s": rec["points"] * (f32_steps - 16) / wall,
                  "ms_per_step": wall / (f32_steps - 16) * 1e3,
                  "max_abs_energy_diff_vs_f64_same_steps": max(diffs),
                  "median_rel_energy_diff_vs_f64_same_steps": float(np.median(rels))}
    # FP32 with the Legendre products as split-BF16 x3 on BF16 MFMA
    gb = W.GroundStateBatch(so, count=count, seed0=4000, dt=dt, max_steps=f32_steps, tol=0.0, dtype="f32",
                            gemm="bf16x3")
    gb.advance(16)
    so.synchronize()
    t0 = time.perf_counter()
    gb.advance(f32_steps - 16)
    so.synchronize()
    wall = time.perf_counter() - t0
    gb.check()
    diffs = [abs(it["energy"] - e) for it, e in zip(gb.inst, e64)]
    rels = [d / abs(e) for d, e in zip(diffs, e64)]
    del gb
    rec["f32_split_bf16x3"] = {"steps": f32_steps - 16, "pt_steps_per_s": rec["points"] * (f32_steps - 16) / wall,
                               "ms_per_step": wall / (f32_steps - 16) * 1e3,
                               "max_abs_energy_diff_vs_f64_same_steps": max(diffs),
                               "median_rel_energy_diff_vs_f64_same_steps": float(np.median(rels))}
    return rec


def learned_warm_start(so, args):
    """Warm start of imaginary time from the network vs the plain cold start, on 1024 x 32 grids
    (r in [0.5, 50], 32 Gauss-Legendre theta nodes / l = 0..31), dt = 0.05, FP64 propagation with the
    decay-estimator energy. The reference energy of each held-out potential comes from an untimed cold
    run of 49152 steps; both timed runs retire an instance once E - E_ref <= 1e-8 |E_ref| (imaginary time
    lowers the energy monotonically)."""
    import numpy as np

    from split_operator_amd import warm_start as WS

    def log(msg):
        print("warm start: " + msg, file=sys.stderr, flush=True)

    ws = WS.LearnedWarmStart(n_r=1024, n_theta=32, dt=0.05, hidden=(256, 256, 256), seed=5)
    rng = np.random.default_rng(5000)
    lo, hi = WS.PARAM_LO, WS.PARAM_HI
    ptrain = lo + (hi - lo) * rng.uniform(size=(args.c5_train, 3))
    peval = lo + (hi - lo) * np.random.default_rng(6000).uniform(size=(args.c5_eval, 3))
    t0 = time.perf_counter()
    targets = ws.train(ptrain, args.c5_cold_steps, args.c5_warm_steps, first=256, updates=args.c5_updates, log=log)
    so.synchronize()
    t_train = time.perf_counter() - t0
    losses = ws.losses
    del targets
    log("reference run")
    ref, _ = ws.relax(peval, 49152)
    e_ref = [it["hist"][-1][2] for it in ref]
    del ref
    log("cold run")
    cold, t_cold = ws.relax(peval, 49152, e_ref=e_ref)
    t0 = time.perf_counter()
    psi0 = ws.predict(peval)
    t_inf = time.perf_counter() - t0
    log("warm run")
    warm, t_warm = ws.relax(peval, 49152, e_ref=e_ref, psi0s=psi0)
    met = lambda runs: [it["met"] for it in runs]  # noqa: E731
    ok_c, ok_w = all(m is not None for m in met(cold)), all(m is not None for m in met(warm))
    err_w = max((it["hist"][-1][2] - e) / abs(e) for it, e in zip(warm, e_ref))
    return {
        "grid": [1024, 32], "dt": 0.05, "storage": "complex f64 propagation, FP32 training solutions",
        "network": {"widths": ws.widths, "weights": "bf16 (fp32 Adam masters)", "accumulate": "fp32"},
        "train": {"potentials": args.c5_train, "curriculum": "first 256 cold, the rest warm-started from the first fit",
                  "cold_steps": args.c5_cold_steps, "warm_steps": args.c5_warm_steps, "updates_per_fit": args.c5_updates,
                  "generate_and_fit_s": t_train, "loss_first": losses[0], "loss_last": losses[-1],
                  "loss_unit": "per sample, sum over outputs of (error / output scale)^2"},
        "eval_potentials": args.c5_eval, "criterion": "E - E_ref <= 1e-8 |E_ref|, E_ref after 49152 cold steps",
        "cold": {"time_s": t_cold, "all_met": ok_c, "steps_max": max(it["steps"] for it in cold),
                 "steps_mean": float(np.mean([it["steps"] for it in cold]))},
        "warm": {"time_s": t_warm + t_inf, "inference_s": t_inf, "all_met": ok_w,
                 "steps_max": max(it["steps"] for it in warm),
                 "steps_mean": float(np.mean([it["steps"] for it in warm])), "max_rel_energy_above_ref": err_w},
        "speedup": t_cold / (t_warm + t_inf),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding of the path is not launched", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import split_operator_amd as so
    from split_operator_amd import workloads as W

    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    seed = 1 * 1000 + 0
    p, (radial, polar), (c6, c12, a) = W.radial_angular(
        so, n_r=N_R, n_theta=N_THETA, dt=DT, steps=args.steps, seed=seed, dtype="f32",
        gl=so.gauss_legendre(N_THETA))
    if not args.no_autotune:
        p.set_autotune(True)
    p.steps(args.warmup)
    so.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.steps(args.steps)
    so.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pts = N_R * N_THETA
    value = pts * args.steps / wall
    stats = p.stats()

    # FP32 and FP32-storage/FP64-arithmetic against FP64 storage after exactly 1000 steps of the same stack
    fin = {}
    for dt_ in ("f32", "f32x", "f64"):
        r, _, _ = W.radial_angular(so, n_r=N_R, n_theta=N_THETA, dt=DT, steps=1000, seed=seed, dtype=dt_,
                                   gl=so.gauss_legendre(N_THETA))
        r.steps(1000)
        fin[dt_] = r.wave_function().array
        del r
    l2_1000 = float(np.linalg.norm(fin["f32"] - fin["f64"]) / np.linalg.norm(fin["f64"]))
    l2x_1000 = float(np.linalg.norm(fin["f32x"] - fin["f64"]) / np.linalg.norm(fin["f64"]))
    del fin

    # the same workload with FP32 storage and FP64 pass arithmetic: throughput
    px, _, _ = W.radial_angular(so, n_r=N_R, n_theta=N_THETA, dt=DT, steps=args.steps, seed=seed, dtype="f32x",
                                gl=so.gauss_legendre(N_THETA))
    px.steps(args.warmup)
    so.synchronize()
    t0 = time.perf_counter()
    px.steps(args.steps)
    so.synchronize()
    wall_x = time.perf_counter() - t0
    del px

    # accuracy of the FP32 run against the same stack stored in FP64 on the GPU
    q, _, _ = W.radial_angular(so, n_r=N_R, n_theta=N_THETA, dt=DT, steps=args.steps, seed=seed, dtype="f64",
                               gl=so.gauss_legendre(N_THETA))
    q.steps(args.warmup + args.steps)
    a32 = p.wave_function().array
    a64 = q.wave_function().array
    l2 = float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64))
    del q

    # per-kernel timing of the steady-state step (overwrites p's state: done last)
    passes = [pass_record(t) for t in p.time_passes(args.kernel_iters)]
    dom = max(passes, key=lambda r: r["avg_us"])
    if dom["kernel"] == "legendre_gemm":
        achieved = dom["tflops"]
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_MFMA_PEAK_TFLOPS}
    else:
        roof = {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dom["gbs"] / HBM_PEAK_GBS}
    roof["kernel"] = dom["kernel"]
    roof["traffic"] = None
    tfile = os.path.join(ROOT, "profiles", "traffic_radial_angular_2048x128.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roof["traffic"] = json.load(f).get(dom["kernel"])

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        n_cpu = args.cpu_steps
        po, _, _ = W.radial_angular(O, n_r=N_R, n_theta=N_THETA, dt=DT, steps=n_cpu, seed=seed)
        po.steps(1)
        t1 = time.perf_counter()
        po.steps(n_cpu)
        cw = time.perf_counter() - t1
        cpu = {"value": pts * n_cpu / cw, "unit": "grid-point-steps/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"f64 C oracle, same 2048x128 stack, {n_cpu} steps after 1 warm-up step ({cw:.1f} s)"}

    extra = {}
    if args.c3_steps > 0:
        extra["cartesian_3d_256"] = cartesian_256(so, W, args.c3_steps, args.kernel_iters, not args.no_autotune)
    if args.c4_instances > 0:
        extra["ground_states_batched"] = ground_states(so, W, args.c4_instances, args.c4_max_steps,
                                                       args.c4_f32_steps, args.c4_solo)
    if args.c5_train > 0:
        extra["learned_warm_start"] = learned_warm_start(so, args)

    line = {
        "metric": METRIC, "value": value, "unit": "grid-point-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "radial_angular_realtime", "n_r": N_R, "n_theta": N_THETA, "dt": DT,
                   "potential": {"C6": c6, "C12": c12, "a": a}, "storage": "complex f32, f64 norms",
                   "border_damping": "last 10% of r, loss checked", "seed": seed},
        "roofline": roof,
        "cpu_baseline": cpu,
        "launches_per_step": stats["launches_per_step"],
        "rel_l2_vs_f64_gpu": l2,
        "rel_l2_vs_f64_gpu_1000_steps": l2_1000,
        "f32_storage_f64_arithmetic": {"value": N_R * N_THETA * args.steps / wall_x,
                                       "ms_per_step": wall_x / args.steps * 1e3,
                                       "rel_l2_vs_f64_gpu_1000_steps": l2x_1000},
        "passes": passes,
        "extra_workloads": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
