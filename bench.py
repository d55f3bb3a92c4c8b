"""Benchmark: steps/s of the 40k-vertex cloth drape on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one SimulationBase::step() of the 200x200-quad drape (40,401
vertices, 80,000 triangles, sphere R=0.5 + ground plane, mu=0.1, dt=1/200,
20 local-global iterations, self collision on) with state resident in HBM.
The cloth is advanced untimed to step --first-step (110: on the sphere and
folding onto itself), then --warmup steps, then --steps steps are timed. The
whole 500-step drape is also timed once. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

# thread placement of the CPU oracle's OpenMP runtime (read when it starts)
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel, info, name="drape40k"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    profiles/<name>_pmc_traffic.json (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    the same counter's reading for a device copy of known size), when it was measured on this
    workload, kernel and level structure (a summary without levels per launch or split counts was
    measured with one launch per level and unsplit reductions); otherwise None."""
    path = os.path.join(ROOT, "profiles", "%s_pmc_traffic.json" % name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("kernel") != kernel or d.get("levels_swept") != info["n_levels_swept"] or \
            d.get("n_vertices") != info["n_vertices"] or d.get("nnz_F") != info["nnz_F"] or \
            d.get("top_fold") != info["top_fold"] or d.get("top_chunks") != info["top_chunks"] or \
            d.get("parts_backward", [1] * info["n_levels_swept"]) != info["parts_backward"][:info["launches_backward"]] or \
            d.get("parts_forward", [1] * info["n_levels_swept"]) != info["parts_forward"][:info["launches_forward"]] or \
            d.get("split_backward", [1] * info["n_levels_swept"]) != info["split_backward"][:info["n_levels_swept"]]:
        return None
    return d.get("traffic_bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, budget_s=20.0, start=2, state=None, window=None, runs=3):
    """The CPU oracle (reference algorithm and OpenMP structure) on a bounded sample.
    With `state` (q, v at the first timed GPU step `start`), each of `runs` runs starts from
    that state and times the same steps the GPU timed (at most `window`, cut to the budget);
    the median run is reported. Without it the oracle steps from the scene's start to `start`."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc = po.Oracle(spec)
    if state is None:
        orc.step(start)  # warm-up into the first steps of the same trajectory
    else:
        orc.set_state(*state)
    t0 = time.perf_counter()
    orc.step(1)
    t1 = time.perf_counter() - t0
    nr = runs if state is not None else 1
    k = max(2, min(window or 100, int(budget_s / nr / max(t1, 1e-6))))
    rates = []
    for r in range(nr):
        if state is not None:
            orc.set_state(*state)
        t0 = time.perf_counter()
        orc.step(k)
        rates.append(k / (time.perf_counter() - t0))
    po.set_num_threads(1)
    if state is not None:
        orc.set_state(*state)
    t0 = time.perf_counter()
    orc.step(1)
    dt1 = time.perf_counter() - t0
    po.set_num_threads(threads)
    s0 = start if state is not None else start + 1
    out = {"value": float(np.median(rates)), "unit": "steps/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
           "runs": [round(x, 4) for x in rates],
           "sample": "steps %d..%d (%s), %d OpenMP threads, median of %d runs, factorisation excluded"
                     % (s0, s0 + k - 1, "from the GPU's state at step %d" % start if state is not None
                        else "the oracle's own trajectory", threads, len(rates)),
           "one_thread": {"value": 1 / dt1, "unit": "steps/s", "sample": "step %d, 1 thread" % s0}}
    return out


def warm_start_comparison(steps=2, cap=40000):
    """The learned warm start on its configured workload: the 40k drape from its state at step
    150 (self-collision on), step 151 untimed (it gives the previous and learned starts their
    previous reactions), then `steps` steps run to a 1e-6 mean friction residual with the plain,
    previous-step, learned and exact warm starts (climatenet_amd.learn.evaluate_drape40k); the
    exact converged reactions of the same step are the bound any obstacle-reaction warm start
    can reach."""
    from climatenet_amd import learn
    if not os.path.exists(learn.MODEL):
        return None
    res = learn.evaluate_drape40k(learn.load(learn.MODEL), start=150, steps=steps, cap=cap, log=lambda m: None)
    res["target_iteration_ratio"] = 1.5
    res["target_met"] = res["iteration_ratio_plain_over_learned"] >= 1.5
    return res


def kernel_roofline(sim, info, timing_steps, traffic_name):
    """Average launch time of each kernel family over `timing_steps` more steps with HIP events
    around every launch (the local kernel, each whole sweep divided by its launches, the top
    block), the dominant family by total time, and its HBM roofline: algorithmic bytes per launch
    (pdfc_info) over the live average launch time."""
    sim.set_kernel_timing(True)
    sim.step(timing_steps)
    kt = sim.kernel_times()
    sim.set_kernel_timing(False)
    per_fwd = max(1, info["launches_forward"])
    per_bwd = max(1, info["launches_backward"])
    fam = {
        "k_local": (kt["local_us"], info["bytes_local"], kt["local_launches"]),
        "k_forward": (kt["forward_us"], info["bytes_forward"] / per_fwd, kt["forward_launches"]),
        "k_backward": (kt["backward_us"], info["bytes_backward"] / per_bwd, kt["backward_launches"]),
        "k_top": (kt["top_us"], info["bytes_top"], kt["top_launches"]),
    }
    dom = max(fam, key=lambda k: fam[k][0] * fam[k][2])
    us, bytes_per_launch, _ = fam[dom]
    achieved = bytes_per_launch / (us * 1e-6) / 1e9 if us > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(dom, info, traffic_name),
                "kernel": dom, "avg_launch_us": round(us, 3), "algorithmic_bytes_per_launch": bytes_per_launch}
    times = {k: round(v[0], 3) for k, v in fam.items()}
    if kt["self_iterations"] > 0:
        times["k_self_levels"] = round(kt["self_us"], 3)  # all self-contact levels of an iteration
    return roofline, times


def cpu_baseline_batched(specs, q, v, ranges, budget_s=20.0, runs=3):
    """The 64 scenes one after another on the CPU oracle, each from its part of the batch's
    state: aggregate instance-steps/s over k steps of every scene (k cut to the budget), median
    of `runs` runs from the same states."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orcs = [po.Oracle(sp) for sp in specs]

    def reset():
        for o, (a, b) in zip(orcs, ranges):
            o.set_state(q[3 * a:3 * b].copy(), v[3 * a:3 * b].copy())

    reset()
    t0 = time.perf_counter()
    for o in orcs:
        o.step(1)
    one = time.perf_counter() - t0
    k = max(1, min(20, int(budget_s / runs / max(one, 1e-6))))
    rates = []
    for _ in range(runs):
        reset()
        t0 = time.perf_counter()
        for o in orcs:
            o.step(k)
        rates.append(len(orcs) * k / (time.perf_counter() - t0))
    return {"value": float(np.median(rates)), "unit": "instance-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "runs": [round(x, 3) for x in rates],
            "sample": "%d steps of each of the 64 scenes, one scene after another, from the batch's state after "
                      "its timed steps; %d OpenMP threads; median of %d runs; factorisation excluded"
                      % (k, threads, runs)}


def bench_batched(args):
    """64 independent cloths (N ~ U{32..128}, one sphere each) advanced as one batch, against
    the same 64 scenes run one after another; instance-steps/s of each."""
    from climatenet_amd import native, scenes
    specs = scenes.batched_instances(64, seed=0)
    t0 = time.perf_counter()
    sim = native.BatchSimulation(specs, record_iterations=False, use_graph=not args.no_graph)
    create_s = time.perf_counter() - t0
    sim.step(args.warmup)
    t0 = time.perf_counter()
    sim.step(args.steps)
    batched_s = time.perf_counter() - t0
    info = sim.info()
    q_end, v_end = sim.state()
    roofline, ktimes = kernel_roofline(sim, info, args.timing_steps, "batched64")
    seq_s = 0.0
    for sp in ([] if args.no_sequential else specs):
        one = native.Simulation(sp, record_iterations=False, use_graph=not args.no_graph)
        one.step(args.warmup)
        t1 = time.perf_counter()
        one.step(args.steps)
        seq_s += time.perf_counter() - t1
        one.close()
    value = len(specs) * args.steps / batched_s
    out = {
        "metric": "instance-steps/s, 64 batched cloths",
        "value": round(value, 3), "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * batched_s / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "64 cloths of N x N quads, N ~ U{32..128} (seed 0), 1 m side, one sphere each "
                               "(R ~ U[0.2, 0.6], mu ~ U[0, 0.8]), dt=1/100, 15 local-global iterations",
                   "total_vertices": info["n_vertices"], "levels": info["n_levels"],
                   "levels_swept": info["n_levels_swept"], "top_block_rows": info["top_m"],
                   "graph_replay": not args.no_graph, "create_seconds": round(create_s, 2),
                   "tuning": sim.tuning(), "waves_forward": info["waves_forward"][:info["n_levels_swept"]],
                   "waves_backward": info["waves_backward"][:info["n_levels_swept"]],
                   "levels_per_launch_forward": info["parts_forward"][:info["launches_forward"]],
                   "levels_per_launch_backward": info["parts_backward"][:info["launches_backward"]],
                   "split_backward": info["split_backward"][:info["n_levels_swept"]],
                   "top_pregather": info["top_pregather"],
                   "top_chunks": info["top_chunks"], "top_fold": info["top_fold"], "nnz_F": info["nnz_F"],
                   "autotune_ms": round(info["autotune_ms"], 1)},
        "sequential": None if args.no_sequential else {
            "value": round(len(specs) * args.steps / seq_s, 3), "unit": "instance-steps/s",
            "note": "the same 64 scenes, one simulation after another on the GPU"},
        "speedup_vs_sequential": None if args.no_sequential else round(seq_s / batched_s, 3),
        "roofline": roofline,
        "kernel_times_us": ktimes,
        # whole batch step: algorithmic bytes of its 15 iterations at the measured rate
        "step_roofline": {"bytes_per_step": info["bytes_per_iteration"] * specs[0].iterations,
                          "achieved": round(info["bytes_per_iteration"] * specs[0].iterations * args.steps
                                            / batched_s / 1e9, 3),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(info["bytes_per_iteration"] * specs[0].iterations * args.steps
                                        / batched_s / 1e9 / HBM_PEAK_GBS, 6)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batched(specs, q_end, v_end, sim.ranges)
    print(json.dumps(out), flush=True)


def bench_mlp(args):
    """The learned part's network alone: training (forward + MSE gradient + backward + Adam,
    BF16 MFMA with FP32 accumulation) and inference on 262,144 synthetic contact rows. Rates
    include the host-device copies of each call; the kernels' own times and MFMA counters come
    from rocprofv3 on this workload."""
    from climatenet_amd import abi, native
    rows = 1 << 18
    rng = np.random.default_rng(0)
    x = rng.standard_normal((rows, abi.MLP_FEATURES)).astype(np.float32)
    t = np.tanh(x[:, :3] * x[:, 3:6] + 0.5 * x[:, 6:9]).astype(np.float32)
    mlp = native.ContactMLP(seed=0)
    norm = np.concatenate([np.zeros(12), np.ones(12), np.ones(3)]).astype(np.float32)
    p, _ = mlp.params()
    mlp.set_params(p, norm)
    mlp.forward(x[:4096])
    mlp.train(x[:4096], t[:4096], 1, batch=512)
    epochs = max(1, args.steps // 50)
    t0 = time.perf_counter()
    losses = mlp.train(x, t, epochs, batch=512, lr=1e-3)
    train_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        mlp.forward(x)
    fwd_s = (time.perf_counter() - t0) / reps
    flop_fwd = 2 * (12 * 64 + 64 * 64 + 64 * 64 + 64 * 3)  # 18,304 per row
    out = {
        "metric": "MLP training rows/s (contact warm-start network)",
        "value": round(rows * epochs / train_s, 1), "unit": "rows/s", "n_gpus": 1, "steps": epochs,
        "warmup": 1, "ms_per_step": round(1e3 * train_s / epochs, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "12-64-64-64-3 SiLU MLP, %d rows, batch 512, Adam" % rows, "epochs": epochs},
        "inference_rows_per_s": round(rows / fwd_s, 1),
        "flop_per_row": {"forward": flop_fwd, "training": 3 * flop_fwd},
        "end_to_end_tflops": {"training": round(3 * flop_fwd * rows * epochs / train_s / 1e12, 4),
                              "inference": round(flop_fwd * rows / fwd_s / 1e12, 4)},
        "final_loss": float(losses[-1]),
    }
    print(json.dumps(out), flush=True)


def bench_stacked(args):
    """The self-collision-heavy configuration: three 150 x 150-quad sheets 5 mm apart (68,403
    vertices) on a sphere, mu 0.3, dt 1/200, 30 iterations, self tolerance 4e-3, in FP32
    substitution with FP64 refinement. After the timed steps, one more step from the same
    state in FP32+refinement and in FP64 gives the per-step error of the precision mode."""
    from climatenet_amd import abi, native, scenes
    spec = scenes.stacked_layers()
    sim = native.Simulation(spec, precision=abi.FP32_REFINE64, record_iterations=False)
    sim.step(args.warmup)
    t0 = time.perf_counter()
    sim.step(args.steps)
    elapsed = time.perf_counter() - t0
    info = sim.info()
    q, v = sim.state()
    ref = native.Simulation(spec, precision=abi.FP64, record_iterations=False)
    ref.set_state(q, v)
    sim.step(1)
    ref.step(1)
    (q32, v32), (q64, v64) = sim.state(), ref.state()
    err = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    out = {
        "metric": "steps/s, 3 stacked 150x150 sheets (self-collision heavy)",
        "value": round(args.steps / elapsed, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64-refine", "data": "synthetic",
        "config": {"workload": "3 x 150^2 quads (68,403 vertices), sphere R=0.4, mu=0.3, dt=1/200, 30 iterations, "
                               "self-collision tol 4e-3, FP32 substitution + 2 FP64 refinement sweeps",
                   "n_vertices": info["n_vertices"], "levels": info["n_levels"], "top_block_rows": info["top_m"],
                   "self_contacts_last_step": info["n_self_contacts"], "contacts_last_step": info["n_contacts"]},
        "error_vs_fp64_one_step": {"q": err(q32, q64), "v": err(v32, v64)},
    }
    print(json.dumps(out), flush=True)


def bench_meshdrape(args):
    """A scene shaped like the reference's replicability drapes: a 5,929-vertex cloth on a
    key-framed rotating sphere mesh (200 frames, 5,120 faces) and a ground plane; steps/s
    over the timed steps after the cloth has reached the sphere, with the device time per step
    split into the part before the first iteration (mesh refresh, detection) and the iterations."""
    from climatenet_amd import native, scenes
    spec = scenes.rotating_mesh_drape()
    sim = native.Simulation(spec, record_iterations=True)
    sim.step(60 + args.warmup)  # the cloth reaches the sphere after about 40 steps
    t0 = time.perf_counter()
    sim.step(args.steps)
    elapsed = time.perf_counter() - t0
    info = sim.info()
    ps = sim.phase_stats()
    out = {
        "metric": "steps/s, cloth on a key-framed rotating mesh sphere",
        "value": round(args.steps / elapsed, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 60 + args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "76x76-quad 2.4 m cloth (5,929 vertices), icosphere R=0.5 as a 200-frame moving "
                               "mesh (5,120 faces, tolerance 0.015) + plane, mu=0.1, dt=1/200, 20 iterations",
                   "contacts_last_step": info["n_contacts"], "levels": info["n_levels"],
                   "top_block_rows": info["top_m"], "tuning": sim.tuning()},
        "device_us_per_step": {"to_first_iteration": round(ps["collision_us"], 1),
                               "iterations": round(ps["iterations_us"], 1)},
    }
    if not args.no_cpu_baseline:  # the oracle from step 60, when the cloth rests on the sphere
        out["cpu_baseline"] = cpu_baseline(spec, budget_s=10.0, start=60)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=["fp64", "fp32r"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the plain vs learned warm-start comparison (iterations to a 1e-6 residual)")
    ap.add_argument("--timing-steps", type=int, default=5)
    ap.add_argument("--first-step", type=int, default=110,
                    help="drape: untimed steps before the warm-up (110: the cloth is on the sphere and "
                         "folds onto itself, the self-contact phase)")
    ap.add_argument("--no-full-drape", action="store_true",
                    help="drape: skip timing the whole 500-step drape (to profile the window alone)")
    ap.add_argument("--no-self-collision", action="store_true",
                    help="disable self-collision (not the benchmark configuration)")
    ap.add_argument("--workload", choices=["drape", "batched", "stacked", "meshdrape", "mlp"], default="drape",
                    help="drape: the 40k-vertex drape (default); batched: 64 independent cloths; stacked: three "
                         "stacked sheets, FP32 + FP64 refinement; meshdrape: cloth on a key-framed mesh sphere; "
                         "mlp: the warm-start network alone")
    ap.add_argument("--no-graph", action="store_true", help="batched: launch kernels eagerly")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched: skip the 64 scenes run one after another on the GPU (profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device 4 times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8, "copies": 4}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "mlp":
        bench_mlp(args)
        return
    if args.workload == "stacked":
        bench_stacked(args)
        return
    if args.workload == "meshdrape":
        bench_meshdrape(args)
        return

    from climatenet_amd import abi, native, scenes

    spec = scenes.sphere_drape_40k()
    if args.no_self_collision:
        spec.self_collision = False
    prec = abi.FP64 if args.precision == "fp64" else abi.FP32_REFINE64
    # the whole drape once (500 steps from the flat start), then the timed window
    full_s = None
    if not args.no_full_drape:
        full = native.Simulation(spec, precision=prec, refinement_sweeps=2, record_iterations=True)
        full.step(1)
        t0 = time.perf_counter()
        full.step(spec.steps - 1)
        full_s = time.perf_counter() - t0
        full.close()
    sim = native.Simulation(spec, precision=prec, refinement_sweeps=2, record_iterations=True)
    sim.step(args.first_step + args.warmup)  # untimed advance into the contact phase, then warm-up
    window_state = sim.state()
    t0 = time.perf_counter()
    sim.step(args.steps)  # returns after the stream is synchronised
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed
    phases = sim.phase_stats()
    first_timed = args.first_step + args.warmup + 1

    # kernel times: instrumented continuation of the same run (events around every launch)
    info = sim.info()
    K = spec.iterations
    roofline, ktimes = kernel_roofline(sim, info, args.timing_steps, "drape40k")
    info = sim.info()

    out = {
        "metric": "steps/s on 40k-vertex cloth drape",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.FP64 else "f32+f64-refine",
        "data": "synthetic",
        "config": {"workload": "drape_200x200: 40,401 vertices, sphere R=0.5 + plane, mu=0.1, dt=1/200, "
                               "20 local-global iterations, self-collision %s" % ("on (tol 5e-3)" if spec.self_collision else "off"),
                   "first_step": first_timed, "untimed_advance": args.first_step,
                   "timed_steps": "%d..%d of %d" % (first_timed, first_timed + args.steps - 1, spec.steps),
                   "window_contacts_per_step": round(phases["contacts_sum"] / args.steps, 1),
                   "window_self_contacts_per_step": round(phases["self_contacts_sum"] / args.steps, 1),
                   "window_self_levels_max": phases["self_levels_max"],
                   "self_contacts_last_step": info["n_self_contacts"],
                   "iterations": K, "n_vertices": info["n_vertices"], "nnz_L": info["nnz_L"],
                   "nnz_F": info["nnz_F"], "levels": info["n_levels"], "levels_swept": info["n_levels_swept"],
                   "top_block_rows": info["top_m"], "contacts_last_step": info["n_contacts"],
                   "waves_forward": info["waves_forward"][:info["n_levels_swept"]],
                   "waves_backward": info["waves_backward"][:info["n_levels_swept"]],
                   "levels_per_launch_forward": info["parts_forward"][:info["launches_forward"]],
                   "levels_per_launch_backward": info["parts_backward"][:info["launches_backward"]],
                   "split_backward": info["split_backward"][:info["n_levels_swept"]],
                   "top_pregather": info["top_pregather"],
                   "leaf_size": info["leaf_size"], "top_chunks": info["top_chunks"], "top_fold": info["top_fold"],
                   "autotune_ms": round(info["autotune_ms"], 1), "tuning": sim.tuning()},
        "roofline": roofline,
        # whole step: the algorithmic bytes of its iterations at the measured rate
        "step_roofline": {"bytes_per_step": info["bytes_per_iteration"] * K,
                          "achieved": round(info["bytes_per_iteration"] * K * steps_per_s / 1e9, 3),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(info["bytes_per_iteration"] * K * steps_per_s / 1e9 / HBM_PEAK_GBS, 6)},
        "kernel_times_us": ktimes,
        "self_collision_host_us_per_step": {"round_trip": round(phases["detection_us"], 1),
                                            "pairing": round(phases["pairing_us"], 1),
                                            "of_which_levels": round(phases["order_us"], 1),
                                            "of_which_upload": round(phases["upload_us"], 1)},
    }
    if full_s is not None:
        out["full_drape"] = {"value": round((spec.steps - 1) / full_s, 3), "unit": "steps/s",
                             "sample": "steps 2..%d of the drape from its flat start (step 1 untimed)" % spec.steps}
    if not args.no_warm_start:
        out["warm_start"] = warm_start_comparison()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, budget_s=24.0, start=first_timed, state=window_state,
                                           window=args.steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
