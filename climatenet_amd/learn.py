"""Learned warm start: contact data from the simulator, training on the GPU with the
library's own MLP kernels, and the comparison against the plain warm start.

    python -m climatenet_amd.learn generate --out data.npz [--samples 1000000]
    python -m climatenet_amd.learn train --data data.npz --out contact_mlp.npz
    python -m climatenet_amd.learn evaluate --model contact_mlp.npz [--steps 40]

Data: runs of the 20x20 sphere drape (jitter seeds 0, 1, ...) and of the three stacked
150x150 sheets, every step iterated until the mean friction (Alart-Curnier) residual
is below 1e-6 (at most 1000 iterations). A sample is one obstacle contact: the 12
features at the start of its step and the converged reaction per unit mass in the
contact basis (pdfc_get_warm_data). Only converged steps are kept.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import abi, native, scenes

TOL = 1e-6
CAP = 1000


def generate(samples=1_000_000, tol=TOL, cap=CAP, log=print):
    """Converged contact samples from the 20x20 drape (seeds 0..) and the stacked sheets."""
    X, Y = [], []
    total = 0

    def run(spec, steps, precision):
        nonlocal total
        sim = native.Simulation(spec, precision=precision, residual_tolerance=tol, max_iterations=cap)
        t0 = time.perf_counter()
        kept = 0
        for s in range(steps):
            sim.step(1)
            ps = sim.phase_stats()
            if ps["converged_steps"] != 1:
                continue
            x, y = sim.warm_data()
            if len(x):
                X.append(x)
                Y.append(y)
                kept += len(x)
                total += len(x)
            if total >= samples:
                break
        log("  %s: %d samples, %.1f s" % (spec.name, kept, time.perf_counter() - t0))

    # the stacked sheets first: most of the samples, FP32 substitution + FP64 refinement
    run(scenes.stacked_layers(), 400, abi.FP32_REFINE64)
    seed = 0
    while total < samples and seed < 200:
        run(scenes.sphere_drape_small(jitter_seed=seed), 200, abi.FP64)
        seed += 1
    X = np.concatenate(X)[:samples]
    Y = np.concatenate(Y)[:samples]
    return X, Y


def normalisation(X, Y):
    """Feature mean and std (std floored), per-output scale = std of the target."""
    norm = np.zeros(abi.MLP_NORM, dtype=np.float32)
    norm[:12] = X.mean(axis=0)
    norm[12:24] = np.maximum(X.std(axis=0), 1e-6)
    norm[24:27] = np.maximum(Y.std(axis=0), 1e-9)
    return norm


def train(X, Y, epochs=30, batch=512, lr=2e-3, seed=0, log=print):
    mlp = native.ContactMLP(seed=seed)
    mlp.set_params(None, normalisation(X, Y))
    t0 = time.perf_counter()
    losses = []
    # step decay: three stages of the learning rate
    stages = [(lr, epochs // 2), (lr / 4, epochs // 3), (lr / 16, epochs - epochs // 2 - epochs // 3)]
    for k, (r, e) in enumerate(stages):
        if e <= 0:
            continue
        losses += list(mlp.train(X, Y, epochs=e, batch=batch, lr=r, seed=seed + k))
    dt = time.perf_counter() - t0
    log("  trained %d epochs on %d samples in %.1f s, loss %.4g -> %.4g" % (epochs, len(X), dt, losses[0], losses[-1]))
    return mlp, np.array(losses), dt


def save(path, mlp):
    p, z = mlp.params()
    np.savez(path, params=p, norm=z)


def load(path):
    d = np.load(path, allow_pickle=False)
    mlp = native.ContactMLP()
    mlp.set_params(d["params"], d["norm"])
    return mlp


def evaluate(spec, mlp, steps, warmup_steps=0, modes=(abi.WARM_PLAIN, abi.WARM_PREVIOUS, abi.WARM_LEARNED),
             precision=abi.FP64, tol=TOL, cap=CAP, log=print):
    """Runs `steps` steps of `spec` from the state after `warmup_steps` fixed-iteration steps,
    once per warm start; returns per mode the mean iterations to the tolerance over steps
    with contacts, steps/s, fallbacks and the final state."""
    base = native.Simulation(spec, precision=precision)
    if warmup_steps:
        base.step(warmup_steps)
    q0, v0 = base.state()
    base.close()
    out = {}
    for mode in modes:
        sim = native.Simulation(spec, precision=precision, residual_tolerance=tol, max_iterations=cap,
                                warm_start=mode, mlp=mlp if mode == abi.WARM_LEARNED else None)
        sim.set_state(q0, v0)
        it_tol, it_exec, conv, fb, csteps = 0, 0, 0, 0, 0
        t0 = time.perf_counter()
        for s in range(steps):
            sim.step(1)
            ps = sim.phase_stats()
            it_tol += ps["iterations_to_tolerance"]
            it_exec += ps["iterations"]
            conv += ps["converged_steps"]
            fb += ps["fallbacks"]
            csteps += ps["contact_steps"]
        dt = time.perf_counter() - t0
        q, v = sim.state()
        name = {abi.WARM_PLAIN: "plain", abi.WARM_PREVIOUS: "previous", abi.WARM_LEARNED: "learned"}[mode]
        out[name] = {"iterations_to_tolerance": it_tol / max(1, csteps), "iterations_executed": it_exec / steps,
                     "converged_steps": conv, "contact_steps": csteps, "fallbacks": fb,
                     "steps_per_s": steps / dt, "q": q, "v": v}
        log("  %-8s iterations to %.0e: %.1f (executed %.1f), converged %d/%d, fallbacks %d, %.2f steps/s"
            % (name, tol, out[name]["iterations_to_tolerance"], out[name]["iterations_executed"], conv, csteps, fb,
               steps / dt))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--out", required=True)
    g.add_argument("--samples", type=int, default=1_000_000)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=512)
    t.add_argument("--lr", type=float, default=2e-3)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", required=True)
    e.add_argument("--steps", type=int, default=40)
    e.add_argument("--warmup", type=int, default=150)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "generate":
        X, Y = generate(a.samples, log=log)
        np.savez(a.out, X=X, Y=Y)
        log("%d samples -> %s" % (len(X), a.out))
    elif a.cmd == "train":
        d = np.load(a.data, allow_pickle=False)
        mlp, losses, dt = train(d["X"], d["Y"], a.epochs, a.batch, a.lr, log=log)
        save(a.out, mlp)
        log(json.dumps({"epochs": a.epochs, "samples": int(len(d["X"])), "seconds": dt,
                        "loss_first": float(losses[0]), "loss_last": float(losses[-1])}))
    else:
        mlp = load(a.model)
        res = evaluate(scenes.sphere_drape_40k(), mlp, a.steps, a.warmup, log=log)
        log(json.dumps({k: {kk: vv for kk, vv in v.items() if kk not in ("q", "v")} for k, v in res.items()}))


if __name__ == "__main__":
    sys.exit(main())
