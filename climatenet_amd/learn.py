"""Learned warm start: contact data from the simulator, training on the GPU with the
library's own MLP kernels, and the comparison against the plain warm start.

    python -m climatenet_amd.learn generate --out data.npz
    python -m climatenet_amd.learn train --data data.npz --out contact_mlp.npz
    python -m climatenet_amd.learn evaluate --model contact_mlp.npz

Data: runs of the 20x20 sphere drape (jitter seeds 0..4, 200 steps, FP64) and of the
three stacked 150x150 sheets (FP32 substitution + FP64 refinement, from the first
contacts), every step iterated until the mean friction (Alart-Curnier) residual is
below 1e-6 or the iteration cap is reached. A sample is one obstacle contact: the 12
features at the start of its step and the final reaction per unit mass in the contact
basis (pdfc_get_warm_data), with the residual its step reached.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import abi, native, scenes

TOL = 1e-6
CAP = 8000
HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "data", "contact_mlp.npz")


def run_for_data(spec, steps, precision=abi.FP64, start_steps=0, cap=CAP, tol=TOL, log=print):
    """(features, targets, residual reached) of every contact step of `steps` adaptive steps."""
    sim = native.Simulation(spec, precision=precision)
    if start_steps:
        sim.step(start_steps)
    q, v = sim.state()
    sim.close()
    sim = native.Simulation(spec, precision=precision, residual_tolerance=tol, max_iterations=cap)
    sim.set_state(q, v)
    X, Y, R = [], [], []
    t0 = time.perf_counter()
    for s in range(steps):
        sim.step(1)
        if not sim.phase_stats()["contact_steps"]:
            continue
        x, y = sim.warm_data()
        m, _ = sim.iteration_stats()
        X.append(x)
        Y.append(y)
        R.append(np.full(len(x), m[-1], dtype=np.float32))
    n = sum(len(x) for x in X)
    log("  %s: %d samples, %.1f s" % (spec.name, n, time.perf_counter() - t0))
    if not X:
        return np.zeros((0, 12), np.float32), np.zeros((0, 3), np.float32), np.zeros(0, np.float32)
    return np.concatenate(X), np.concatenate(Y), np.concatenate(R)


def generate(small_seeds=5, stacked_steps=25, stacked_cap=4000, log=print):
    parts = [run_for_data(scenes.sphere_drape_small(jitter_seed=s), 200, log=log) for s in range(small_seeds)]
    parts.append(run_for_data(scenes.stacked_layers(), stacked_steps, abi.FP32_REFINE64, start_steps=40,
                              cap=stacked_cap, log=log))
    return tuple(np.concatenate([p[k] for p in parts]) for k in range(3))


def normalisation(X, Y):
    """Feature mean and std, per-output scale = std of the target. The std is floored at a
    quarter of the mean's magnitude, so a feature that is constant in the data (mu, when all
    runs share one friction coefficient) maps other values to moderate inputs."""
    norm = np.zeros(abi.MLP_NORM, dtype=np.float32)
    norm[:12] = X.mean(axis=0)
    norm[12:24] = np.maximum(X.std(axis=0), np.maximum(0.25 * np.abs(X.mean(axis=0)), 1e-6))
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
    g.add_argument("--small-seeds", type=int, default=5)
    g.add_argument("--stacked-steps", type=int, default=25)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--out", default=MODEL)
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=MODEL)
    e.add_argument("--steps", type=int, default=10)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "generate":
        X, Y, R = generate(a.small_seeds, a.stacked_steps, log=log)
        np.savez(a.out, X=X, Y=Y, R=R)
        log("%d samples -> %s (residual reached: median %.2e, max %.2e)" % (len(X), a.out, np.median(R), R.max()))
    elif a.cmd == "train":
        d = np.load(a.data, allow_pickle=False)
        mlp, losses, dt = train(d["X"], d["Y"], a.epochs, a.batch, a.lr, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        save(a.out, mlp)
        X, Y = d["X"], d["Y"]
        scale = np.linalg.norm(Y, axis=1).mean()
        err = np.linalg.norm(mlp.forward(X) - Y, axis=1) / (np.linalg.norm(Y, axis=1) + 0.1 * scale)
        log(json.dumps({"epochs": a.epochs, "samples": int(len(X)), "seconds": round(dt, 2),
                        "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                        "relative_error_median": float(np.median(err)),
                        "relative_error_p90": float(np.percentile(err, 90))}))
    else:
        mlp = load(a.model)
        res = evaluate(scenes.sphere_drape_small(jitter_seed=7), mlp, a.steps, 60, log=log)
        log(json.dumps({k: {kk: vv for kk, vv in v.items() if kk not in ("q", "v")} for k, v in res.items()}))


if __name__ == "__main__":
    sys.exit(main())
