"""Writes the golden fixtures of this directory from the CPU oracle (run in this repository;
the reference itself is not involved):

drape20_oracle.npz  the 20x20 sphere drape (441 vertices, mu 0.3, dt 1/100, 10 iterations, FP64,
                    initial jitter 1e-4 from mt19937_64 seed 0) for 200 steps: q and v after steps
                    1, 99, 100, 199, 200; the obstacle contacts (vertex, reaction in the contact
                    basis, normal) and the per-iteration mean / max friction residual of steps
                    100 and 200.
incline_mesh_oracle.npz  the closed-form sliding example on a triangle-mesh support: the mean
                    velocity of the cloth after every one of 800 steps.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from climatenet_amd import scenes  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

KEEP = (1, 99, 100, 199, 200)


def drape():
    spec = scenes.sphere_drape_small()
    orc = po.Oracle(spec)
    out = {}
    for s in range(1, 201):
        orc.step(1)
        if s in KEEP:
            q, v = orc.state()
            out["q%d" % s] = q
            out["v%d" % s] = v
        if s in (100, 200):
            cs = orc.contacts()
            out["cv%d" % s] = np.array([c.vertex for c in cs], dtype=np.int64)
            out["cr%d" % s] = np.array([list(c.r_local) for c in cs]).reshape(-1, 3)
            out["cn%d" % s] = np.array([list(c.normal) for c in cs]).reshape(-1, 3)
            mean, mx = orc.iteration_stats()
            out["acmean%d" % s] = mean
            out["acmax%d" % s] = mx
    np.savez_compressed(os.path.join(HERE, "drape20_oracle.npz"), **out)


def incline():
    spec = scenes.analytic_incline_mesh()
    orc = po.Oracle(spec)
    V = []
    for _ in range(spec.steps):
        orc.step(1)
        V.append(orc.state()[1].reshape(-1, 3).mean(axis=0))
    np.savez_compressed(os.path.join(HERE, "incline_mesh_oracle.npz"), v_mean=np.array(V))


if __name__ == "__main__":
    drape()
    incline()
