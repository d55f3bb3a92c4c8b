"""GPU: what the trained contact network (climatenet_amd/data/contact_mlp.npz, trained on the
GPU by the library's own kernels on converged contacts of the 20x20 drape and the stacked
sheets) gains over the plain warm start, in iterations to a 1e-6 friction residual.

Measured on three steps of the 20x20 drape (jitter seed 7, a trajectory not in the
training data) from the same state: plain, learned, and the exact converged reactions
of the same step held for 1000 iterations (the bound on what any reaction warm start
can give). The printed numbers are the result; the assertions are the guarantees:
the learned start converges to the plain start's solution (no accuracy loss), costs
at most the fallback's stagnation window more than plain, and the bound shows the
warm-start path itself works.
"""
import os

import numpy as np
import pytest

from climatenet_amd import abi, learn, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-6
CAP = 8000


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_committed_network_against_plain_and_exact(gpu):
    assert os.path.exists(learn.MODEL)
    mlp = learn.load(learn.MODEL)
    spec = scenes.sphere_drape_small(jitter_seed=7)
    base = gpu.Simulation(spec)
    base.step(60)
    q, v = base.state()
    plain = gpu.Simulation(spec, residual_tolerance=TOL, max_iterations=CAP)
    learned = gpu.Simulation(spec, residual_tolerance=TOL, max_iterations=CAP, warm_start=abi.WARM_LEARNED, mlp=mlp)
    exact = gpu.Simulation(spec, residual_tolerance=TOL, max_iterations=CAP)
    plain.set_state(q, v)
    learned.set_state(q, v)
    rows = []
    for s in range(3):
        q0, v0 = plain.state()
        plain.step(1)
        learned.step(1)  # its own trajectory: its previous-step reactions are its own
        exact.set_state(q0, v0)
        exact.set_warm_reactions(plain.contacts(), 1000)
        exact.step(1)
        pp, pl, pe = plain.phase_stats(), learned.phase_stats(), exact.phase_stats()
        for ps in (pp, pl, pe):
            assert ps["contact_steps"] == 1 and ps["converged_steps"] == 1, ps
        rows.append((pp["iterations_to_tolerance"], pl["iterations_to_tolerance"], pe["iterations_to_tolerance"],
                     pl["fallbacks"]))
        # no accuracy loss: both at the tolerance, the same fixed point every step
        qp, vp = plain.state()
        ql, vl = learned.state()
        assert rel_err(ql, qp) < 1e-5 and rel_err(vl, vp) < 1e-2
    P, L, E = (np.mean([r[k] for r in rows]) for k in range(3))
    print("iterations to %.0e: plain %.0f, learned %.0f (fallbacks %d), exact reactions held 1000: %.0f"
          % (TOL, P, L, sum(r[3] for r in rows), E))
    print("learned / plain = %.3f, exact / plain = %.3f" % (L / P, E / P))
    assert L <= P + 1000
    assert E < 0.6 * P
