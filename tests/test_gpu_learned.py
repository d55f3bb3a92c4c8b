"""GPU: the learned warm start on its configured workload, the 40k-vertex drape.

The committed network (climatenet_amd/data/contact_mlp.npz, climatenet_amd.learn.build) was
trained on the GPU by the library's own kernels on the configured data set: 1e6 obstacle
contacts from runs of the 20x20 drape and of the three stacked sheets, each step iterated to a
1e-6 friction residual or 1,000 iterations. It predicts the change of each contact's reaction
from the previous step's; the change is scaled by the factor that fits held-out contacts best,
then by the held-start objective on drapes outside the training data, which keeps the previous
reactions alone (factor 0) unless the change saves at least 3% of the iterations (it saved 1%).

From the drape's state at step 150 (self-collision on, as configured), step 151 is run to a
1e-6 mean Alart-Curnier residual (coulomb_util.cpp:101-152) untimed, which gives the previous
and learned starts their previous reactions, then step 152 with each warm start: plain (the
reference), previous (step 151's converged reactions), learned, and exact (the plain run's
converged reactions of step 152, held while the residual keeps falling), the bound of any warm
start of the obstacle reactions. The printed numbers are the result; the assertions are the
guarantees: every start reaches the tolerance from the same state, the learned start ends at
the plain start's solution as closely as the other converged starts do (runs to a 1e-6 friction
residual agree to about 1e-6 relative) and needs no more iterations than plain, and the exact
reactions do save iterations (the warm-start path works).
"""
import os

import numpy as np
import pytest

from climatenet_amd import learn

pytestmark = pytest.mark.gpu


def test_learned_warm_start_on_the_40k_drape(gpu):
    assert os.path.exists(learn.MODEL)
    mlp = learn.load(learn.MODEL)
    print("40k drape: 150 steps, step 151 untimed, then step 152 per warm start to 1e-6", flush=True)
    res = learn.evaluate_drape40k(mlp, start=150, steps=1, cap=40000)
    P = res["plain"]["iterations_to_tolerance"]
    L = res["learned"]["iterations_to_tolerance"]
    E = res["exact"]["iterations_to_tolerance"]
    V = res["previous"]["iterations_to_tolerance"]
    print("iterations to 1e-6: plain %.0f, previous %.0f, learned %.0f, exact reactions %.0f" % (P, V, L, E))
    print("plain / learned = %.4f (target 1.5), plain / previous = %.4f, plain / exact = %.4f (bound)"
          % (P / L, P / V, P / E))
    diffs = {m: res[m]["max_rel_diff_q_vs_plain"] for m in ("previous", "learned", "exact")}
    print("max relative difference of q from the plain run:", diffs)
    assert res["same_state_after_first_step"]
    for mode in ("plain", "previous", "learned", "exact"):
        assert res[mode]["converged_steps"] == 1, (mode, res[mode])
    assert diffs["learned"] <= 10 * max(diffs["previous"], diffs["exact"], 1e-9)
    assert L <= P
    assert E < P
