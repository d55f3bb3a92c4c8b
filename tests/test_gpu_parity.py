"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Per-step parity is checked from identical states (both sides start every
step from the oracle's state), because contact predicates are discontinuous
and free-running trajectories of a chaotic system drift apart.
Tolerances (FP64 mode): positions, velocities and contact reactions within
1e-9 relative (max-norm, relative to the largest magnitude of the quantity);
contact sets identical.
"""
import re

import numpy as np
import pytest

from climatenet_amd import abi, scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

REL = 1e-9


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def contact_arrays(cs):
    key = np.array([(c.obstacle, c.vertex) for c in cs], dtype=np.int64).reshape(-1, 2)
    r = np.array([list(c.r_local) for c in cs]).reshape(-1, 3)
    nrm = np.array([list(c.normal) for c in cs]).reshape(-1, 3)
    act = np.array([c.active for c in cs], dtype=np.int64)
    return key, r, nrm, act


@pytest.fixture(params=["top", "levels"])
def solve_mode(request, monkeypatch):
    """'top': default dense top block (the whole matrix for small meshes);
    'levels': supernodal sweeps only (PDFC_TOP_MB=0)."""
    if request.param == "levels":
        monkeypatch.setenv("PDFC_TOP_MB", "0")
    else:
        monkeypatch.delenv("PDFC_TOP_MB", raising=False)
    return request.param


def test_solve_matches_oracle(gpu, solve_mode):
    spec = scenes.sphere_drape_small()
    sim = gpu.Simulation(spec)
    info = sim.info()
    assert (info["top_m"] > 0) == (solve_mode == "top")
    orc = po.Oracle(spec)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(3 * spec.n_vertices)
    x = sim.solve(b)
    xo = orc.solve(b)
    assert rel_err(x, xo) < 1e-12
    # residual against the oracle's K (independent assembly)
    K = orc.K()
    X = x.reshape(-1, 3)
    res = np.abs(K @ X - b.reshape(-1, 3)).max() / np.abs(b).max()
    assert res < 1e-12


def test_solve_40k_residual(gpu, solve_mode):
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    sim = gpu.Simulation(spec)
    info = sim.info()
    assert (info["top_m"] > 0) == (solve_mode == "top")
    if solve_mode == "top":
        assert info["n_levels_swept"] < info["n_levels"]
    orc = po.Oracle(spec, iterations=1)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(3 * spec.n_vertices)
    x = sim.solve(b)
    K = orc.K()
    res = np.abs(K @ x.reshape(-1, 3) - b.reshape(-1, 3)).max() / np.abs(b).max()
    assert res < 1e-11
    assert rel_err(x, orc.solve(b)) < 1e-11


@pytest.mark.parametrize("steps", [200])
def test_small_drape_identical_state_parity(gpu, steps, solve_mode):
    """20x20 drape: every step from the oracle's state; q, v, contacts, reactions."""
    spec = scenes.sphere_drape_small()
    sim = gpu.Simulation(spec)
    orc = po.Oracle(spec)
    worst = dict(q=0.0, v=0.0, r=0.0)
    n_contact_steps = 0
    for s in range(steps):
        q0, v0 = orc.state()
        sim.set_state(q0, v0)
        orc.step(1)
        sim.step(1)
        q, v = sim.state()
        qo, vo = orc.state()
        worst["q"] = max(worst["q"], rel_err(q, qo))
        worst["v"] = max(worst["v"], rel_err(v, vo))
        ko, ro, no, ao = contact_arrays(orc.contacts())
        kg, rg, ng, ag = contact_arrays(sim.contacts())
        assert np.array_equal(ko, kg), "contact sets differ at step %d" % s
        if len(ko):
            n_contact_steps += 1
            assert np.array_equal(ao, ag), "take-off flags differ at step %d" % s
            assert rel_err(ng, no) < 1e-14
            worst["r"] = max(worst["r"], rel_err(rg, ro))
        mo, Mo = orc.iteration_stats()
        mg, Mg = sim.iteration_stats()
        if len(ko):
            assert np.allclose(mg, mo, rtol=1e-6, atol=1e-12), (s, mg, mo)
    print("worst relative errors:", worst, "steps with contact:", n_contact_steps)
    assert n_contact_steps > 100
    assert worst["q"] < REL and worst["v"] < REL and worst["r"] < REL, worst


def test_free_fall_closed_form(gpu):
    """No contact, undeformed cloth: v_{n+1} = v_n + dt g, q_{n+1} = q_n + dt v_{n+1}."""
    rest, tris = scenes.square_grid(10, 1.0, 5.0)
    spec = scenes.SceneSpec("freefall", rest, tris, None,
                            [scenes.ObstacleSpec("sphere", (0.0, 0.0, 0.0), radius=0.25)],
                            np.array([[0.3]]), time_step=0.01, iterations=10)
    sim = gpu.Simulation(spec)
    q = rest.copy().reshape(-1, 3)
    v = np.zeros_like(q)
    g = np.array(spec.gravity)
    for _ in range(20):
        sim.step(1)
        v = v + spec.time_step * g
        q = q + spec.time_step * v
    qs, vs = sim.state()
    assert np.abs(vs.reshape(-1, 3) - v).max() < 1e-9
    assert np.abs(qs.reshape(-1, 3) - q).max() < 1e-9


def test_inclined_plane_closed_form(gpu):
    """Sliding on a 45 degree plane with mu = 0.1: tangential acceleration
    g (sin th - mu cos th) (replicability/analytic_example_data_processing.py:53-147)."""
    mu, th = 0.1, np.deg2rad(45.0)
    spec = scenes.inclined_plane(mu=mu)
    sim = gpu.Simulation(spec)
    dt = spec.time_step
    tangent = np.array([np.cos(th), 0.0, np.sin(th)])
    speeds = []
    for s in range(400):
        sim.step(1)
        _, v = sim.state()
        speeds.append(v.reshape(-1, 3).mean(axis=0) @ -tangent)
    speeds = np.array(speeds)
    # after landing the tangential speed grows linearly with slope g (sin - mu cos)
    acc = np.polyfit(np.arange(250, 400) * dt, speeds[250:400], 1)[0]
    expect = 9.81 * (np.sin(th) - mu * np.cos(th))
    assert abs(acc - expect) < 1e-6 * expect + 1e-9, (acc, expect)
    assert len(sim.contacts()) == 4


def test_fp32_refinement_error(gpu, solve_mode):
    spec = scenes.sphere_drape_small()
    s64 = gpu.Simulation(spec)
    s32 = gpu.Simulation(spec, precision=abi.FP32_REFINE64, refinement_sweeps=3)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(3 * spec.n_vertices)
    x64 = s64.solve(b)
    x32 = s32.solve(b)
    assert rel_err(x32, x64) < 1e-12


def test_fp32_refinement_error_40k(gpu, solve_mode):
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    s64 = gpu.Simulation(spec)
    s32 = gpu.Simulation(spec, precision=abi.FP32_REFINE64, refinement_sweeps=2)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(3 * spec.n_vertices)
    err = rel_err(s32.solve(b), s64.solve(b))
    print("fp32 + 2 refinement sweeps, max rel error vs fp64:", err)
    assert err < 1e-10


def pinned(probe, top_mode, fold):
    """probe's substitution schedule with the top block's mode and fold set, every level in a
    launch of its own (the fold runs in the first backward level kernel)."""
    base = re.sub(r";t=\d+;o=\d|;g=[\d,]+|;h=[\d,]+", "", probe.tuning())
    return "%s;t=%d;o=%d" % (base, top_mode, fold)


@pytest.mark.parametrize("top_mode", [0, 8])
@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_REFINE64])
def test_folded_top_finish_is_bit_identical(gpu, monkeypatch, precision, top_mode):
    """The top block's ordered finish run inside the first backward level sums the same partial
    products in the same order as the separate finish kernel: with the same substitution
    schedule, the two paths give identical states (symmetric tiles, t=0, and the split
    matvec over 8 column ranges)."""
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    probe = gpu.Simulation(spec, precision=precision, record_iterations=False)
    assert probe.info()["top_m"] > 0
    sims = []
    for fold in (1, 0):
        monkeypatch.setenv("PDFC_TUNE", pinned(probe, top_mode, fold))
        s = gpu.Simulation(spec, precision=precision, record_iterations=False)
        assert s.info()["top_fold"] == fold and s.info()["top_chunks"] == top_mode
        sims.append(s)
    a, b = sims
    a.step(5)
    b.step(5)
    qa, va = a.state()
    qb, vb = b.state()
    assert np.array_equal(qa, qb) and np.array_equal(va, vb)


def test_symmetric_tiles_match_split_matvec(gpu, monkeypatch):
    """The symmetric tile products (upper block triangle of S^-1, each block used for its rows
    and, transposed, for its columns) solve K x = b as the split matvec does, to rounding:
    S^-1 is symmetric up to the rounding of its unit solves."""
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    probe = gpu.Simulation(spec, record_iterations=False)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(3 * spec.n_vertices)
    xs = {}
    for mode in (0, 8):
        monkeypatch.setenv("PDFC_TUNE", pinned(probe, mode, 1))
        s = gpu.Simulation(spec, record_iterations=False)
        assert s.info()["top_chunks"] == mode
        xs[mode] = s.solve(b)
    err = rel_err(xs[0], xs[8])
    print("symmetric tiles vs split matvec, max rel difference:", err)
    assert err < 1e-12


def test_air_damping_identical_state_parity(gpu):
    """Air damping on (P5): the coefficient is zeroed at contact vertices (SimulationBase.cpp:
    347-355). 20x20 drape with damping 2e-3 (dt alpha / m = 0.08 at an interior vertex; the
    reference's explicit damping term diverges once that exceeds 1), every step from the
    oracle's state."""
    spec = scenes.sphere_drape_small()
    spec.air_damping = 2e-3
    sim = gpu.Simulation(spec)
    orc = po.Oracle(spec)
    worst = dict(q=0.0, v=0.0, r=0.0)
    n_contact_steps = 0
    for s in range(120):
        q0, v0 = orc.state()
        sim.set_state(q0, v0)
        orc.step(1)
        sim.step(1)
        q, v = sim.state()
        qo, vo = orc.state()
        worst["q"] = max(worst["q"], rel_err(q, qo))
        worst["v"] = max(worst["v"], rel_err(v, vo))
        ko, ro, _, _ = contact_arrays(orc.contacts())
        kg, rg, _, _ = contact_arrays(sim.contacts())
        assert np.array_equal(ko, kg), "contact sets differ at step %d" % s
        if len(ko):
            n_contact_steps += 1
            worst["r"] = max(worst["r"], rel_err(rg, ro))
    print("air damping 2e-3: worst relative errors", worst, "steps with contact", n_contact_steps)
    assert n_contact_steps > 50
    assert max(worst.values()) < REL, worst
    # the damping acts: the same steps without it differ
    plain = scenes.sphere_drape_small()
    o2 = po.Oracle(plain)
    o2.step(120)
    assert rel_err(o2.state()[1], orc.state()[1]) > 1e-3


def test_drape_40k_full_step_parity_in_self_contact_phase(gpu):
    """The 40k drape at full size, its whole step: advanced on the GPU to step 118, where the cloth
    on the sphere folds onto itself (thousands of obstacle contacts, hundreds of self contacts),
    then 3 steps from identical states against the oracle: q, v, obstacle and self reactions
    within 1e-9, identical contact sets, node-node pairs and levels (SimulationBase.cpp:103-173)."""
    spec = scenes.sphere_drape_40k()
    sim = gpu.Simulation(spec)
    print("40k drape: 118 GPU steps", flush=True)
    sim.step(118)
    orc = po.Oracle(spec)
    worst = dict(q=0.0, v=0.0, r=0.0, rs=0.0)
    for s in range(3):
        q0, v0 = sim.state()
        orc.set_state(q0, v0)
        sim.step(1)
        orc.step(1)
        q, v = sim.state()
        qo, vo = orc.state()
        worst["q"] = max(worst["q"], rel_err(q, qo))
        worst["v"] = max(worst["v"], rel_err(v, vo))
        ko, ro, no, ao = contact_arrays(orc.contacts())
        kg, rg, ng, ag = contact_arrays(sim.contacts())
        assert np.array_equal(ko, kg), "contact sets differ at step %d" % s
        assert np.array_equal(ao, ag), "take-off flags differ at step %d" % s
        worst["r"] = max(worst["r"], rel_err(rg, ro))
        so, sg = orc.self_contacts(), sim.self_contacts()
        assert [(c.vertex, c.partner, c.level) for c in so] == [(c.vertex, c.partner, c.level) for c in sg]
        worst["rs"] = max(worst["rs"], rel_err(np.array([list(c.r_local) for c in sg]),
                                               np.array([list(c.r_local) for c in so])))
        mo, _ = orc.iteration_stats()
        mg, _ = sim.iteration_stats()
        assert np.allclose(mg, mo, rtol=1e-6, atol=1e-12), (s, mg, mo)
        print("step", 119 + s, "contacts", len(ko), "self contacts", len(so),
              "levels", 1 + max(c.level for c in so), flush=True)
        assert len(ko) > 1000 and len(so) > 100
    print("worst relative errors", worst)
    assert max(worst.values()) < REL, worst


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_REFINE64])
def test_merged_levels_are_bit_identical(gpu, monkeypatch, precision):
    """Consecutive substitution levels merged into one launch (a workgroup per subtree group,
    levels separated by the workgroup barrier) compute every task as the level kernel does with
    the same waves per task: states after 5 steps of the 40k drape are identical for the
    level-by-level sweeps, two and three launches per sweep, and one launch per sweep."""
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    monkeypatch.setenv("PDFC_CROSS", "0")  # one level per bisection: enough levels for three ranges
    probe = gpu.Simulation(spec, precision=precision, record_iterations=False)
    L = probe.info()["n_levels_swept"]
    assert L >= 3
    waves = ",".join(["8"] * L)  # 2 slots per merged level: fits the merged launch's LDS everywhere
    parts = [",".join(["1"] * L), "1,%d" % (L - 1), "%d,1" % (L - 1), "1,1,%d" % (L - 2), str(L)]
    states = []
    for g in parts:
        h = ",".join(reversed(g.split(",")))
        monkeypatch.setenv("PDFC_TUNE", "f=%s;b=%s;t=0;o=0;g=%s;h=%s" % (waves, waves, g, h))
        s = gpu.Simulation(spec, precision=precision, record_iterations=False)
        I = s.info()
        assert I["launches_forward"] == len(g.split(",")) and I["launches_backward"] == len(h.split(","))
        print("g=%s h=%s: %d + %d launches per solve" % (g, h, I["launches_forward"], I["launches_backward"]),
              flush=True)
        s.step(5)
        states.append(s.state())
    q0, v0 = states[0]
    for (q, v), g in zip(states[1:], parts[1:]):
        assert np.array_equal(q, q0) and np.array_equal(v, v0), g


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_REFINE64])
def test_split_backward_reductions_match(gpu, monkeypatch, precision):
    """Backward reductions split over 2 or 4 workgroups per task (partials summed in part order by
    the next level, which also writes their final values) solve K x = b as the unsplit sweep
    does, to rounding, and 5 steps of the 40k drape agree to 1e-10."""
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    probe = gpu.Simulation(spec, precision=precision, record_iterations=False)
    L = probe.info()["n_levels_swept"]
    assert L >= 2
    base = re.sub(r";p=[\d,]+", "", probe.tuning())
    rng = np.random.default_rng(12)
    b = rng.standard_normal(3 * spec.n_vertices)
    out = {}
    mixed = ([4, 1, 2][:L - 1] + [1] * L)[:L - 1] + [1]  # the last (leaf) level is never split
    for name, parts in (("whole", [1] * L), ("two", [2] * (L - 1) + [1]), ("four", [4] * (L - 1) + [1]),
                        ("mixed", mixed)):
        monkeypatch.setenv("PDFC_TUNE", base + ";p=" + ",".join(map(str, parts)))
        s = gpu.Simulation(spec, precision=precision, record_iterations=False)
        assert s.info()["split_backward"][:L] == parts
        x = s.solve(b)
        s.step(5)
        out[name] = (x, s.state())
    xw, (qw, vw) = out["whole"]
    for name in ("two", "four", "mixed"):
        x, (q, v) = out[name]
        print(name, "solve rel diff %.3e, q %.3e, v %.3e" % (rel_err(x, xw), rel_err(q, qw), rel_err(v, vw)), flush=True)
        assert rel_err(x, xw) < 1e-12
        assert rel_err(q, qw) < 1e-10 and rel_err(v, vw) < 1e-10


def test_pregathered_tiles_are_bit_identical(gpu, monkeypatch):
    """The symmetric tiles reading b_T from a gather launch of its own compute exactly what the
    tiles gathering b_T themselves compute (same pull order): identical states after 5 steps."""
    spec = scenes.sphere_drape_40k()
    spec.self_collision = False
    probe = gpu.Simulation(spec, record_iterations=False)
    base = re.sub(r";t=\d+;o=\d;q=\d", "", probe.tuning())
    states = []
    for q in (0, 1):
        monkeypatch.setenv("PDFC_TUNE", "%s;t=0;o=0;q=%d" % (base, q))
        s = gpu.Simulation(spec, record_iterations=False)
        assert s.info()["top_pregather"] == q
        s.step(5)
        states.append(s.state())
    assert np.array_equal(states[0][0], states[1][0]) and np.array_equal(states[0][1], states[1][1])
