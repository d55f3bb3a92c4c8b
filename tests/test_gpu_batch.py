"""GPU: batched scenes (one block-diagonal system, one kernel sequence per step) and
the HIP-graph replay of a step.

Batched parity is per scene against the oracle run on that scene alone, from
identical states, to the FP64 bar (1e-9 relative, identical contact sets); the
batch also agrees with each scene run alone on the GPU (1e-10: only the summation
order of the autotuned substitution differs). A replayed graph runs the
same kernels with the same arguments as the eager step, so its states are
bit-identical to the eager ones; replays are interleaved with eager work on the
same stream and checked one by one.
"""
import numpy as np
import pytest

from climatenet_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

REL = 1e-9


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def small_batch():
    return scenes.batched_instances(count=5, seed=3, sizes=(6, 16))


def test_batch_matches_oracle_per_scene(gpu):
    specs = small_batch()
    sim = gpu.BatchSimulation(specs)
    orcs = [po.Oracle(sp) for sp in specs]
    for o in orcs:
        o.step(4)  # a few steps in; contact starts within the checked window
    n_contact = 0
    for s in range(40):
        Q, Vv = [], []
        for o in orcs:
            q, v = o.state()
            Q.append(q)
            Vv.append(v)
        sim.set_state(np.concatenate(Q), np.concatenate(Vv))
        sim.step(1)
        for o in orcs:
            o.step(1)
        q, v = sim.state()
        cs = sim.contacts()
        for k, (o, (a, b)) in enumerate(zip(orcs, sim.ranges)):
            qo, vo = o.state()
            assert rel_err(q[3 * a:3 * b], qo) < REL, (s, k)
            assert rel_err(v[3 * a:3 * b], vo) < REL, (s, k)
            mine = [(c.vertex - a, list(c.r_local)) for c in cs if c.obstacle == k]
            assert all(a <= c.vertex < b for c in cs if c.obstacle == k)
            ref = [(c.vertex, list(c.r_local)) for c in o.contacts()]
            assert [m[0] for m in mine] == [r[0] for r in ref], (s, k)
            if ref:
                n_contact += 1
                assert rel_err(np.array([m[1] for m in mine]), np.array([r[1] for r in ref])) < REL
    assert n_contact > 50


def test_batch_equals_scenes_run_alone(gpu):
    specs = small_batch()
    sim = gpu.BatchSimulation(specs)
    sim.step(30)
    q, v = sim.state()
    for sp, (a, b) in zip(specs, sim.ranges):
        one = gpu.Simulation(sp)
        one.step(30)
        q1, v1 = one.state()
        # same supernodes and operators; the substitution's per-level wave split (autotuned
        # separately for the batch and for each scene) changes the summation order only
        assert rel_err(q[3 * a:3 * b], q1) < 1e-10
        assert rel_err(v[3 * a:3 * b], v1) < 1e-10


def test_graph_replay_is_bit_identical_to_eager(gpu, monkeypatch):
    spec = scenes.sphere_drape_small()
    eager = gpu.Simulation(spec, record_iterations=False)
    monkeypatch.setenv("PDFC_TUNE", eager.tuning())  # the same summation order
    graph = gpu.Simulation(spec, record_iterations=False, use_graph=True)
    rng = np.random.default_rng(0)
    for s in range(30):
        eager.step(1)
        graph.step(1)
        qe, ve = eager.state()
        qg, vg = graph.state()
        assert np.array_equal(qe, qg) and np.array_equal(ve, vg), s
        if s % 5 == 2:  # eager work on the graph's stream between replays
            graph.solve(rng.standard_normal(3 * spec.n_vertices))
    assert graph.phase_stats()["iterations"] == spec.iterations


def test_graph_replay_of_a_batch(gpu, monkeypatch):
    specs = small_batch()
    eager = gpu.BatchSimulation(specs, record_iterations=False)
    monkeypatch.setenv("PDFC_TUNE", eager.tuning())  # the same summation order
    graph = gpu.BatchSimulation(specs, record_iterations=False, use_graph=True)
    for s in range(3):
        eager.step(7)
        graph.step(7)
        qe, _ = eager.state()
        qg, _ = graph.state()
        assert np.array_equal(qe, qg), s


def test_batch_of_boxes_rotating_spheres_and_attachments(gpu):
    """A batch of a box drape, a rotating-sphere drape (turning from step 6) and a cloth hanging
    from two attachment constraints: per scene, per step, from identical states, the oracle's
    states and contacts (with obstacle speeds) to the FP64 bar."""
    specs = [scenes.box_drape(), scenes.rotating_sphere_drape(start=0.05), scenes.hanging_cloth()]
    sim = gpu.BatchSimulation(specs)
    orcs = [po.Oracle(sp) for sp in specs]
    seen_speed = False
    n_contact = 0
    for s in range(60):
        Q, Vv = zip(*[o.state() for o in orcs])
        sim.set_state(np.concatenate(Q), np.concatenate(Vv))
        sim.step(1)
        for o in orcs:
            o.step(1)
        q, v = sim.state()
        cs = sim.contacts()
        for k, (o, (a, b)) in enumerate(zip(orcs, sim.ranges)):
            qo, vo = o.state()
            assert rel_err(q[3 * a:3 * b], qo) < REL, (s, k)
            assert rel_err(v[3 * a:3 * b], vo) < REL, (s, k)
            mine = [c for c in cs if c.obstacle == k]
            ref = o.contacts()
            assert [c.vertex - a for c in mine] == [c.vertex for c in ref], (s, k)
            if ref:
                n_contact += 1
                sp_m = np.array([list(c.speed) for c in mine])
                sp_r = np.array([list(c.speed) for c in ref])
                assert np.abs(sp_m - sp_r).max() < 1e-14, (s, k)
                seen_speed = seen_speed or np.abs(sp_r).max() > 0
                assert rel_err(np.array([list(c.r_local) for c in mine]),
                               np.array([list(c.r_local) for c in ref])) < REL, (s, k)
    assert n_contact > 60 and seen_speed


def test_batch_64_full_size_parity(gpu):
    """The batched configuration at full size: 64 cloths of 32^2..128^2 quads (517,876 vertices) in
    one batch, advanced 30 steps (every cloth on its sphere), then 3 steps from identical states,
    each scene against the oracle run on that scene alone (SimulationSpeed.cpp:171-187 per
    instance): q, v and reactions within 1e-9, identical contact sets."""
    specs = scenes.batched_instances(64, seed=0)
    sim = gpu.BatchSimulation(specs)
    print("64 cloths: 30 batched steps", flush=True)
    sim.step(30)
    orcs = [po.Oracle(sp) for sp in specs]
    worst = dict(q=0.0, v=0.0, r=0.0)
    in_contact = 0
    for s in range(3):
        q, v = sim.state()
        for o, (a, b) in zip(orcs, sim.ranges):
            o.set_state(q[3 * a:3 * b].copy(), v[3 * a:3 * b].copy())
            o.step(1)
        sim.step(1)
        q, v = sim.state()
        cs = sim.contacts()
        for k, (o, (a, b)) in enumerate(zip(orcs, sim.ranges)):
            qo, vo = o.state()
            worst["q"] = max(worst["q"], rel_err(q[3 * a:3 * b], qo))
            worst["v"] = max(worst["v"], rel_err(v[3 * a:3 * b], vo))
            mine = [(c.vertex - a, list(c.r_local)) for c in cs if c.obstacle == k]
            ref = [(c.vertex, list(c.r_local)) for c in o.contacts()]
            assert [m[0] for m in mine] == [r[0] for r in ref], (s, k)
            if ref:
                in_contact += 1
                worst["r"] = max(worst["r"], rel_err(np.array([m[1] for m in mine]),
                                                     np.array([r[1] for r in ref])))
    print("64 cloths: worst relative errors", worst, "scene-steps in contact", in_contact)
    assert in_contact == 3 * 64
    assert max(worst.values()) < REL, worst
