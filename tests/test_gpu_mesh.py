"""GPU: triangle-mesh obstacles (StaticMesh, MovingMesh) against the oracle.

The 20x20 drape falls on a triangle-mesh sphere (icosphere, 1,280 faces), static or
moving through key frames (turning and rising: non-zero point speeds). Every step starts
from the oracle's state; the contact sets (vertex, obstacle) are identical, and the
contact points, normals and speeds agree to rounding. Positions, velocities and
reactions agree within 1e-9 relative.
"""
import numpy as np
import pytest

from climatenet_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

REL = 1e-9


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("moving", [False, True])
def test_mesh_obstacle_identical_state_parity(gpu, moving):
    spec = scenes.mesh_sphere_drape(moving=moving)
    sim = gpu.Simulation(spec)
    orc = po.Oracle(spec)
    worst = dict(q=0.0, v=0.0, r=0.0, point=0.0, speed=0.0)
    n_contact = 0
    max_speed = 0.0
    for s in range(120):
        q0, v0 = orc.state()
        sim.set_state(q0, v0)
        orc.step(1)
        sim.step(1)
        q, v = sim.state()
        qo, vo = orc.state()
        worst["q"] = max(worst["q"], rel_err(q, qo))
        worst["v"] = max(worst["v"], rel_err(v, vo))
        cg, co = sim.contacts(), orc.contacts()
        assert [(c.obstacle, c.vertex) for c in cg] == [(c.obstacle, c.vertex) for c in co], s
        if co:
            n_contact += 1
            A = lambda cs, f: np.array([list(getattr(c, f)) for c in cs])
            worst["r"] = max(worst["r"], rel_err(A(cg, "r_local"), A(co, "r_local")))
            worst["point"] = max(worst["point"], np.abs(A(cg, "point") - A(co, "point")).max())
            assert np.abs(A(cg, "normal") - A(co, "normal")).max() < 1e-14
            sp = A(co, "speed")
            max_speed = max(max_speed, np.abs(sp).max())
            worst["speed"] = max(worst["speed"], np.abs(A(cg, "speed") - sp).max())
    print("worst:", worst, "contact steps:", n_contact, "max obstacle speed:", max_speed)
    assert n_contact > 60
    assert worst["q"] < REL and worst["v"] < REL and worst["r"] < REL, worst
    assert worst["point"] < 1e-15 and worst["speed"] < 1e-14
    if moving:
        assert max_speed > 0.1


def test_analytic_incline_on_a_mesh_support(gpu):
    """The reference's closed-form sliding example (45-degree support as a two-triangle mesh
    obstacle, threshold 1e-3, mu 0.1): on the GPU the cloth's speed equals the closed form at
    every one of 800 steps, before and after the impact."""
    from tests.test_oracle_mesh import analytic_speed
    spec = scenes.analytic_incline_mesh()
    sim = gpu.Simulation(spec)
    down = np.array([0.0, np.cos(np.pi / 4), -np.sin(np.pi / 4)])
    worst = 0.0
    for s in range(800):
        sim.step(1)
        _, v = sim.state()
        V = v.reshape(-1, 3).mean(axis=0)
        t = (s + 1) * spec.time_step
        sp = V @ down if sim.contacts() else -V[2]
        worst = max(worst, abs(sp - analytic_speed(t)) / analytic_speed(t))
    print("worst relative deviation from the closed form", worst)
    assert len(sim.contacts()) == 4
    assert worst < 1e-10, worst


def test_reference_shaped_moving_mesh_drape_parity(gpu):
    """The scene shaped like the reference's replicability drapes (5,929-vertex cloth on a
    200-frame, 5,120-face key-framed sphere mesh, tolerance 0.015, plus a plane): from the
    oracle's state after 36 steps, 8 steps with hundreds of mesh contacts agree (contact sets
    identical; positions, velocities and reactions within 1e-9). Exercises the device-built
    grid with buckets of tens of triangles and the 16-lane query."""
    spec = scenes.rotating_mesh_drape()
    sim = gpu.Simulation(spec)
    orc = po.Oracle(spec)
    orc.step(36)
    sim.step(36)  # the mesh clocks advance together; the states are synchronised below
    worst = dict(q=0.0, v=0.0, r=0.0, speed=0.0)
    n_contacts = []
    for s in range(8):
        q0, v0 = orc.state()
        sim.set_state(q0, v0)
        orc.step(1)
        sim.step(1)
        q, v = sim.state()
        qo, vo = orc.state()
        worst["q"] = max(worst["q"], rel_err(q, qo))
        worst["v"] = max(worst["v"], rel_err(v, vo))
        cg, co = sim.contacts(), orc.contacts()
        assert [(c.obstacle, c.vertex) for c in cg] == [(c.obstacle, c.vertex) for c in co], s
        A = lambda cs, f: np.array([list(getattr(c, f)) for c in cs])
        worst["r"] = max(worst["r"], rel_err(A(cg, "r_local"), A(co, "r_local")))
        worst["speed"] = max(worst["speed"], np.abs(A(cg, "speed") - A(co, "speed")).max())
        n_contacts.append(len(co))
    print("worst:", worst, "contacts per step:", n_contacts)
    assert min(n_contacts) > 100
    assert worst["q"] < REL and worst["v"] < REL and worst["r"] < REL, worst
    assert worst["speed"] < 1e-14


def test_moving_mesh_with_self_collision_parity(gpu):
    """A folded cloth (two layers 4 mm apart) falls on the moving mesh sphere with self
    collision on: from identical states, the mesh contacts and the self-contact pairs with
    their computation levels are identical and positions and velocities agree within 1e-9,
    over 60 steps in which both kinds of contact occur."""
    from tests.test_gpu_self import folded_state
    spec = scenes.mesh_sphere_drape(moving=True)
    spec.self_collision = True
    spec.self_tolerance = 5e-3
    q = folded_state(spec, 0.004)
    sim = gpu.Simulation(spec)
    orc = po.Oracle(spec)
    sim.set_state(q, np.zeros_like(q))
    orc.set_state(q, np.zeros_like(q))
    worst = dict(q=0.0, v=0.0)
    both = 0
    for s in range(60):
        q0, v0 = orc.state()
        sim.set_state(q0, v0)
        orc.step(1)
        sim.step(1)
        q, v = sim.state()
        qo, vo = orc.state()
        worst["q"] = max(worst["q"], rel_err(q, qo))
        worst["v"] = max(worst["v"], rel_err(v, vo))
        cg, co = sim.contacts(), orc.contacts()
        assert [(c.obstacle, c.vertex) for c in cg] == [(c.obstacle, c.vertex) for c in co], s
        sg = [(c.vertex, c.partner, tuple(c.face), c.level) for c in sim.self_contacts()]
        so = [(c.vertex, c.partner, tuple(c.face), c.level) for c in orc.self_contacts()]
        assert sg == so, s
        both += bool(co) and bool(so)
    print("worst:", worst, "steps with mesh and self contacts:", both)
    assert both >= 20
    assert worst["q"] < REL and worst["v"] < REL, worst
