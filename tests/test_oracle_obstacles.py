"""Pins of the oracle's axis-aligned boxes, rotating spheres and attachment constraints.

* Box: a vertex inside the box touches the nearest face, at its projection on that face,
  with that face's outward normal; a vertex outside has no contact; at the centre of a
  cube every face is equally near and the Side enum's first (TOP, +y) wins.
* Rotating sphere: the contact speed is omega x (contact - centre) with omega = speed *
  unit axis, zero until the rotation's start time has been stepped through; a vertex held
  on the sphere by gravity with a large friction coefficient sticks to the moving surface
  (relative tangential velocity -> 0).
* Attachments: with springs and bending switched off and no gravity, every attached vertex
  follows implicit Euler of a linear spring of stiffness w^2 to its anchor,
  v1 = (m v0 + dt w^2 (P - q0)) / (m + dt^2 w^2), exactly, and K gains w^2 dt^2 on the
  attached vertices' diagonal entries.
"""
import numpy as np

from climatenet_amd import scenes
from oracle import pyoracle as po
from tests.test_oracle_mesh import make_spec


def nearest_face(p, lo, hi):
    """Brute force: distance to each of the six face planes of a point inside [lo, hi]."""
    faces = []
    for k in range(3):
        faces.append((p[k] - lo[k], k, -1.0))
        faces.append((hi[k] - p[k], k, 1.0))
    return faces


def test_box_contact_is_nearest_face_projection():
    lo = np.array([-0.3, -0.2, -0.1])
    size = np.array([0.6, 0.5, 0.4])
    hi = lo + size
    ob = scenes.ObstacleSpec("box", tuple(lo), size=tuple(size))
    rng = np.random.default_rng(3)
    inside = lo + rng.uniform(0.02, 0.98, (50, 3)) * size
    outside = np.array([[0.0, 0.0, hi[2] + 1e-3], [lo[0] - 1e-3, 0.0, 0.0], [0.0, hi[1] + 0.05, 0.0]])
    pts = np.vstack([inside, outside])
    spec = make_spec(pts, ob)
    orc = po.Oracle(spec)
    orc.detect()
    got = {c.vertex: c for c in orc.contacts()}
    assert set(got) == set(range(len(inside)))
    for i, p in enumerate(inside):
        d, k, sgn = min(nearest_face(p, lo, hi), key=lambda t: t[0])
        want = p.copy()
        want[k] = hi[k] if sgn > 0 else lo[k]
        n = np.zeros(3)
        n[k] = sgn
        assert np.abs(np.array(list(got[i].point)) - want).max() < 1e-15
        assert list(got[i].normal) == list(n)


def test_box_tie_goes_to_top_side():
    ob = scenes.ObstacleSpec("box", (-0.5, -0.5, -0.5), size=(1.0, 1.0, 1.0))
    spec = make_spec([[0.0, 0.0, 0.0], [0.2, 0.0, 0.2]], ob)
    orc = po.Oracle(spec)
    orc.detect()
    cs = {c.vertex: c for c in orc.contacts()}
    # centre: all six faces at 0.5 -> TOP (+y); (0.2, 0, 0.2): RIGHT (+x) and FRONT (+z)
    # tie at 0.3 -> RIGHT comes first in the enum (TOP, BOTTOM, LEFT, RIGHT, FRONT, BACK)
    # (the point goes through relative coordinates and back: rounding only)
    assert list(cs[0].normal) == [0.0, 1.0, 0.0] and np.allclose(list(cs[0].point), [0.0, 0.5, 0.0], atol=1e-16)
    assert list(cs[1].normal) == [1.0, 0.0, 0.0] and np.allclose(list(cs[1].point), [0.5, 0.0, 0.2], atol=1e-16)


def test_rotating_sphere_speed_and_start():
    c = np.array([0.1, -0.2, 0.05])
    R, w = 0.3, 2.5
    axis = np.array([1.0, 2.0, 2.0])
    ob = scenes.ObstacleSpec("rotating_sphere", tuple(c), radius=R, axis=tuple(axis), angular_speed=w,
                             rotation_start=0.025)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(20, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = c + d * (R - 0.005)
    spec = make_spec(pts, ob)
    spec.gravity = (0.0, 0.0, 0.0)
    orc = po.Oracle(spec)
    omega = w * axis / np.linalg.norm(axis)
    # the clock: 0.025, 0.015, 0.005, then -0.005 at the fourth step's detection
    for k in range(5):
        orc.set_state(np.vstack([pts, spec.rest.reshape(-1, 3)[len(pts):]]).ravel(), None)
        orc.detect()
        cs = orc.contacts()
        assert len(cs) == len(pts)
        for cc in cs:
            sp = np.array(list(cc.speed))
            if k < 3:
                assert not sp.any()
            else:
                want = np.cross(omega, np.array(list(cc.point)) - c)
                assert np.abs(sp - want).max() < 1e-15
                # tangent to the sphere, |speed| = |omega| R sin(angle to the axis)
                assert abs(sp @ np.array(list(cc.normal))) < 1e-14
        orc.step(1)


def test_sticking_on_rotating_sphere_follows_the_surface():
    """A vertex pressed onto a turning sphere with mu = 50 sticks: after the step its
    velocity relative to the surface point has no tangential part."""
    c = np.zeros(3)
    R = 0.25
    ob = scenes.ObstacleSpec("rotating_sphere", tuple(c), radius=R, axis=(0.0, 0.0, 1.0), angular_speed=4.0,
                             rotation_start=-1.0, material=0)
    p = np.array([[0.05, 0.0, R - 0.002]])
    spec = make_spec(p, ob)
    spec.friction = np.array([[50.0]])
    spec.stretching = 1e-6
    spec.bending = 0.0
    spec.iterations = 300
    orc = po.Oracle(spec)
    orc.step(1)
    q, v = orc.state()
    cs = [cc for cc in orc.contacts() if cc.vertex == 0]
    assert len(cs) == 1 and cs[0].active
    n = np.array(list(cs[0].normal))
    rel = v.reshape(-1, 3)[0] - np.array(list(cs[0].speed))
    tang = rel - (rel @ n) * n
    speed = np.linalg.norm(list(cs[0].speed))
    assert speed > 0.1
    assert np.linalg.norm(tang) < 1e-6 * speed


def test_attachment_is_implicit_euler_of_a_linear_spring():
    N = 6
    rest, tris = scenes.square_grid(N, 1.0, 0.0)
    P = rest.reshape(-1, 3)
    idx = np.array([0, 3, 17, 40], dtype=np.int32)
    rng = np.random.default_rng(5)
    anchors = P[idx] + rng.uniform(-0.1, 0.1, (len(idx), 3))
    wts = np.array([10.0, 50.0, 3.0, 200.0])
    spec = scenes.SceneSpec("attach", rest, tris, None, [], np.array([[0.0]]), stretching=0.0, bending=0.0,
                            time_step=0.01, iterations=3, gravity=(0.0, 0.0, 0.0),
                            attachments=(idx, anchors, wts))
    orc = po.Oracle(spec)
    v0 = rng.normal(size=P.shape) * 0.1
    orc.set_state(rest, v0.ravel())
    m = orc.masses()
    K = orc.K()
    dt = spec.time_step
    for a, i in enumerate(idx):
        assert abs(K[i, i] - (m[i] + dt * dt * wts[a] ** 2)) <= 1e-15 * K[i, i]
    orc.step(1)
    q1, v1 = orc.state()
    V1 = v1.reshape(-1, 3)
    for a, i in enumerate(idx):
        want = (m[i] * v0[i] + dt * wts[a] ** 2 * (anchors[a] - P[i])) / (m[i] + dt * dt * wts[a] ** 2)
        assert np.abs(V1[i] - want).max() < 1e-13 * max(1.0, np.abs(want).max())
    free = np.setdiff1d(np.arange(len(P)), idx)
    assert np.abs(V1[free] - v0[free]).max() < 1e-15
