"""Pins of the oracle's triangle-mesh obstacles (StaticMesh / MovingMesh).

* A flat triangulated obstacle: contacts are exactly the vertices within the tolerance
  of it, at their projections, with the winding's normal.
* Near an icosphere: the chosen contact point is the closest point of the whole mesh
  (brute force over all faces with a plain closest-point routine).
* A key-framed mesh translating at constant velocity: the contact speed is that
  velocity and the contact point follows the mesh as the clock advances.
"""
import numpy as np

from climatenet_amd import scenes
from oracle import pyoracle as po


def flat_mesh(n=4, side=2.0, z=0.0):
    P, T = scenes.square_grid(n, side, z)
    return P.reshape(-1, 3), T.reshape(-1, 3)


def make_spec(points, obstacle):
    """A probe 'cloth': each probe point is the corner of its own triangle whose other two
    corners sit far away (z = 50), so every probe vertex has mass and only the probes can
    meet the obstacle."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    far = []
    tris = []
    for i in range(n):
        far += [[50.0 + i, 50.0, 50.0], [50.0 + i, 50.1, 50.0]]
        tris.append([i, n + 2 * i, n + 2 * i + 1])
    rest = np.vstack([pts, np.array(far)])
    return scenes.SceneSpec("mesh_probe", rest.ravel(), np.array(tris, dtype=np.int32).ravel(), None, [obstacle],
                            np.array([[0.2]]), time_step=0.01, iterations=1)


def plain_closest(p, a, b, c):
    """Closest point of triangle abc to p by projection and edge clamping (independent of
    the oracle's Ericson routine)."""
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # inside test with barycentric coordinates
    v0, v1, v2 = b - a, c - a, q - a
    d00, d01, d11, d20, d21 = v0 @ v0, v0 @ v1, v1 @ v1, v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    if v >= 0 and w >= 0 and v + w <= 1:
        return q
    best = None
    for s, e in ((a, b), (b, c), (c, a)):
        t = np.clip(np.dot(p - s, e - s) / np.dot(e - s, e - s), 0.0, 1.0)
        r = s + t * (e - s)
        if best is None or np.sum((r - p) ** 2) < np.sum((best - p) ** 2):
            best = r
    return best


def test_flat_mesh_contacts_closed_form():
    P, T = flat_mesh()
    ob = scenes.ObstacleSpec("mesh", tolerance=1e-2, frames=P[None], triangles=T)
    rng = np.random.default_rng(0)
    pts = np.column_stack([rng.uniform(-0.9, 0.9, 40), rng.uniform(-0.9, 0.9, 40), rng.uniform(-0.02, 0.02, 40)])
    spec = make_spec(pts, ob)
    orc = po.Oracle(spec)
    orc.detect()
    cs = orc.contacts()
    got = {c.vertex: c for c in cs}
    for i, p in enumerate(pts):
        if abs(p[2]) < 1e-2:
            assert i in got, (i, p)
            c = got[i]
            assert np.allclose(list(c.point), [p[0], p[1], 0.0], atol=1e-15)
            assert np.allclose(list(c.normal), [0, 0, 1], atol=1e-15)
            assert list(c.speed) == [0.0, 0.0, 0.0]
        else:
            assert i not in got, (i, p)


def test_icosphere_contact_is_the_closest_point():
    P, T = scenes.icosphere(0.25, subdivisions=2)
    ob = scenes.ObstacleSpec("mesh", tolerance=1e-2, frames=P[None], triangles=T)
    rng = np.random.default_rng(1)
    d = rng.normal(size=(60, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * (0.25 + rng.uniform(-0.004, 0.008, 60))[:, None]
    spec = make_spec(pts, ob)
    orc = po.Oracle(spec)
    orc.detect()
    got = {c.vertex: c for c in orc.contacts()}
    checked = 0
    for i, p in enumerate(pts):
        cands = [plain_closest(p, P[a], P[b], P[c]) for a, b, c in T]
        dist = [np.linalg.norm(q - p) for q in cands]
        k = int(np.argmin(dist))
        if dist[k] < 1e-2 - 1e-9:
            assert i in got
            assert np.linalg.norm(np.array(list(got[i].point)) - cands[k]) < 1e-12
            checked += 1
        elif dist[k] > 1e-2 + 1e-9:
            assert i not in got
    assert checked > 40


def test_moving_mesh_translation_speed_and_clock():
    P, T = flat_mesh()
    u = np.array([0.3, -0.2, 0.05])
    dt_frame = 0.035
    frames = np.array([P + k * dt_frame * u for k in range(6)])
    times = np.array([k * dt_frame for k in range(6)])
    ob = scenes.ObstacleSpec("mesh", tolerance=1e-2, frames=frames, triangles=T, times=times)
    pts = np.array([[0.1, 0.2, 0.004], [-0.3, 0.4, -0.006], [0.5, -0.5, 0.5]])
    spec = make_spec(pts, ob)
    spec.gravity = (0.0, 0.0, 0.0)
    orc = po.Oracle(spec)
    orc.detect()
    cs = orc.contacts()
    assert sorted(c.vertex for c in cs) == [0, 1]
    for c in cs:
        assert np.allclose(list(c.speed), u, rtol=1e-12, atol=1e-15)
    # after k steps of dt = 0.01 the mesh sits at P + t u while t <= the last key frame
    for k in range(1, 16):
        orc.step(1)
        q, _ = orc.state()
        t = 0.01 * k
        orc.set_state(np.concatenate([(P[0] + t * u + np.array([0.0, 0.0, 0.003]))[None],
                                      q.reshape(-1, 3)[1:]]).ravel(), None)
        orc.detect()
        c0 = [c for c in orc.contacts() if c.vertex == 0]
        assert len(c0) == 1
        # the vertex sits 3 mm above the moved mesh's first vertex position: its contact
        # point is the moved mesh point right below
        assert np.allclose(list(c0[0].point), P[0] + t * u, atol=1e-12), (k, list(c0[0].point), P[0] + t * u)


def analytic_speed(t, g=9.81, th=np.pi / 4, mu=0.1, h=0.2):
    """replicability/analytic_example_data_processing.py:53-112, restated: free fall until
    t* = sqrt(2h/g), then speed along the support g t* (cos th - mu sin th) + g (sin th - mu cos th)(t - t*)."""
    ts = np.sqrt(2 * h / g)
    if t < ts:
        return g * t
    return g * ts * (np.cos(th) - np.sin(th) * mu) + g * (np.sin(th) - np.cos(th) * mu) * (t - ts)


def test_analytic_incline_on_a_mesh_support():
    """The reference's closed-form sliding example with its support as a triangle-mesh
    obstacle (threshold 1e-3): the cloth's speed (vertical before impact, along the support
    after) equals the closed form at every step."""
    spec = scenes.analytic_incline_mesh()
    orc = po.Oracle(spec)
    down = np.array([0.0, np.cos(np.pi / 4), -np.sin(np.pi / 4)])
    worst = 0.0
    for s in range(800):
        orc.step(1)
        _, v = orc.state()
        V = v.reshape(-1, 3).mean(axis=0)
        t = (s + 1) * spec.time_step
        sp = V @ down if orc.contacts() else -V[2]
        worst = max(worst, abs(sp - analytic_speed(t)) / analytic_speed(t))
    assert len(orc.contacts()) == 4
    assert worst < 1e-10, worst


def test_reference_shaped_scene_frames():
    """The moving-mesh bench scene: 5,929 cloth vertices; 200 key frames of a 2,562-vertex,
    5,120-face sphere of radius 0.5 about (0, 0, -0.52), 0.035 s apart, turning rigidly about
    z by 60 degrees in all (each frame a rotation of frame 0)."""
    spec = scenes.rotating_mesh_drape()
    o = spec.obstacles[0]
    assert spec.n_vertices == 5929 and o.frames.shape == (200, 2562, 3)
    assert o.triangles.reshape(-1, 3).shape == (5120, 3)
    assert np.allclose(np.diff(o.times), 0.035) and o.times[0] == 0.0
    c = np.array([0.0, 0.0, -0.52])
    r = np.linalg.norm(o.frames - c, axis=2)
    assert np.abs(r - 0.5).max() < 1e-12
    # frame k = R_z(60 deg * k / 199) frame 0, checked through the angle of every vertex
    d0, d1 = o.frames[0] - c, o.frames[199] - c
    moved = np.hypot(d0[:, 0], d0[:, 1]) > 1e-3
    ang = np.arctan2(d1[moved, 1], d1[moved, 0]) - np.arctan2(d0[moved, 1], d0[moved, 0])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(ang - np.deg2rad(60.0)).max() < 1e-12
    assert np.abs(d1[:, 2] - d0[:, 2]).max() < 1e-15
