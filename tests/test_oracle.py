"""Pins of the CPU oracle (no GPU).

The reference ships no tests or golden vectors, so each oracle function is
checked against what mathematics fixes: closed forms, invariants, an
independent scipy assembly/solve, and brute force.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from climatenet_amd import scenes
from oracle import pyoracle as po

SQ2 = np.sqrt(2.0)


def grid_spec(N=6, side=1.0, z=0.3, **kw):
    rest, tris = scenes.square_grid(N, side, z)
    obs = kw.pop("obstacles", [scenes.ObstacleSpec("sphere", (0.0, 0.0, -10.0), radius=0.1)])
    return scenes.SceneSpec("grid", rest, tris, None, obs, np.array([[0.3]]), **kw)


def test_mt19937_64_standard_value():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937_64 (seed 5489)
    g = scenes.MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_bending_coefficients_regular_grid_closed_form():
    """Flat grid, spacing h, diagonals SW-NE: S,N -> 2(sqrt2-1)/h; SW,NE -> 1/h; W,E -> sqrt2/h.
    The W/E versus S/N asymmetry is the reference's second angle taken at the previous
    ring vertex (geometry_util.cpp:690-704), so a ring-direction or angle-vertex mistake fails."""
    N, side = 6, 1.2
    h = side / N
    o = po.Oracle(grid_spec(N, side))
    rest = o.spec.rest.reshape(-1, 3)
    expect = {(0, -1): 2 * (SQ2 - 1), (-1, -1): 1.0, (-1, 0): SQ2, (0, 1): 2 * (SQ2 - 1), (1, 1): 1.0, (1, 0): SQ2}
    nb = o.sizes()["B"]
    assert nb == (N - 1) ** 2
    for b in range(nb):
        idx, cf = o.bending_row(b)
        c = idx[0]
        assert abs(cf[0] - (-cf[1:]).sum()) < 1e-12 * abs(cf[0])
        for j, coef in zip(idx[1:], cf[1:]):
            off = tuple(np.round((rest[j, :2] - rest[c, :2]) / h).astype(int))
            assert abs(-coef - expect[off] / h) < 1e-6 / h, (off, -coef * h, expect[off])


def test_bending_ring_direction_matters_on_jittered_ring():
    """On an irregular ring the quirk makes the coefficients depend on the ring direction."""
    rng = np.random.default_rng(0)
    ang = np.deg2rad([0, 55, 130, 180, 250, 300]) + rng.uniform(-0.1, 0.1, 6)
    rad = 1.0 + rng.uniform(-0.2, 0.2, 6)
    ring = np.stack([rad * np.cos(ang), rad * np.sin(ang), np.zeros(6)], axis=1)
    c = np.zeros(3)
    fwd = po.mean_value_coefficients(c, ring)
    rev = po.mean_value_coefficients(c, ring[::-1].copy())[::-1]
    assert np.abs(fwd - rev).max() > 1e-3
    # tan(theta/2) identity for the angle at the centre: with a symmetric regular hexagon
    hexr = np.stack([np.cos(np.arange(6) * np.pi / 3), np.sin(np.arange(6) * np.pi / 3), np.zeros(6)], axis=1)
    w = po.mean_value_coefficients(c, hexr)
    # centre angle 60 deg -> tan 30; angle at the previous ring vertex between v_i and c is 60 deg too
    assert np.allclose(w, 2 * np.tan(np.pi / 6), rtol=1e-14)


def test_masses_and_rest_lengths():
    N, side = 8, 2.0
    spec = grid_spec(N, side, area_density=0.37)
    o = po.Oracle(spec)
    m = o.masses()
    assert abs(m.sum() - 0.37 * side * side) < 1e-12
    h = side / N
    ij, w, L0 = o.springs()
    assert len(L0) == 2 * N * (N + 1) + N * N
    assert np.all((np.abs(L0 - h) < 1e-6) | (np.abs(L0 - h * SQ2) < 1e-6))
    assert np.allclose(w, np.sqrt(spec.stretching))


def assemble_K_independently(o, spec):
    """diag(m) + dt^2 A^T A from the oracle's rows, assembled with scipy."""
    n = spec.n_vertices
    ij, w, L0 = o.springs()
    E = len(w)
    rows, cols, vals = [], [], []
    for e in range(E):
        rows += [e, e]
        cols += [ij[e, 0], ij[e, 1]]
        vals += [w[e], -w[e]]
    for b in range(o.sizes()["B"]):
        idx, cf = o.bending_row(b)
        rows += [E + b] * len(idx)
        cols += list(idx)
        vals += list(np.sqrt(spec.bending) * cf)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(E + o.sizes()["B"], n))
    return sp.diags(o.masses()) + spec.time_step ** 2 * (A.T @ A)


def test_K_assembly_and_solve_against_scipy():
    spec = grid_spec(12, 1.0, time_step=0.005)
    o = po.Oracle(spec)
    K = o.K()
    Ki = assemble_K_independently(o, spec)
    assert abs(K - Ki).max() < 1e-13 * abs(Ki).max()
    assert abs(K - K.T).max() == 0.0
    rng = np.random.default_rng(1)
    b = rng.standard_normal(3 * spec.n_vertices)
    x = o.solve(b).reshape(-1, 3)
    xs = spla.spsolve(Ki.tocsc(), b.reshape(-1, 3))
    assert np.abs(x - xs).max() < 1e-11 * np.abs(xs).max()


def test_free_fall_closed_form():
    spec = grid_spec(6, 1.0, z=5.0, time_step=0.01)
    o = po.Oracle(spec)
    q = spec.rest.reshape(-1, 3).copy()
    v = np.zeros_like(q)
    g = np.array(spec.gravity)
    for _ in range(25):
        o.step(1)
        v = v + spec.time_step * g
        q = q + spec.time_step * v
    qo, vo = o.state()
    # 25 solves with cond(K) ~ 1e3: rounding accumulates to ~1e-11 relative
    assert np.abs(vo.reshape(-1, 3) - v).max() < 1e-10 * np.abs(v).max()
    assert np.abs(qo.reshape(-1, 3) - q).max() < 1e-10 * np.abs(q).max()


def test_inclined_plane_closed_form():
    """replicability/analytic_example_data_processing.py:53-147: after landing the cloth
    slides with acceleration g (sin th - mu cos th)."""
    mu, th = 0.1, np.deg2rad(45.0)
    spec = scenes.inclined_plane(mu=mu)
    o = po.Oracle(spec)
    tangent = np.array([np.cos(th), 0.0, np.sin(th)])
    sp_ = []
    for _ in range(400):
        o.step(1)
        _, v = o.state()
        sp_.append(v.reshape(-1, 3).mean(axis=0) @ -tangent)
    acc = np.polyfit(np.arange(250, 400) * spec.time_step, np.array(sp_)[250:400], 1)[0]
    expect = 9.81 * (np.sin(th) - mu * np.cos(th))
    assert abs(acc - expect) < 1e-9 * expect


def test_sticking_on_incline_stays_at_rest():
    """mu = 1 > tan(30 deg): a cloth resting on the plane does not move."""
    th = np.deg2rad(30.0)
    spec = scenes.inclined_plane(mu=1.0, angle_deg=30.0, height=0.005)
    o = po.Oracle(spec)
    o.step(200)
    _, v = o.state()
    assert np.abs(v).max() < 1e-10
    cs = o.contacts()
    assert len(cs) == 4 and all(c.active for c in cs)
    # normal reaction balances gravity: sum r_N = m g cos(th) dt
    rn = sum(c.r_local[0] for c in cs)
    m = o.masses().sum()
    assert abs(rn - m * 9.81 * np.cos(th) * spec.time_step) < 1e-10 * rn


def test_cone_projection_invariants():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        f = rng.standard_normal(3) * 10 ** rng.uniform(-3, 3)
        mu = rng.uniform(0, 1.5)
        act, r = po.cone(f, mu)
        if f[0] > 0:
            assert not act and np.all(r == 0)
            continue
        assert act
        assert r[0] == -f[0]
        rt = np.hypot(r[1], r[2])
        assert rt <= mu * r[0] * (1 + 1e-14) + 1e-300
        ft = np.hypot(f[1], f[2])
        if ft <= mu * -f[0]:
            assert np.array_equal(r, -f)
        else:
            assert abs(rt - mu * r[0]) <= 1e-14 * max(rt, 1e-300) + 1e-300
            assert abs(r[1] * f[2] - r[2] * f[1]) <= 1e-12 * ft * max(rt, 1e-300)  # parallel
            assert r[1] * f[1] + r[2] * f[2] <= 0


def test_alart_curnier_zero_at_exact_solutions():
    mu = 0.4
    # sticking: u = 0, r inside the cone
    assert po.alart_curnier([0, 0, 0], [1.0, 0.1, -0.2], mu)[0] < 1e-15
    # sliding: u_N = 0, r_T = -mu r_N u_T/|u_T|
    uT = np.array([0.3, -0.4])
    r = np.array([2.0, *(-mu * 2.0 * uT / np.linalg.norm(uT))])
    assert po.alart_curnier([0.0, *uT], r, mu)[0] < 1e-14
    # take-off: r = 0, u_N >= 0
    assert po.alart_curnier([0.5, 0.2, 0.1], [0, 0, 0], mu)[0] == 0.0
    # violation: approaching with zero reaction -> |AC_N| = |u_N|
    nrm, ac = po.alart_curnier([-0.7, 0.0, 0.0], [0, 0, 0], mu)
    assert abs(ac[0] - 0.7) < 1e-15
    # wrong sliding direction is penalised
    assert po.alart_curnier([0.0, *uT], np.array([2.0, *(mu * 2.0 * uT / np.linalg.norm(uT))]), mu)[0] > 0.1


def test_contact_basis_orthonormal():
    rng = np.random.default_rng(3)
    for nrm in list(rng.standard_normal((200, 3))) + [np.array([1.0, 0, 0]), np.array([-1.0, 0, 0])]:
        nrm = nrm / np.linalg.norm(nrm)
        B = po.contact_basis(nrm)
        assert np.allclose(B @ B.T, np.eye(3), atol=1e-14)
        assert np.allclose(np.cross(B[0], B[1]), B[2], atol=1e-14)
        assert np.array_equal(B[0], nrm)


def test_closest_point_brute_force():
    rng = np.random.default_rng(4)
    uu, vv = np.meshgrid(np.linspace(0, 1, 401), np.linspace(0, 1, 401))
    m = uu + vv <= 1
    U, V = uu[m], vv[m]
    for _ in range(60):
        a, b, c = rng.standard_normal((3, 3))
        p = rng.standard_normal(3) * 1.5
        cp, bary = po.closest_point(p, a, b, c, 1e9)
        assert np.allclose(bary.sum(), 1.0)
        assert np.allclose(cp, bary[0] * a + bary[1] * b + bary[2] * c, atol=1e-12)
        pts = a[None] + U[:, None] * (b - a)[None] + V[:, None] * (c - a)[None]
        dmin = np.sqrt(((pts - p) ** 2).sum(1)).min()
        d = np.linalg.norm(cp - p)
        assert d <= dmin + 1e-12
        assert d >= dmin - 5e-3 * max(np.linalg.norm(b - a), np.linalg.norm(c - a))


def test_closest_point_early_exit_uses_plain_tolerance():
    """geometry_util.cpp:525-533 compares |ap|^2-|ab|^2-|ac|^2 against tol (not tol^2)."""
    a, b, c = np.array([0.0, 0, 0]), np.array([0.1, 0, 0]), np.array([0, 0.1, 0])
    p = np.array([0.0, 0.0, 0.2])  # |ap|^2 - |ab|^2 - |ac|^2 = 0.02
    cp, bary = po.closest_point(p, a, b, c, 0.01)
    assert np.array_equal(bary, [1.0, 0.0, 0.0])
    cp, bary = po.closest_point(p, a, b, c, 0.03)
    assert np.array_equal(bary, [1.0, 0.0, 0.0])  # vertex region of a anyway
    p2 = np.array([0.02, 0.02, 0.2])
    cp, bary = po.closest_point(p2, a, b, c, 0.01)
    assert np.array_equal(bary, [1.0, 0.0, 0.0])  # early exit
    cp, bary = po.closest_point(p2, a, b, c, 0.1)
    assert np.allclose(cp, [0.02, 0.02, 0.0])     # true face projection


def test_obstacle_margins():
    """Sphere contact iff |x-c| < R + 1e-2 (Sphere.cpp:93-98); plane iff (x-p).n < 1.5e-2."""
    rest, tris = scenes.square_grid(2, 1.0, 0.0)
    R = 0.25
    spec = scenes.SceneSpec("m", rest, tris, None,
                            [scenes.ObstacleSpec("sphere", (0.0, 0.0, 0.0), radius=R),
                             scenes.ObstacleSpec("plane", (0.0, 0.0, -1.0), (0.0, 0.0, 2.0))],
                            np.array([[0.2]]))
    o = po.Oracle(spec)
    q = rest.reshape(-1, 3).copy()
    q[:] = [0.0, 0.0, 5.0]
    q[4] = [0.0, 0.0, R + 1e-2 - 1e-9]   # centre vertex just inside the margin
    q[0] = [0.0, 0.0, R + 1e-2 + 1e-9]   # just outside
    q[1] = [3.0, 0.0, -1.0 + 1.5e-2 - 1e-9]
    q[2] = [3.0, 0.0, -1.0 + 1.5e-2 + 1e-9]
    o.set_state(q.ravel(), np.zeros(q.size))
    o.detect()
    cs = [(c.obstacle, c.vertex) for c in o.contacts()]
    assert cs == [(0, 4), (1, 1)]
    c = o.contacts()
    assert np.allclose(c[0].point, [0, 0, R]) and np.allclose(c[0].normal, [0, 0, 1])
    assert np.allclose(c[1].normal, [0, 0, 1]) and abs(c[1].point[2] + 1.0) < 1e-15


def _stacked(N=10, gap=0.003, tol=0.004):
    s = scenes.stacked_layers(N=N, side=0.5, zs=(0.3, 0.3 + gap, 0.3 + 2 * gap))
    s.self_tolerance = tol
    return s


def test_self_candidates_grid_equals_brute_force():
    spec = _stacked()
    o = po.Oracle(spec)
    rng = np.random.default_rng(5)
    q = spec.rest.copy() + rng.uniform(-1e-3, 1e-3, spec.rest.size)
    o.set_state(q, np.zeros_like(q))
    o.detect(brute=False)
    a = [(c.vertex, c.triangle, tuple(c.barycentric)) for c in o.candidates()]
    o.detect(brute=True)
    b = [(c.vertex, c.triangle, tuple(c.barycentric)) for c in o.candidates()]
    assert len(a) > 100 and a == b
    assert a == sorted(a)


def test_self_pairing_rules():
    spec = _stacked()
    o = po.Oracle(spec)
    rng = np.random.default_rng(6)
    q = spec.rest.copy() + rng.uniform(-1e-3, 1e-3, spec.rest.size)
    o.set_state(q, np.zeros_like(q))
    o.detect()
    sc = o.self_contacts()
    assert len(sc) > 0
    tris = spec.triangles.reshape(-1, 3)
    seen = set()
    for s in sc:
        assert s.vertex not in s.face and s.partner in s.face
        # node-node pairs are unique and symmetric at most once
        key = (s.vertex, s.partner)
        assert key not in seen
        seen.add(key)
        assert abs(np.linalg.norm(s.normal) - 1) < 1e-14
    assert any(tuple(t) == tuple(s.face) for s in sc[:1] for t in tris)


def _triangle(z, obstacles, mu=0.3, damping=0.0, iterations=10):
    rest = np.array([[0.0, 0.0, z], [1.0, 0.0, z], [0.0, 1.0, z]]).ravel()
    return scenes.SceneSpec("triangle", rest, np.array([0, 1, 2], dtype=np.int32), None, obstacles,
                            np.array([[mu]]), area_density=0.1, time_step=0.01, iterations=iterations,
                            air_damping=damping)


def test_air_damping_free_fall_closed_form():
    """Damping D = alpha (SimulationSpeed.cpp:238-278, SimulationBase.cpp:347-355) on a rigidly
    translating triangle: A^T p - A^T A q^k = 0 and A^T A v = 0, so every iterate solves
    m v^{k+1} = m t - dt alpha v^k with t = v_n + dt g: v^K = t * sum_{j=0..K} (-c)^j, c = dt alpha / m."""
    alpha, K = 0.05, 10
    spec = _triangle(5.0, [scenes.ObstacleSpec("sphere", (0.0, 0.0, -10.0), radius=0.1)], damping=alpha,
                     iterations=K)
    o = po.Oracle(spec)
    v0 = np.tile([0.2, -0.1, 0.3], 3)
    o.set_state(spec.rest.copy(), v0)
    o.step(1)
    q1, v1 = o.state()
    m = o.masses()
    assert np.allclose(m, 0.1 * 0.5 / 3, rtol=1e-15)
    c = spec.time_step * alpha / m[0]
    t = np.array([0.2, -0.1, 0.3]) + spec.time_step * np.array(spec.gravity)
    want = t * sum((-c) ** j for j in range(K + 1))
    assert np.abs(v1.reshape(3, 3) - want).max() < 1e-13
    assert np.abs(q1 - (spec.rest + spec.time_step * v1)).max() < 1e-15
    # without damping the same step is plain free fall
    o2 = po.Oracle(_triangle(5.0, [scenes.ObstacleSpec("sphere", (0.0, 0.0, -10.0), radius=0.1)]))
    o2.set_state(spec.rest.copy(), v0)
    o2.step(1)
    assert np.abs(o2.state()[1].reshape(3, 3) - t).max() < 1e-14


def test_air_damping_is_off_at_contact_vertices():
    """The damping coefficient is zeroed at obstacle-contact vertices (SimulationBase.cpp:347-355):
    a triangle sliding on a frictionless plane keeps its tangential velocity exactly, while the
    same triangle in the air loses it to damping."""
    alpha = 0.5
    plane = [scenes.ObstacleSpec("plane", (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))]
    v0 = np.tile([0.4, -0.3, 0.0], 3)
    o = po.Oracle(_triangle(0.005, plane, mu=0.0, damping=alpha))
    o.set_state(o.spec.rest.copy(), v0)
    o.step(1)
    assert len(o.contacts()) == 3
    v1 = o.state()[1].reshape(3, 3)
    assert np.abs(v1[:, :2] - [0.4, -0.3]).max() < 1e-14
    air = po.Oracle(_triangle(0.5, plane, mu=0.0, damping=alpha))
    air.set_state(air.spec.rest.copy(), v0)
    air.step(1)
    assert len(air.contacts()) == 0
    assert np.abs(air.state()[1].reshape(3, 3)[:, 0]).max() < 0.4 * (1 - 0.5 * 0.01 * alpha / (0.1 * 0.5 / 3))
