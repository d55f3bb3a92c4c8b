"""Synthetic cloth scenes (square grids over spheres and planes).

Each scene is what the reference would load from a JSON file plus OBJ meshes:
positions are rounded to float32 exactly like the reference's OBJ reader
(tinyobjloader with ``real_t = float``, extern/tinyObjLoader/tiny_obj_loader.h:100-106),
then widened to float64.

Grid convention: vertex (i, j) has index j*(N+1)+i and sits at
(-side/2 + i*h, -side/2 + j*h, z) with h = side/N; quad (a=(i,j), b=(i+1,j),
c=(i,j+1), d=(i+1,j+1)) splits into triangles (a,b,d) and (a,d,c), counter-
clockwise seen from +z.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi

_MASK64 = (1 << 64) - 1


class MT19937_64:
    """The standard 64-bit Mersenne Twister (std::mt19937_64), seeded like the C++ engine."""

    def __init__(self, seed: int = 0):
        self.mt = [0] * 312
        self.mt[0] = seed & _MASK64
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & _MASK64
        self.index = 312

    def _twist(self):
        mt = self.mt
        upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF
        for i in range(312):
            x = (mt[i] & upper) | (mt[(i + 1) % 312] & lower)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.index = 0

    def __call__(self) -> int:
        if self.index >= 312:
            self._twist()
        y = self.mt[self.index]
        self.index += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK64

    def uniform01(self) -> float:
        """(u64 >> 11) * 2^-53, in [0, 1)."""
        return (self() >> 11) * (1.0 / 9007199254740992.0)


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def square_grid(N: int, side: float, z: float, x0: float = 0.0, y0: float = 0.0):
    """(positions 3n float64 rounded to float32, triangles 3T int32) of an N x N quad grid."""
    h = side / N
    i = np.arange(N + 1, dtype=np.float64)
    X, Y = np.meshgrid(x0 - side / 2 + i * h, y0 - side / 2 + i * h, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="xy")
    a = (jj * (N + 1) + ii).ravel()
    b = a + 1
    c = a + (N + 1)
    d = c + 1
    tris = np.empty((2 * a.size, 3), dtype=np.int32)
    tris[0::2] = np.stack([a, b, d], axis=1)
    tris[1::2] = np.stack([a, d, c], axis=1)
    return f32(pos.ravel()), tris.ravel()


def jitter_xy(pos: np.ndarray, amplitude: float, seed: int = 0) -> np.ndarray:
    """x, y of each vertex += U(-a, a) from mt19937_64(seed), x then y, vertex order."""
    rng = MT19937_64(seed)
    out = pos.reshape(-1, 3).copy()
    for k in range(out.shape[0]):
        out[k, 0] += -amplitude + 2.0 * amplitude * rng.uniform01()
        out[k, 1] += -amplitude + 2.0 * amplitude * rng.uniform01()
    return f32(out.ravel())


@dataclass
class ObstacleSpec:
    kind: str  # "sphere" | "plane" | "mesh" | "box" | "rotating_sphere"
    point: tuple = (0.0, 0.0, 0.0)
    normal: tuple = (0.0, 0.0, 1.0)
    radius: float = 0.0
    material: int = 0
    # triangle mesh: frames [F, nv, 3] (F = 1: static), triangles [T, 3], frame times [F]
    tolerance: float = 0.0
    frames: Optional[np.ndarray] = None
    triangles: Optional[np.ndarray] = None
    times: Optional[np.ndarray] = None
    # box: point is the corner with the smallest coordinates, size the edge lengths
    size: tuple = (0.0, 0.0, 0.0)
    # rotating sphere: axis, angular speed (rad/s), time before the surface turns (s)
    axis: tuple = (0.0, 0.0, 1.0)
    angular_speed: float = 0.0
    rotation_start: float = 0.0


@dataclass
class SceneSpec:
    name: str
    rest: np.ndarray
    triangles: np.ndarray
    positions: Optional[np.ndarray]
    obstacles: List[ObstacleSpec]
    friction: np.ndarray  # n_mat x n_mat
    material: int = 0
    area_density: float = 0.1
    bending: float = 5e-4
    stretching: float = 2e2
    time_step: float = 0.01
    iterations: int = 10
    steps: int = 200
    self_collision: bool = False
    self_tolerance: float = 0.0
    gravity: tuple = (0.0, 0.0, -9.81)
    air_damping: float = 0.0
    # attachment constraints: vertex indices, target positions [k, 3], row weights
    attachments: Optional[tuple] = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_vertices(self) -> int:
        return self.rest.size // 3

    @property
    def n_triangles(self) -> int:
        return self.triangles.size // 3

    def initial_positions(self) -> np.ndarray:
        return self.rest.copy() if self.positions is None else self.positions.copy()

    def c_scene(self) -> abi.Scene:
        rest = np.ascontiguousarray(self.rest, dtype=np.float64)
        pos = None if self.positions is None else np.ascontiguousarray(self.positions, dtype=np.float64)
        tris = np.ascontiguousarray(self.triangles, dtype=np.int32)
        fr = np.ascontiguousarray(self.friction, dtype=np.float64)
        obs = (abi.Obstacle * max(1, len(self.obstacles)))()
        keep = []
        kinds = {"sphere": abi.OBSTACLE_SPHERE, "plane": abi.OBSTACLE_PLANE, "mesh": abi.OBSTACLE_MESH,
                 "box": abi.OBSTACLE_BOX, "rotating_sphere": abi.OBSTACLE_ROTATING_SPHERE}
        for k, o in enumerate(self.obstacles):
            obs[k].type = kinds[o.kind]
            obs[k].material = o.material
            for c in range(3):
                obs[k].point[c] = o.point[c]
                obs[k].normal[c] = o.normal[c]
            obs[k].radius = o.radius
            for c in range(3):
                obs[k].size[c] = o.size[c]
                obs[k].rotation_axis[c] = o.axis[c]
            obs[k].angular_speed = o.angular_speed
            obs[k].rotation_start = o.rotation_start
            if o.kind == "mesh":
                frames = np.ascontiguousarray(o.frames, dtype=np.float64).reshape(-1, np.asarray(o.frames).shape[-2], 3)
                mt = np.ascontiguousarray(o.triangles, dtype=np.int32)
                obs[k].tolerance = o.tolerance
                obs[k].n_mesh_vertices = frames.shape[1]
                obs[k].n_mesh_triangles = mt.size // 3
                obs[k].mesh_triangles = abi.iptr(mt)
                obs[k].n_frames = frames.shape[0]
                obs[k].mesh_frames = abi.dptr(frames)
                tm = None
                if frames.shape[0] >= 2:
                    tm = np.ascontiguousarray(o.times, dtype=np.float64)
                    obs[k].frame_times = abi.dptr(tm)
                keep += [frames, mt, tm]
        s = abi.Scene()
        if self.attachments is not None:
            av = np.ascontiguousarray(self.attachments[0], dtype=np.int32)
            ap = np.ascontiguousarray(self.attachments[1], dtype=np.float64).reshape(-1)
            aw = np.ascontiguousarray(self.attachments[2], dtype=np.float64)
            s.n_attachments = av.size
            s.attachment_vertices = abi.iptr(av)
            s.attachment_positions = abi.dptr(ap)
            s.attachment_weights = abi.dptr(aw)
            keep += [av, ap, aw]
        self._keep = [rest, pos, tris, fr, obs] + keep
        s.n_vertices = self.n_vertices
        s.rest_positions = abi.dptr(rest)
        s.positions = abi.dptr(pos)
        s.n_triangles = self.n_triangles
        s.triangles = abi.iptr(tris)
        s.area_density = self.area_density
        s.bending = self.bending
        s.stretching = self.stretching
        s.material = self.material
        s.n_obstacles = len(self.obstacles)
        s.obstacles = obs
        s.n_materials = fr.shape[0]
        s.friction = abi.dptr(fr)
        return s

    def c_params(self, precision=abi.FP64, refinement_sweeps=2, warm_start=abi.WARM_PLAIN,
                 record_iterations=True, iterations=None, residual_tolerance=0.0,
                 max_iterations=1000, use_graph=False, warm_hold=1) -> abi.Params:
        p = abi.Params()
        p.time_step = self.time_step
        p.local_global_iterations = self.iterations if iterations is None else iterations
        p.self_collision = int(self.self_collision)
        p.self_collision_tolerance = self.self_tolerance
        for c in range(3):
            p.gravity[c] = self.gravity[c]
        p.air_damping = self.air_damping
        p.precision = precision
        p.refinement_sweeps = refinement_sweeps
        p.warm_start = warm_start
        p.record_iterations = int(record_iterations)
        p.residual_tolerance = residual_tolerance
        p.max_iterations = max_iterations
        p.use_graph = int(use_graph)
        p.warm_hold = warm_hold
        return p


def sphere_drape_small(jitter_seed: int = 0) -> SceneSpec:
    """20x20-quad cloth (441 vertices), side 1 m at z=0.3, over a sphere R=0.25, mu=0.3,
    dt=1/100, 10 local-global iterations, 200 steps; initial x,y jitter 1e-4 (seeded)."""
    rest, tris = square_grid(20, 1.0, 0.3)
    pos = jitter_xy(rest, 1e-4, jitter_seed)
    return SceneSpec("drape_20x20", rest, tris, pos,
                     [ObstacleSpec("sphere", (0.0, 0.0, 0.0), radius=0.25)],
                     np.array([[0.3]]), time_step=0.01, iterations=10, steps=200)


def sphere_drape_40k() -> SceneSpec:
    """200x200-quad cloth (40,401 vertices, 80,000 triangles), side 2 m at z=0.55 over a
    sphere R=0.5 at the origin plus a ground plane z=-0.5, mu=0.1, dt=1/200,
    20 local-global iterations, self-collision on (tolerance 5e-3), 500 steps."""
    rest, tris = square_grid(200, 2.0, 0.55)
    return SceneSpec("drape_200x200", rest, tris, None,
                     [ObstacleSpec("sphere", (0.0, 0.0, 0.0), radius=0.5),
                      ObstacleSpec("plane", (0.0, 0.0, -0.5), (0.0, 0.0, 1.0))],
                     np.array([[0.1]]), time_step=0.005, iterations=20, steps=500,
                     self_collision=True, self_tolerance=5e-3)


def stacked_layers(N: int = 150, side: float = 1.5, zs=(0.45, 0.455, 0.46)) -> SceneSpec:
    """Three N x N sheets stacked 5 mm apart (one mesh, three components) over a sphere
    R=0.4; mu=0.3 for cloth-sphere and cloth-cloth; self tolerance 4e-3; dt=1/200,
    30 iterations, 400 steps."""
    P, T = [], []
    nv = (N + 1) ** 2
    for k, z in enumerate(zs):
        p, t = square_grid(N, side, z)
        P.append(p)
        T.append(t + k * nv)
    rest = np.concatenate(P)
    tris = np.concatenate(T).astype(np.int32)
    return SceneSpec("stacked_3x%d" % N, rest, tris, None,
                     [ObstacleSpec("sphere", (0.0, 0.0, 0.0), radius=0.4, material=1)],
                     np.array([[0.3, 0.3], [0.3, 0.3]]), time_step=0.005, iterations=30,
                     steps=400, self_collision=True, self_tolerance=4e-3)


def inclined_plane(mu: float = 0.1, angle_deg: float = 45.0, height: float = 0.05) -> SceneSpec:
    """A 2x2-vertex square cloth (two triangles) held flat and released `height` above a
    plane through the origin tilted by `angle_deg` about y (the closed-form example of
    replicability/analytic_example_data_processing.py)."""
    th = np.deg2rad(angle_deg)
    nrm = (-np.sin(th), 0.0, np.cos(th))
    s = 0.1
    # square lying parallel to the plane, centred above the origin along the normal
    t1 = np.array([np.cos(th), 0.0, np.sin(th)])
    t2 = np.array([0.0, 1.0, 0.0])
    c0 = np.array(nrm) * height
    pts = [c0 + a * t1 + b * t2 for b in (-s / 2, s / 2) for a in (-s / 2, s / 2)]
    rest = f32(np.array(pts).ravel())
    tris = np.array([0, 1, 3, 0, 3, 2], dtype=np.int32)
    return SceneSpec("inclined", rest, tris, None,
                     [ObstacleSpec("plane", (0.0, 0.0, 0.0), nrm)],
                     np.array([[mu]]), area_density=1.0, bending=0.0, stretching=1e3,
                     time_step=1e-3, iterations=50, steps=400)


def batched_instances(count: int = 64, seed: int = 0, sizes=(32, 128)):
    """Independent 1 m cloths of N x N quads, N ~ U{sizes[0]..sizes[1]}, each dropped 5 cm above its
    own sphere: R ~ U[0.2, 0.6], centre offset ~ U[-0.1, 0.1]^2 in the plane, mu ~ U[0, 0.8];
    dt=1/100, 15 iterations, 300 steps. Draws per instance from one mt19937_64(seed), in
    the order N, mu, R, offset x, offset y."""
    rng = MT19937_64(seed)
    specs = []
    for k in range(count):
        span = sizes[1] - sizes[0]
        N = sizes[0] + min(span, int(rng.uniform01() * (span + 1)))
        mu = 0.8 * rng.uniform01()
        R = 0.2 + 0.4 * rng.uniform01()
        cx = -0.1 + 0.2 * rng.uniform01()
        cy = -0.1 + 0.2 * rng.uniform01()
        rest, tris = square_grid(N, 1.0, R + 0.05)
        specs.append(SceneSpec("batch%d_%dx%d" % (k, N, N), rest, tris, None,
                               [ObstacleSpec("sphere", (cx, cy, 0.0), radius=R)],
                               np.array([[mu]]), time_step=0.01, iterations=15, steps=300))
    return specs


def icosphere(radius: float, center=(0.0, 0.0, 0.0), subdivisions: int = 3):
    """Subdivided icosahedron projected on the sphere; triangles wound counter-clockwise seen
    from outside (outward normals). Returns (vertices [nv, 3], triangles [T, 3])."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        G = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            G += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = G
    P = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    return P, np.array(F, dtype=np.int32)


def mesh_sphere_drape(moving: bool = False, jitter_seed: int = 0) -> SceneSpec:
    """The 20x20 drape (1 m, mu=0.3, dt=1/200, 20 iterations) dropped from 2 cm above a
    triangle-mesh sphere of radius 0.25 (icosphere, 1,280 faces), collision tolerance 1.5e-2
    (a mesh obstacle only sees vertices within the tolerance of its surface, so a vertex must
    not cross it in one step). Moving: 21 key frames 0.035 s apart of the sphere turning about
    z by 0.1 rad per frame while rising 1 mm per frame (a MovingMesh with point velocities)."""
    rest, tris = square_grid(20, 1.0, 0.27)
    pos = jitter_xy(rest, 1e-4, jitter_seed)
    P, F = icosphere(0.25, (0.0, 0.0, 0.0), 3)
    if moving:
        frames, times = [], []
        for k in range(21):
            a = 0.1 * k
            R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
            frames.append(P @ R.T + np.array([0.0, 0.0, 1e-3 * k]))
            times.append(0.035 * k)
        ob = ObstacleSpec("mesh", tolerance=1.5e-2, frames=np.array(frames), triangles=F, times=np.array(times))
    else:
        ob = ObstacleSpec("mesh", tolerance=1.5e-2, frames=P[None], triangles=F)
    return SceneSpec("mesh_drape_20x20" + ("_moving" if moving else ""), rest, tris, pos, [ob],
                     np.array([[0.3]]), time_step=0.005, iterations=20, steps=200)


def box_drape(jitter_seed: int = 0) -> SceneSpec:
    """The 20x20 drape (1 m, mu=0.3, dt=1/100, 10 iterations) dropped from z=0.3 on an
    axis-aligned box 0.5 x 0.5 x 0.75 whose top face is at z=0.25 (a box sees only the
    vertices inside it; the nearest face gives the contact)."""
    rest, tris = square_grid(20, 1.0, 0.3)
    pos = jitter_xy(rest, 1e-4, jitter_seed)
    return SceneSpec("box_drape_20x20", rest, tris, pos,
                     [ObstacleSpec("box", (-0.25, -0.25, -0.5), size=(0.5, 0.5, 0.75))],
                     np.array([[0.3]]), time_step=0.01, iterations=10, steps=200)


def rotating_sphere_drape(jitter_seed: int = 0, start: float = 0.1) -> SceneSpec:
    """The 20x20 drape on a sphere R=0.25 that starts turning about z at 3 rad/s after
    `start` seconds (a RotatingSphere: fixed shape, moving surface), mu=0.6."""
    rest, tris = square_grid(20, 1.0, 0.3)
    pos = jitter_xy(rest, 1e-4, jitter_seed)
    return SceneSpec("rotating_sphere_drape_20x20", rest, tris, pos,
                     [ObstacleSpec("rotating_sphere", (0.0, 0.0, 0.0), radius=0.25, axis=(0.0, 0.0, 2.0),
                                   angular_speed=3.0, rotation_start=start)],
                     np.array([[0.6]]), time_step=0.01, iterations=10, steps=200)


def hanging_cloth(weight: float = 100.0) -> SceneSpec:
    """The 20x20 cloth held by two attachment constraints at the corners of one edge (their
    rest positions, row weight `weight`); it swings down onto a sphere R=0.2 below the
    free half."""
    N = 20
    rest, tris = square_grid(N, 1.0, 0.3)
    P = rest.reshape(-1, 3)
    corners = np.array([0, N], dtype=np.int32)
    return SceneSpec("hanging_20x20", rest, tris, None,
                     [ObstacleSpec("sphere", (0.0, 0.25, -0.1), radius=0.2)],
                     np.array([[0.3]]), time_step=0.01, iterations=10, steps=200,
                     attachments=(corners, P[corners].copy(), np.full(2, weight)))


def analytic_incline_mesh(mu: float = 0.1) -> SceneSpec:
    """The setting of the reference's closed-form sliding example (replicability/scenes/Analytic,
    replicability/analytic_example_data_processing.py): a 2 m x 2 m cloth of 4 vertices (area
    density 1, stretching = bending = 2e-4) held parallel to a 45-degree support and released
    0.2 m above it; the support is a triangle-mesh obstacle (a 4 m x 20 m rectangle, two
    triangles, collision threshold 1e-3) through the origin, descending towards +y. mu=0.1,
    dt=1e-3, 20 iterations, 800 steps."""
    th = np.pi / 4
    down = np.array([0.0, np.cos(th), -np.sin(th)])   # along the support, descending
    across = np.array([1.0, 0.0, 0.0])
    S = np.array([c * across + d * down for c, d in ((2.0, -10.0), (2.0, 10.0), (-2.0, 10.0), (-2.0, -10.0))])
    support_tris = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    lift = np.array([0.0, 0.0, 0.2])                  # 0.2 m above the support, vertically
    C = np.array([c * across + d * down + lift for c, d in ((1.0, -1.0), (1.0, 1.0), (-1.0, -1.0), (-1.0, 1.0))])
    tris = np.array([0, 1, 3, 0, 3, 2], dtype=np.int32)
    return SceneSpec("analytic_incline", f32(C.ravel()), tris, None,
                     [ObstacleSpec("mesh", tolerance=1e-3, frames=S[None], triangles=support_tris)],
                     np.array([[mu]]), area_density=1.0, bending=2e-4, stretching=2e-4,
                     time_step=1e-3, iterations=20, steps=800)


def rotating_mesh_drape() -> SceneSpec:
    """A scene shaped like the reference's replicability drapes (replicability/scenes/Square1):
    a 2.4 m square cloth of 76 x 76 quads (5,929 vertices; area density 3, stretching 200,
    bending 5e-4) at z = 0.093 over a sphere of radius 0.5 given as a key-framed triangle mesh
    (icosphere, 2,562 vertices, 5,120 faces; 200 frames 0.035 s apart turning 60 degrees about
    z; collision tolerance 0.015), plus a ground plane at z = -1; mu = 0.1, dt = 1/200,
    20 iterations."""
    rest, tris = square_grid(76, 2.4, 0.093)
    P, F = icosphere(0.5, center=(0.0, 0.0, -0.52), subdivisions=4)
    c = np.array([0.0, 0.0, -0.52])
    frames, times = [], []
    for k in range(200):
        a = np.deg2rad(60.0) * k / 199.0
        R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
        frames.append((P - c) @ R.T + c)
        times.append(0.035 * k)
    return SceneSpec("rotating_mesh_drape_76x76", rest, tris, None,
                     [ObstacleSpec("mesh", tolerance=0.015, frames=np.array(frames), triangles=F,
                                   times=np.array(times)),
                      ObstacleSpec("plane", (0.0, 0.0, -1.0), (0.0, 0.0, 1.0))],
                     np.array([[0.1]]), area_density=3.0, bending=5e-4, stretching=200.0,
                     time_step=0.005, iterations=20, steps=2450)
