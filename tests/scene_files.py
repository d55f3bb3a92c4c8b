"""Writers of the reference's scene formats (JSON + OBJ) for tests."""
import json
import os

import numpy as np


def write_obj(path, positions, triangles):
    P = np.asarray(positions).reshape(-1, 3)
    T = np.asarray(triangles).reshape(-1, 3)
    with open(path, "w") as f:
        for p in P:
            f.write("v %.9g %.9g %.9g\n" % tuple(p))
        for t in T:
            f.write("f %d %d %d\n" % tuple(t + 1))


def read_obj_positions(path):
    """Vertex positions of an OBJ file as the scene reader stores them (float32)."""
    rows = [l.split()[1:4] for l in open(path) if l.startswith("v ")]
    return np.array(rows, dtype=np.float32).astype(np.float64)


def write_scene(dirpath, spec, frames=2, steps_per_frame=5, log_vertices=(0, 1)):
    """A JSON scene equivalent to a scenes.SceneSpec, with its OBJ files."""
    rest_obj = os.path.join(dirpath, "rest.obj")
    write_obj(rest_obj, spec.rest, spec.triangles)
    mesh = {"Obj filename": rest_obj, "Area Density": float(spec.area_density), "Bending": float(spec.bending),
            "Stretching": float(spec.stretching), "Material Identifier": int(spec.material)}
    if spec.positions is not None:
        cur = os.path.join(dirpath, "current.obj")
        write_obj(cur, spec.positions, spec.triangles)
        mesh["Current state obj filename"] = cur
    obstacles = []
    for o in spec.obstacles:
        if o.kind == "sphere":
            obstacles.append({"Type": "Sphere", "Position": [float(x) for x in o.point], "Radius": float(o.radius),
                              "Material Identifier": int(o.material)})
        elif o.kind == "box":
            obstacles.append({"Type": "Axis Aligned Box", "Position": [float(x) for x in o.point],
                              "Size": [float(x) for x in o.size], "Material Identifier": int(o.material)})
        elif o.kind == "rotating_sphere":
            obstacles.append({"Type": "Rotating Sphere", "Position": [float(x) for x in o.point],
                              "Radius": float(o.radius), "Rotation Axe": [float(x) for x in o.axis],
                              "Radial Speed": float(o.angular_speed), "Rotation Start Time": float(o.rotation_start),
                              "Material Identifier": int(o.material)})
        elif o.kind == "plane":
            obstacles.append({"Type": "Plane", "Point": [float(x) for x in o.point],
                              "Normal": [float(x) for x in o.normal], "Material Identifier": int(o.material)})
        elif o.times is None:
            path = os.path.join(dirpath, "obstacle%d.obj" % len(obstacles))
            write_obj(path, o.frames[0], o.triangles)
            obstacles.append({"Type": "Mesh", "Path": path, "Threshold": float(o.tolerance),
                              "Material Identifier": int(o.material)})
        else:
            prefix = os.path.join(dirpath, "obstacle%d_" % len(obstacles))
            between = float(o.times[1] - o.times[0])
            assert np.allclose(o.times, between * np.arange(len(o.times)), rtol=0, atol=1e-15)
            for k, fr in enumerate(o.frames):
                write_obj(prefix + "%03d.obj" % k, fr, o.triangles)
            obstacles.append({"Type": "Moving Mesh", "Obj Prefix": prefix, "Suffix Size": 3,
                              "Number Frame": len(o.frames), "Time Between Frame": between,
                              "Collision tolerance": float(o.tolerance), "Material Identifier": int(o.material)})
    doc = {
        "Parameters": {"Frame number": frames, "LocalGlobal": int(spec.iterations), "Timestep": float(spec.time_step),
                       "Type": "friction", "Steps per frame": steps_per_frame,
                       "Self-collision": bool(spec.self_collision),
                       "Self-collision tolerance": float(spec.self_tolerance) if spec.self_collision else 0.0,
                       "Gravity": [float(g) for g in spec.gravity], "Air damping": float(spec.air_damping)},
        "Friction coefficients": [[float(x) for x in row] for row in np.asarray(spec.friction)],
        "Meshes": [mesh],
        "Obstacles": obstacles,
        "log": [{"type": "position", "vertices": list(log_vertices)}],
    }
    path = os.path.join(dirpath, "scene.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path
