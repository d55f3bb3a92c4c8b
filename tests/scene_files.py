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
        else:
            obstacles.append({"Type": "Plane", "Point": [float(x) for x in o.point],
                              "Normal": [float(x) for x in o.normal], "Material Identifier": int(o.material)})
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
