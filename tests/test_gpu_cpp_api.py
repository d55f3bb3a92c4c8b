"""GPU: a reference-format JSON scene (OBJ cloth, OBJ obstacle frames) run through the C++
host interface (getSceneFromJSONDocument -> getSimulation -> step(), with the log
callbacks) ends where the same scene handed to the C ABI directly ends.

The direct run uses the positions as the scene reader stores them (OBJ values read as
float, as the reference's tinyobjloader does), so both runs start from the same state;
both use the same substitution schedule (PDFC_TUNE), so they agree bit for bit."""
import copy
import os

import numpy as np
import pytest

from climatenet_amd import scenes
from tests.scene_files import read_obj_positions, write_scene
from tests.test_cpp_api import api_check, run  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

STEPS = 40


def as_read(spec, dirpath):
    """The spec with every OBJ-borne position replaced by what the reader stores."""
    s = copy.deepcopy(spec)
    s.rest = read_obj_positions(os.path.join(dirpath, "rest.obj")).ravel()
    if s.attachments is not None:  # anchored at the rest positions as read
        idx, _, w = s.attachments
        s.attachments = (idx, s.rest.reshape(-1, 3)[idx].copy(), w)
    s.positions = read_obj_positions(os.path.join(dirpath, "current.obj")).ravel()
    for k, o in enumerate(s.obstacles):
        if o.kind != "mesh":
            continue
        if o.times is None:
            o.frames = read_obj_positions(os.path.join(dirpath, "obstacle%d.obj" % k))[None]
        else:
            o.frames = np.array([read_obj_positions(os.path.join(dirpath, "obstacle%d_%03d.obj" % (k, f)))
                                 for f in range(len(o.frames))])
    return s


SCENES = {"sphere_drape_small": scenes.sphere_drape_small, "mesh_drape_moving": lambda: scenes.mesh_sphere_drape(True),
          "box_drape": scenes.box_drape, "rotating_sphere_drape": scenes.rotating_sphere_drape,
          "hanging_cloth": scenes.hanging_cloth}


@pytest.mark.parametrize("name", list(SCENES))
def test_json_scene_matches_direct_abi(gpu, api_check, tmp_path, name):  # noqa: F811
    spec = SCENES[name]()
    if spec.positions is None:
        spec.positions = spec.rest.copy()
    d = str(tmp_path)
    path = write_scene(d, spec)
    out_file = os.path.join(d, "final.txt")
    print("json scene", name, STEPS, "steps", flush=True)
    extra = []
    if spec.attachments is not None:  # PhysicScene::addConstraint(AttachmentConstraint) before getSimulation
        idx, _, w = spec.attachments
        assert len(set(w)) == 1
        extra = [",".join(str(int(i)) for i in idx) + ":%r" % float(w[0])]
    # the same substitution schedule in both processes (the autotuner may otherwise choose
    # another summation order, and contact-rich runs amplify rounding differences)
    sim = gpu.Simulation(as_read(spec, d))
    env = dict(os.environ, PDFC_TUNE=sim.tuning())
    rc, out, raw = run(api_check, path, "sim", str(STEPS), out_file, *extra, env=env)
    assert rc == 0, raw
    q_json = np.loadtxt(out_file)
    sim.step(STEPS)
    q, _ = sim.state()
    err = np.abs(q_json - q).max() / np.abs(q).max()
    moved = np.abs(q - as_read(spec, d).positions).max()
    print(name, "max rel difference", err, "max displacement", moved)
    assert moved > 1e-3
    assert err == 0.0
