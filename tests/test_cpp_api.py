"""C++ host interface (reference API mirror): JSON/OBJ readers and the reference's
error behaviour, on the CPU. The simulation itself is exercised in the GPU tests."""
import json
import os
import subprocess

import numpy as np
import pytest

from climatenet_amd import build as B, scenes
from tests.scene_files import write_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def api_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("api") / "api_check")
    lib = os.path.join(ROOT, "climatenet_amd", "lib")
    if not os.path.exists(os.path.join(lib, "libpdfc_cpp.so")):
        pytest.skip("C++ interface not built")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "climatenet_amd", "cpp"),
                    os.path.join(ROOT, "tests", "native", "api_check.cpp"), "-L" + lib, "-lpdfc_cpp", "-lpdfc",
                    "-Wl,-rpath," + lib, "-o", exe], check=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe] + list(args), capture_output=True, text=True)
    return r.returncode, dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l), r.stdout


def test_scene_json_round_trip(api_check, tmp_path):
    spec = scenes.sphere_drape_small()
    path = write_scene(str(tmp_path), spec)
    rc, out, raw = run(api_check, path)
    assert rc == 0, raw
    assert int(out["vertices"]) == 441 and int(out["triangles"]) == 800
    assert int(out["obstacles"]) == 1 and int(out["materials"]) == 1
    assert float(out["timestep"]) == 0.01 and int(out["iterations"]) == 10
    # "Current state obj filename" positions, read back as float like the reference's OBJ reader
    v1 = np.array([float(x) for x in out["v1"].split()])
    assert np.array_equal(v1, spec.positions.reshape(-1, 3)[1])


def _edit(path, fn):
    doc = json.load(open(path))
    fn(doc)
    text = json.dumps(doc)
    open(path, "w").write(text)
    return text


def test_integer_where_double_required_is_rejected(api_check, tmp_path):
    """jsonRequireDouble uses IsDouble (include/jsonHelper.hpp:136-139): 0 fails, 0.0 passes."""
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    doc = json.load(open(path))
    doc["Parameters"]["Timestep"] = 1
    text = json.dumps(doc).replace('"Timestep": 1,', '"Timestep": 1,')
    open(path, "w").write(text)
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "type error with key Timestep" in raw


def test_missing_key_and_optional_damping(api_check, tmp_path):
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    _edit(path, lambda d: d["Parameters"].pop("Air damping"))
    rc, out, raw = run(api_check, path)
    assert rc == 0 and float(out["damping"]) == 0.0
    _edit(path, lambda d: d["Parameters"].pop("LocalGlobal"))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "cannot found key LocalGlobal" in raw


def test_unsupported_obstacle_and_bad_friction(api_check, tmp_path):
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    _edit(path, lambda d: d["Obstacles"].append({"Type": "Rotating Sphere", "Material Identifier": 0}))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "not supported" in raw
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    _edit(path, lambda d: d.__setitem__("Friction coefficients", [[0.1, 0.2]]))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "square matrix" in raw


def test_malformed_json_and_missing_obj(api_check, tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"Parameters": {"Timestep": 1e-3,, }')
    rc, out, raw = run(api_check, str(p))
    assert rc == 3 and "JSON parse error" in raw
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    _edit(path, lambda d: d["Meshes"][0].__setitem__("Obj filename", str(tmp_path / "nope.obj")))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "Could not load obj file" in raw


def test_simulation_without_device_reports_error(api_check, tmp_path):
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("GPU present")
    except ImportError:
        pass
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    rc, out, raw = run(api_check, path, "sim", "1")
    assert rc == 4 and "no HIP device" in raw
