"""C++ host interface (reference API mirror): JSON/OBJ readers and the reference's
error behaviour, on the CPU. The simulation itself is exercised in the GPU tests."""
import json
import os
import subprocess

import numpy as np
import pytest

from climatenet_amd import build as B, scenes
from tests.scene_files import read_obj_positions, write_scene

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


def run(exe, *args, env=None):
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, env=env)
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


def test_unknown_obstacle_and_bad_friction(api_check, tmp_path):
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    _edit(path, lambda d: d["Obstacles"].append({"Type": "Cylinder", "Material Identifier": 0}))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "Invalid type for static object type" in raw
    path = write_scene(str(tmp_path), scenes.sphere_drape_small())
    _edit(path, lambda d: d["Obstacles"].append({"Type": "Rotating Sphere", "Material Identifier": 0}))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "cannot found key Radius" in raw
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


@pytest.mark.parametrize("moving", [False, True])
def test_mesh_obstacle_readers(api_check, tmp_path, moving):
    """"Mesh" (src/scene_building.cpp:222-246) and "Moving Mesh" (:248-277): OBJ frames
    <prefix><index zero-padded to Suffix Size>.obj at times index * Time Between Frame."""
    spec = scenes.mesh_sphere_drape(moving=moving)
    path = write_scene(str(tmp_path), spec)
    rc, out, raw = run(api_check, path)
    assert rc == 0, raw
    ob = spec.obstacles[0]
    f = out["obstacle0"].split()
    assert int(f[0]) == 2 and int(f[1]) == ob.material and float(f[2]) == ob.tolerance
    assert int(f[3]) == ob.frames.shape[1] and int(f[4]) == len(ob.triangles)
    assert int(f[5]) == len(ob.frames) and int(f[6]) == ob.triangles.reshape(-1)[-1]
    assert float(f[7]) == (-1.0 if not moving else (len(ob.frames) - 1) * float(ob.times[1] - ob.times[0]))
    last = np.array([float(x) for x in f[8:11]])
    obj = os.path.join(str(tmp_path), "obstacle0_%03d.obj" % (len(ob.frames) - 1) if moving else "obstacle0.obj")
    assert np.array_equal(last, read_obj_positions(obj)[0])
    assert np.abs(last - ob.frames[-1][0]).max() < 1e-7


def test_moving_mesh_missing_frame(api_check, tmp_path):
    spec = scenes.mesh_sphere_drape(moving=True)
    path = write_scene(str(tmp_path), spec)
    _edit(path, lambda d: d["Obstacles"][0].__setitem__("Number Frame", len(spec.obstacles[0].frames) + 1))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "Could not load obj file" in raw and "_021.obj" in raw
    _edit(path, lambda d: d["Obstacles"][0].__setitem__("Number Frame", 1))
    rc, out, raw = run(api_check, path)
    assert rc == 3 and "at least two frames" in raw


def test_box_and_rotating_sphere_readers(api_check, tmp_path):
    """"Axis Aligned Box" (src/scene_building.cpp:204-211) and "Rotating Sphere" (:279-293)."""
    spec = scenes.box_drape()
    spec.obstacles.append(scenes.rotating_sphere_drape().obstacles[0])
    path = write_scene(str(tmp_path), spec)
    rc, out, raw = run(api_check, path)
    assert rc == 0, raw
    f = [float(x) for x in out["obstacle0"].split()]
    assert int(f[0]) == 3 and f[6:12] == [-0.25, -0.25, -0.5, 0.5, 0.5, 0.75]
    f = [float(x) for x in out["obstacle1"].split()]
    assert int(f[0]) == 4 and f[6:12] == [0.25, 0.0, 0.0, 2.0, 3.0, 0.1]


def test_stats_writer_format(api_check, tmp_path):
    """computeAndWriteStats (template/physics/io.tpp:23-63), including its start value
    std::numeric_limits<double>::min() (the smallest positive double) for the maximum,
    which an all-negative series reports."""
    out = tmp_path / "stats.txt"
    rc, _, raw = run(api_check, "--stats", str(out))
    assert rc == 0, raw
    sep = "\n-------------------\n\n"
    want = ("Empty\n\t(No data)\n" + sep + "One\n\t Value : 2.5\n" + sep +
            "Three\n\t Number      : 3\n\t Min value   : 1\n\t Mean value  : 2\n\t Max value   : 3\n" + sep +
            "Negative\n\t Number      : 2\n\t Min value   : -2\n\t Mean value  : -1.5\n"
            "\t Max value   : 2.22507e-308\n" + sep)
    assert out.read_text() == want


REFERENCE = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "scenes")), reason="reference scenes not present")
def test_reference_scene_files_parse(api_check):
    """Every scene file the reference ships (scenes/, replicability/scenes/) reads with the C++
    interface, from a directory one level below the reference root as its relative OBJ paths
    assume; the two whose meshes the reference does not ship fail the way the reference's reader
    does ("Could not load obj file")."""
    import glob
    files = sorted(glob.glob(os.path.join(REFERENCE, "scenes", "*.json")) +
                   glob.glob(os.path.join(REFERENCE, "replicability", "scenes", "*", "*.json")))
    assert len(files) >= 20
    ok = 0
    for f in files:
        r = subprocess.run([api_check, f], capture_output=True, text=True, cwd=os.path.join(REFERENCE, "scenes"))
        out = dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l)
        if r.returncode == 3 and "Could not load obj file" in r.stdout:
            missing = r.stdout.split("Could not load obj file: ")[1].split()[0]
            assert not os.path.exists(os.path.join(REFERENCE, "scenes", missing)), f
            continue
        assert r.returncode == 0, (f, r.stdout[-500:])
        assert int(out["vertices"]) > 0 and int(out["obstacles"]) >= 1
        ok += 1
    assert ok >= 18
