"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the JSON parser, the OBJ
reader, the scene builders and the stats writer (g++ host build of the C++ interface's test
program; no GPU code is instrumented or run). Valid scenes of every obstacle type and
malformed inputs (broken and truncated JSON, 100,000-deep nesting, a cut \\u escape, out-of-range
and malformed numbers, an OBJ with short, non-numeric and out-of-range records) must run
without a sanitizer report, the malformed ones ending in the reader's error."""
import json
import os
import subprocess

import pytest

from climatenet_amd import scenes
from tests.scene_files import write_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "climatenet_amd", "lib")


@pytest.fixture(scope="module")
def asan_check(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB, "libpdfc.so")):
        pytest.skip("library not built")
    exe = str(tmp_path_factory.mktemp("asan") / "api_check_asan")
    cpp = os.path.join(ROOT, "climatenet_amd", "cpp")
    srcs = [os.path.join(ROOT, "tests", "native", "api_check.cpp")] + \
        [os.path.join(cpp, f) for f in ("json.cpp", "scene_building.cpp", "simulation.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                        "-I" + os.path.join(ROOT, "include"), "-I" + cpp] + srcs +
                       ["-L" + LIB, "-lpdfc", "-Wl,-rpath," + LIB, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + r.stderr[-300:])
    return exe


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, env=env, timeout=120)
    report = any(k in r.stderr for k in ("AddressSanitizer", "LeakSanitizer", "runtime error"))
    return r.returncode, report, r.stdout + r.stderr


def test_readers_clean_under_sanitizers(asan_check, tmp_path):
    cases = []
    for name, spec in [("small", scenes.sphere_drape_small()), ("mesh", scenes.mesh_sphere_drape(True)),
                       ("box", scenes.box_drape()), ("rot", scenes.rotating_sphere_drape())]:
        sub = tmp_path / name
        sub.mkdir()
        if spec.positions is None:
            spec.positions = spec.rest.copy()
        cases.append((write_scene(str(sub), spec), 0))
    bad = {
        "broken.json": '{"Parameters": {"Timestep": 1e-3,, }',
        "truncated.json": open(cases[0][0]).read()[:200],
        "deep.json": "[" * 100000 + "]" * 100000,
        "escape.json": '{"a": "\\u12',
        "numbers.json": '{"a": 1e999999, "b": -, "c": 0x12}',
    }
    for name, text in bad.items():
        p = tmp_path / name
        p.write_text(text)
        cases.append((str(p), 3))
    obj = tmp_path / "bad.obj"
    obj.write_text("v 1 2\nv a b c\nf 1 2\nf 1 2 99\nf -1 -2 -3\nf 0 1 2\n")
    doc = json.load(open(cases[0][0]))
    doc["Meshes"][0]["Obj filename"] = str(obj)
    doc["Meshes"][0].pop("Current state obj filename", None)
    p = tmp_path / "badobj.json"
    p.write_text(json.dumps(doc))
    cases.append((str(p), 3))
    for path, want in cases:
        rc, report, out = run(asan_check, path)
        assert not report, (path, out[-2000:])
        assert rc == want, (path, rc, out[-500:])
    rc, report, out = run(asan_check, "--stats", str(tmp_path / "stats.txt"))
    assert rc == 0 and not report, out[-2000:]
