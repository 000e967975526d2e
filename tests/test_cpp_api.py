"""The C++ Detector API (include/bdcudatag/detector.hpp) through a compiled
test program, tests/cpp/test_detector_api.cpp: exceptions and messages,
the output-capacity clamp, move semantics and stats (reference
detector/src/detector.cpp:227-380)."""
import os
import subprocess

import pytest

import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "bd-cudatag_amd", "lib", "test_detector_api")


def run(*args):
    if not os.path.exists(PROGRAM):
        pytest.fail(f"{PROGRAM} is not built (run __graft_entry__.build())")
    p = subprocess.run([PROGRAM, *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return [line[3:] for line in p.stdout.splitlines() if line.startswith("ok ")]


def test_cpp_api_errors_without_gpu():
    passed = run("cpu")
    assert "unknown family throws" in passed and "height above 16384 throws" in passed


@pytest.mark.gpu
def test_cpp_api_on_gpu(tmp_path):
    img = scenes.render_tag(0, 11, 640, 480, scenes.make_pose(320, 240, 150, 0.1))
    path = str(tmp_path / "id11.pgm")
    scenes.write_pgm(path, img)
    passed = run("gpu", path)
    for check in ("host detect finds id 11", "device detect equals host detect", "capacity 0 writes nothing",
                  "capacity above max_detections is clamped", "stats count the frames run",
                  "move construction transfers the engine", "move assignment transfers the engine",
                  "moved-from detectDevice throws"):
        assert check in passed, (check, passed)
