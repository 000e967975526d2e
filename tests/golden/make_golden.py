"""Writes tests/golden/known_answers.json: the oracle's detections on the
reference's known-answer frames (detector/tests/test_detector.cpp) and on
four of the reference's benchmark scenes (detector/bench/scenes.hpp, seed
6324).  Frames are regenerated from their parameters at test time, so the
fixture holds only inputs' parameters and expected outputs.

  python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "bd-cudatag_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import pyoracle  # noqa: E402
import scenes  # noqa: E402

CASES = [
    {"name": "centered_36h11", "family": 0, "id": 0, "size": [640, 480], "pose": [320, 240, 170, 0.0, 0.0]},
    {"name": "centered_16h5", "family": 1, "id": 0, "size": [640, 480], "pose": [320, 240, 120, 0.0, 0.0]},
    {"name": "rotation_90", "family": 0, "id": 0, "size": [640, 480], "pose": [320, 240, 170, 1.5707963267948966, 0.0]},
    {"name": "small_38px", "family": 0, "id": 7, "size": [640, 480], "pose": [320, 240, 38, 0.25, 0.0]},
    {"name": "homography", "family": 0, "id": 5, "size": [640, 480], "pose": [320, 240, 170, 0.3, 0.00035]},
    {"name": "flipped_bit", "family": 0, "id": 0, "size": [640, 480], "pose": [320, 240, 170, 0.15, 0.0], "flip": True},
]
BENCH = [[640, 480, 0], [640, 480, 3], [1280, 800, 1], [1280, 800, 11]]
CFG = dict(max_detections=32, min_decision_margin=2.0)


def frame_for(case):
    w, h = case["size"]
    pose = scenes.make_pose(*case["pose"])
    return scenes.render_tag(case["family"], case["id"], w, h, pose, flip_one_bit=case.get("flip", False))


def dets(o):
    return [{"id": d["id"], "hamming": d["hamming"], "margin": float(d["decision_margin"]),
             "corners": [float(v) for v in d["corners"].ravel()],
             "homography": [float(v) for v in d["homography"]]} for d in o["detections"]]


def main():
    out = {"config": CFG, "cases": [], "bench": []}
    for case in CASES:
        w, h = case["size"]
        o = pyoracle.Oracle(pyoracle.default_config(case["family"], w, h, **CFG)).run(frame_for(case))
        out["cases"].append(dict(case, detections=dets(o)))
    for w, h, i in BENCH:
        img, _ = scenes.bench_scene(w, h, i)
        o = pyoracle.Oracle(pyoracle.default_config(0, w, h, **CFG)).run(img)
        out["bench"].append({"size": [w, h], "index": i, "checksum": int(np.sum(img, dtype=np.uint64)),
                             "detections": dets(o)})
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
