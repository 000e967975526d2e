"""Golden fixtures (tests/golden/known_answers.json, written by
tests/golden/make_golden.py): the oracle must keep reproducing them on the
CPU, and the HIP detector must match them bit for bit on the GPU."""
import json
import os

import numpy as np
import pytest

import pyoracle
import scenes

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))
CFG = GOLDEN["config"]


def frame_for(case):
    w, h = case["size"]
    return scenes.render_tag(case["family"], case["id"], w, h, scenes.make_pose(*case["pose"]),
                             flip_one_bit=case.get("flip", False))


def same(dets, want):
    assert [d["id"] for d in dets] == [w["id"] for w in want]
    for d, w in zip(dets, want):
        assert d["hamming"] == w["hamming"]
        assert np.float32(d["decision_margin"]) == np.float32(w["margin"])
        assert np.array_equal(np.asarray(d["corners"], np.float32).ravel(), np.asarray(w["corners"], np.float32))
        assert np.array_equal(np.asarray(d["homography"], np.float32), np.asarray(w["homography"], np.float32))


def all_frames():
    for case in GOLDEN["cases"]:
        yield case["name"], case["family"], frame_for(case), case["detections"]
    for b in GOLDEN["bench"]:
        w, h = b["size"]
        img, _ = scenes.bench_scene(w, h, b["index"])
        assert int(np.sum(img, dtype=np.uint64)) == b["checksum"]
        yield f"bench_{w}x{h}_{b['index']}", 0, img, b["detections"]


def test_oracle_reproduces_golden():
    for name, family, img, want in all_frames():
        h, w = img.shape
        o = pyoracle.Oracle(pyoracle.default_config(family, w, h, **CFG)).run(img)
        same(o["detections"], want)


@pytest.mark.gpu
def test_gpu_matches_golden():
    import bdcudatag
    for name, family, img, want in all_frames():
        h, w = img.shape
        det = bdcudatag.Detector(family, w, h, **CFG)
        same(det.detect(img), want)
