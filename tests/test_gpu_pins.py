"""The pins of the quad fit and decode, through the HIP detector.

  - The reference benchmark's 32 scenes at its four resolutions, through
    bd_detector_detect_host with max_detections 32 (detector/bench/
    bench_detector.cpp): every detection bit-identical with the oracle's,
    100 % of tags found, no wrong id and the published median corner error
    (README.md:140-155) within 0.005 px.
  - The two halo frames of the nearest-peak rule, stage by stage against the
    oracle, with the fit on the tag's border or on the ring's outer edge.
"""
import numpy as np
import pytest

import bdcudatag
import pyoracle
import scenes
from parity import compare_frame, gpu_stage_dump
from test_oracle_pins_fit import HALO_PINS, corner_distance, halo_frame

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("res", sorted(scenes.REFERENCE_CORNER_MEDIAN))
def test_reference_protocol_accuracy_gpu(res):
    W, H = res
    det = bdcudatag.Detector(0, W, H, max_detections=32)
    o = pyoracle.Oracle(pyoracle.default_config(0, W, H, max_detections=32))
    found = total = wrong = 0
    errors = []
    for i in range(32):
        img, tags = scenes.bench_scene(W, H, i)
        got = det.detect(img)
        want = o.run(img)["detections"]
        assert [d["id"] for d in got] == [d["id"] for d in want]
        for a, b in zip(got, want):
            assert np.array_equal(a["corners"], b["corners"]) and np.array_equal(a["homography"], b["homography"])
        f, n, w, e = scenes.protocol_score(tags, got)
        found, total, wrong, errors = found + f, total + n, wrong + w, errors + e
    assert (found, wrong) == (total, 0)
    median = scenes.protocol_percentile(errors, 0.5)
    assert abs(median - scenes.REFERENCE_CORNER_MEDIAN[res]) <= 0.005, median


@pytest.mark.parametrize("level,inner", [(125, True), (105, False)])
def test_nearest_peak_halo_gpu(level, inner):
    img, tag, outer = halo_frame(level)
    H, W = img.shape
    det = bdcudatag.Detector(0, W, H)
    det.batch.capture_stages(True)
    dets = det.detect(img)
    o = pyoracle.Oracle(pyoracle.default_config(0, W, H)).run(img)
    bad = compare_frame(gpu_stage_dump(det.batch, 0), o, dets, det.stats())
    assert not bad, bad
    assert [d["id"] for d in dets] == [7]
    assert corner_distance(dets[0]["corners"], tag if inner else outer) < (0.3 if inner else 0.1)


@pytest.mark.parametrize("side,ring,level,edge", HALO_PINS)
def test_second_pass_range_and_valley_gpu(side, ring, level, edge):
    img, tag, outer = halo_frame(level, ring=ring, side=side)
    H, W = img.shape
    det = bdcudatag.Detector(0, W, H)
    det.batch.capture_stages(True)
    dets = det.detect(img)
    o = pyoracle.Oracle(pyoracle.default_config(0, W, H)).run(img)
    bad = compare_frame(gpu_stage_dump(det.batch, 0), o, dets, det.stats())
    assert not bad, bad
    assert [d["id"] for d in dets] == [7]
    assert corner_distance(dets[0]["corners"], tag if edge == "tag" else outer) < 0.1


def test_decode_fp64_gate_gpu():
    # Heavily tilted frames where some decode quads take the FP64
    # homography: every stage, decode records included, bit-identical.
    L = pyoracle.lib()
    L.or_decode_fp64_count(1)
    det = bdcudatag.Detector(0, 1280, 800)
    det.batch.capture_stages(True)
    for seed in (0, 1, 2, 3):
        img, _ = scenes.scene(3, 1280, 800, seed)
        dets = det.detect(img)
        o = pyoracle.Oracle(pyoracle.default_config(0, 1280, 800)).run(img)
        bad = compare_frame(gpu_stage_dump(det.batch, 0), o, dets, det.stats())
        assert not bad, (seed, bad)
    assert L.or_decode_fp64_count(1) > 0
