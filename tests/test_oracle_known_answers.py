"""The reference's end-to-end known-answer tests (detector/tests/test_detector.cpp)
re-expressed against the CPU oracle, so they run without a GPU.

Same frames (rendered by the port of test_support.hpp's render_tag), same
configuration (max_detections 32, min_decision_margin 2) and the same
tolerances as the reference test file.
"""
import itertools

import numpy as np
import pytest

import pyoracle
import scenes

W, H = 640, 480


def oracle_run(img, family=0, width=W, height=H, **kw):
    cfg = pyoracle.default_config(family, width, height, max_detections=32,
                                  min_decision_margin=2.0)
    cfg.update(kw)
    return pyoracle.Oracle(cfg).run(img)


def check_corners(det, pose, tol):
    # test_detector.cpp:56-67: corners against the pose's projection of the
    # unit square, in continuous coordinates.
    gt = scenes.square_corners(pose)
    assert np.abs(det["corners"] - gt).max() <= tol, (det["corners"], gt)


def test_family_tables():
    # test_detector.cpp:70-102
    for family, count, bits, dmin in ((0, 587, 6, 11), (1, 30, 4, 5)):
        codes, b = pyoracle.family_codes(family)
        assert len(codes) == count and b == bits
        rots = [codes]
        for _ in range(3):
            rots.append(np.array([scenes.rotate_code(int(c), bits) for c in rots[-1]], dtype=np.uint64))
        best = 1000
        for k, r in enumerate(rots):
            x = codes[:, None] ^ r[None, :]
            d = np.array([[bin(int(v)).count("1") for v in row] for row in x])
            if k == 0:
                np.fill_diagonal(d, 1000)
            best = min(best, int(d.min()))
        assert best == dmin


@pytest.mark.parametrize("family,size", [(0, 170), (1, 120)])
def test_centered_tag(family, size):
    # test_detector.cpp:104-137
    pose = scenes.make_pose(W * 0.5, H * 0.5, size, 0.0)
    img = scenes.render_tag(family, 0, W, H, pose)
    o = oracle_run(img, family)
    d = o["detections"]
    assert len(d) == 1
    assert d[0]["id"] == 0 and d[0]["hamming"] == 0 and d[0]["decision_margin"] > 2.0
    check_corners(d[0], pose, 2.5)
    assert o["counts"][0] >= 1


def test_rotations():
    # test_detector.cpp:139-157
    for r in range(4):
        pose = scenes.make_pose(W * 0.5, H * 0.5, 170.0, r * 1.5707963267948966)
        d = oracle_run(scenes.render_tag(0, 0, W, H, pose))["detections"]
        assert len(d) == 1 and d[0]["id"] == 0
        check_corners(d[0], pose, 2.5)


def test_hamming():
    # test_detector.cpp:159-187
    pose = scenes.make_pose(W * 0.5, H * 0.5, 170.0, 0.15)
    img = scenes.render_tag(0, 0, W, H, pose, flip_one_bit=True)
    d = oracle_run(img)["detections"]
    assert len(d) == 1 and d[0]["id"] == 0 and d[0]["hamming"] == 1
    assert oracle_run(img, max_hamming=0)["detections"] == []


def test_multiple():
    # test_detector.cpp:189-228
    w, h = 1280, 720
    cases = [(3, scenes.make_pose(220, 180, 120, 0.2)), (17, scenes.make_pose(640, 180, 140, -0.35)),
             (101, scenes.make_pose(1050, 190, 110, 0.6)), (586, scenes.make_pose(420, 520, 150, 0.9)),
             (42, scenes.make_pose(950, 520, 100, -0.1))]
    img = np.full((h, w), 255, dtype=np.uint8)
    for tid, pose in cases:
        img = np.minimum(img, scenes.render_tag(0, tid, w, h, pose))
    d = oracle_run(img, 0, w, h)["detections"]
    assert len(d) == len(cases)
    assert sorted(x["id"] for x in d) == sorted(t for t, _ in cases)
    for tid, pose in cases:
        for x in d:
            if x["id"] == tid:
                check_corners(x, pose, 3.0)
    assert [x["id"] for x in d] == sorted(x["id"] for x in d)  # sorted by id


def test_small_tag():
    # test_detector.cpp:230-244
    pose = scenes.make_pose(W * 0.5, H * 0.5, 38.0, 0.25)
    d = oracle_run(scenes.render_tag(0, 7, W, H, pose))["detections"]
    assert len(d) == 1 and d[0]["id"] == 7
    check_corners(d[0], pose, 4.0)


def test_blank():
    # test_detector.cpp:246-257
    assert oracle_run(np.full((H, W), 255, np.uint8))["detections"] == []
    assert oracle_run(np.zeros((H, W), np.uint8))["detections"] == []


def test_homography():
    # test_detector.cpp:283-314
    pose = scenes.make_pose(W * 0.5, H * 0.5, 170.0, 0.3, 0.00035)
    d = oracle_run(scenes.render_tag(0, 5, W, H, pose))["detections"]
    assert len(d) == 1
    Hm = d[0]["homography"].astype(np.float64)
    for i, (tx, ty) in enumerate(((-1, -1), (1, -1), (1, 1), (-1, 1))):
        w = Hm[6] * tx + Hm[7] * ty + Hm[8]
        px = (Hm[0] * tx + Hm[1] * ty + Hm[2]) / w
        py = (Hm[3] * tx + Hm[4] * ty + Hm[5]) / w
        assert abs(px - d[0]["corners"][i, 0]) <= 0.01
        assert abs(py - d[0]["corners"][i, 1]) <= 0.01
    assert abs(Hm[8] - 1.0) <= 1e-5
    gx, gy = scenes.project(pose, 0.0, 0.0)
    assert abs(Hm[2] / Hm[8] - gx) <= 3.0 and abs(Hm[5] / Hm[8] - gy) <= 3.0


def test_reference_bench_scenes_recall():
    # bench_detector.cpp:52-85 match rule: same id, centre within 0.75 side;
    # README.md:150 reports 100% recall and 0 wrong ids at 1280x800.
    found = total = wrong = 0
    errs = []
    for index in range(4):
        img, tags = scenes.bench_scene(1280, 800, index)
        d = oracle_run(img, 0, 1280, 800)["detections"]
        ids = {t["id"] for t in tags}
        wrong += sum(1 for x in d if x["id"] not in ids)
        for t in tags:
            gt = t["corners"]
            side = np.mean([np.hypot(*(gt[(k + 1) % 4] - gt[k])) for k in range(4)])
            best = None
            for x in d:
                if x["id"] == t["id"] and np.hypot(*(x["corners"].mean(0) - gt.mean(0))) <= 0.75 * side:
                    e = min(np.mean(np.hypot(*(x["corners"][list(p)] - gt).T))
                            for p in itertools.permutations(range(4)))
                    best = e if best is None else min(best, e)
            total += 1
            if best is not None:
                found += 1
                errs.append(best)
    assert wrong == 0
    assert found == total
    assert np.median(errs) < 0.5
