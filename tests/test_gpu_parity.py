"""Stage-by-stage parity of the HIP detector with the CPU oracle.

Every stage is compared bit-exactly: tile statistics, references, ROI, the
binary image, maximal runs, the component partition (canonical names),
component area/bbox, the ordered candidate list, refined and hull quads,
per-candidate decode results and the final detections with homographies.
The GPU side is read through the C ABI (bd_batch_debug_copy).
"""
import numpy as np
import pytest

import bdcudatag
import pyoracle
import scenes
from parity import (compare_frame, final_outputs, gpu_stage_dump, oracle_components, same_detections,
                    same_final_outputs)

pytestmark = pytest.mark.gpu


def run_both(img, family=0, **cfg):
    H, W = img.shape
    det = bdcudatag.Detector(family, W, H, **cfg)
    plain = det.detect(img)
    plain_out = final_outputs(det.batch, 0)
    det.batch.capture_stages(True)
    dets = det.detect(img)
    assert same_final_outputs(plain_out, final_outputs(det.batch, 0))
    assert same_detections(plain, dets)
    g = gpu_stage_dump(det.batch, 0)
    o = pyoracle.Oracle(pyoracle.default_config(family, W, H, **cfg)).run(img)
    return det, dets, g, o


def check(img, family=0, **cfg):
    det, dets, g, o = run_both(img, family, **cfg)
    bad = compare_frame(g, o, dets, det.stats())
    assert not bad, bad
    return dets, o


KA = dict(max_detections=32, min_decision_margin=2.0)


@pytest.mark.parametrize("family,tag_id,size", [(0, 0, 170), (1, 0, 120), (0, 586, 150), (1, 29, 90)])
def test_centered(family, tag_id, size):
    pose = scenes.make_pose(320, 240, size, 0.0)
    img = scenes.render_tag(family, tag_id, 640, 480, pose)
    dets, _ = check(img, family, **KA)
    assert [d["id"] for d in dets] == [tag_id]
    assert dets[0]["hamming"] == 0


@pytest.mark.parametrize("r", range(4))
def test_rotations(r):
    pose = scenes.make_pose(320, 240, 170, r * np.pi / 2)
    img = scenes.render_tag(0, 0, 640, 480, pose)
    dets, _ = check(img, 0, **KA)
    assert [d["id"] for d in dets] == [0]


@pytest.mark.parametrize("max_hamming,expect", [(-1, 1), (0, 0)])
def test_flipped_bit(max_hamming, expect):
    pose = scenes.make_pose(320, 240, 170, 0.15)
    img = scenes.render_tag(0, 0, 640, 480, pose, flip_one_bit=True)
    dets, _ = check(img, 0, max_hamming=max_hamming, **KA)
    assert len(dets) == expect
    if expect:
        assert dets[0]["hamming"] == 1


def test_small_and_perspective():
    img = scenes.render_tag(0, 7, 640, 480, scenes.make_pose(320, 240, 38, 0.25))
    dets, _ = check(img, 0, **KA)
    assert [d["id"] for d in dets] == [7]
    img = scenes.render_tag(0, 5, 640, 480, scenes.make_pose(320, 240, 170, 0.3, 0.00035))
    dets, _ = check(img, 0, **KA)
    assert [d["id"] for d in dets] == [5]


def test_blank_frames():
    for v in (255, 0, 128):
        img = np.full((480, 640), v, dtype=np.uint8)
        dets, _ = check(img, 0, **KA)
        assert dets == []


@pytest.mark.parametrize("res,index", [((640, 480), 0), ((640, 480), 3), ((1280, 800), 1),
                                       ((1280, 800), 11), ((1920, 1080), 2)])
def test_reference_bench_scenes(res, index):
    img, tags = scenes.bench_scene(res[0], res[1], index)
    dets, _ = check(img, 0, max_detections=32)
    assert len(dets) >= 7


@pytest.mark.parametrize("seed", [0, 1, 2, 5])
def test_workload_eight_tags_1280x800(seed):
    img, tags = scenes.scene(0, 1280, 800, seed)
    check(img, 0)


@pytest.mark.parametrize("family", [0, 1])
def test_dense_1920x1080(family):
    img, tags = scenes.scene(1, 1920, 1080, 7)
    det, dets, g, o = run_both(img, family)
    bad = compare_frame(g, o, dets, det.stats())
    assert not bad, bad
    assert int(o["counts"][0]) > 0


def test_dense_3840x2160_cap():
    img, tags = scenes.scene(1, 3840, 2160, 3)
    det, dets, g, o = run_both(img, 0)
    bad = compare_frame(g, o, dets, det.stats())
    assert not bad, bad


def test_degraded():
    img, tags = scenes.scene(3, 1280, 800, 4)
    check(img, 0)


@pytest.mark.parametrize("W,H", [(16, 16), (1001, 777), (333, 65), (65, 333)])
def test_odd_sizes(W, H):
    rng = np.random.default_rng(W * H)
    img = scenes.render_tag(0, 11, W, H, scenes.make_pose(W * 0.5, H * 0.5, min(W, H) * 0.6, 0.2))
    noise = rng.integers(-4, 5, size=img.shape)
    img = np.clip(img.astype(int) + noise, 0, 255).astype(np.uint8)
    check(img, 0, **KA)


def test_candidate_cap_small():
    img, tags = scenes.scene(1, 1920, 1080, 11)
    det, dets, g, o = run_both(img, 0, max_candidates=16, max_detections=8)
    bad = compare_frame(g, o, dets, det.stats())
    assert not bad, bad


def test_candidate_radix_select():
    # 5035 tag-shaped rings (12x12, 3-px border, fill 0.75): more candidates
    # than the LDS sort holds (2048) and than the 768 kept at 1920x1080, so
    # k_select takes its radix-select path and the cap decides the set.
    W, H = 1920, 1080
    img = np.full((H, W), 200, np.uint8)
    for y in range(10, H - 20, 20):
        for x in range(10, W - 20, 20):
            img[y:y + 12, x:x + 12] = 30
            img[y + 3:y + 9, x + 3:x + 9] = 200
    det, dets, g, o = run_both(img, 0)
    assert int(o["counts"][0]) > 2048
    bad = compare_frame(g, o, dets, det.stats())
    assert not bad, bad


def test_row_table_overflow():
    # Eighty 2-px vertical bars beside a tag put more components in each of
    # the tag's rows than a row's component table holds (64), so the quad
    # fit takes the whole-row scan for those rows.
    img = scenes.render_tag(0, 3, 640, 480, scenes.make_pose(320, 240, 170, 0.0))
    for x in list(range(4, 200, 5)) + list(range(444, 636, 5)):
        img[100:380, x:x + 2] = 20
    det, dets, g, o = run_both(img, 0, **KA)
    comps = oracle_components(o)
    assert min(len({c[2] for c in comps[y]}) for y in range(160, 320)) > 64
    bad = compare_frame(g, o, dets, det.stats())
    assert not bad, bad
    assert [d["id"] for d in dets] == [3]


def same_as_oracle(got, img, family=0, **cfg):
    H, W = img.shape
    want = pyoracle.Oracle(pyoracle.default_config(family, W, H, **cfg)).run(img)["detections"]
    assert [d["id"] for d in got] == [d["id"] for d in want]
    for x, y in zip(got, want):
        assert x["hamming"] == y["hamming"]
        assert np.array_equal(x["corners"], y["corners"])
        assert np.array_equal(x["homography"], y["homography"])


def test_strided_host_frame():
    img, tags = scenes.scene(0, 1280, 800, 9)
    padded = np.zeros((800, 1344), dtype=np.uint8)
    padded[:, :1280] = img
    det = bdcudatag.Detector(0, 1280, 800)
    a = det.detect(padded, stride=1344)
    b = det.detect(img)
    assert [d["id"] for d in a] == [d["id"] for d in b]
    for x, y in zip(a, b):
        assert np.array_equal(x["corners"], y["corners"])
    same_as_oracle(a, img)


def test_device_pointer_path():
    # detector/tests/test_detector.cpp:259-281: id 11 through a device pointer
    import devmem
    img = scenes.render_tag(0, 11, 640, 480, scenes.make_pose(320, 240, 150, 0.1))
    buf = devmem.DeviceBuffer(img)
    det = bdcudatag.Detector(0, 640, 480, **KA)
    a = det.detect_device(buf.ptr, 640)
    b = det.detect(img)
    assert [d["id"] for d in a] == [11]
    same_as_oracle(a, img, **KA)
    assert np.array_equal(a[0]["corners"], b[0]["corners"])
    assert np.array_equal(a[0]["homography"], b[0]["homography"])
    # padded device rows
    padded = np.full((480, 704), 7, dtype=np.uint8)
    padded[:, :640] = img
    pbuf = devmem.DeviceBuffer(padded)
    c = det.detect_device(pbuf.ptr, 704)
    assert [d["id"] for d in c] == [11] and np.array_equal(c[0]["corners"], b[0]["corners"])
    # a device frame spanning 4 GiB or more is rejected before any launch
    with pytest.raises(RuntimeError, match="4 GiB"):
        det.detect_device(buf.ptr, 9_000_000)


def test_device_batch_pointer_changes():
    # Frame descriptors are uploaded only when a pointer, stride or active
    # flag changes: alternate two sets of device frames, then change one
    # frame only, and check every call against a fresh batch.
    import devmem
    W, H, n = 640, 480, 4
    a = [scenes.scene(0, W, H, 200 + i)[0] for i in range(n)]
    b = [scenes.scene(0, W, H, 300 + i)[0] for i in range(n)]
    abuf, bbuf = devmem.DeviceBuffer(np.stack(a)), devmem.DeviceBuffer(np.stack(b))
    pa = [(abuf.ptr + i * W * H, W) for i in range(n)]
    pb = [(bbuf.ptr + i * W * H, W) for i in range(n)]

    def fresh(ptrs):
        return bdcudatag.Batch(0, [(W, H)] * n).detect_device(ptrs)

    def same(x, y):
        assert [[d["id"] for d in f] for f in x] == [[d["id"] for d in f] for f in y]
        for fx, fy in zip(x, y):
            for dx, dy in zip(fx, fy):
                assert np.array_equal(dx["corners"], dy["corners"])

    ra, rb = fresh(pa), fresh(pb)
    assert sum(len(f) for f in ra) > 0 and sum(len(f) for f in rb) > 0
    batch = bdcudatag.Batch(0, [(W, H)] * n)
    mixed = [pa[0], pb[1], pa[2], pa[3]]
    for ptrs, want in [(pa, ra), (pa, ra), (pb, rb), (pa, ra), (mixed, [ra[0], rb[1], ra[2], ra[3]])]:
        same(batch.detect_device(ptrs), want)


def test_misaligned_device_frames():
    # Device frames at 1-3 bytes past a 4-byte boundary with an odd row
    # stride: the pixel-pair loads fold the misalignment in, and the last
    # bytes of the frame are read one at a time.  Results equal the host
    # path's bit for bit.
    import devmem
    img, tags = scenes.scene(0, 1280, 800, 5)
    det = bdcudatag.Detector(0, 1280, 800)
    ref = det.detect(img)
    assert len(ref) >= 6
    stride = 1283
    for off in (1, 2, 3):
        raw = np.full(off + 800 * stride, 111, dtype=np.uint8)
        view = raw[off:].reshape(-1)[:800 * stride].reshape(800, stride)
        view[:, :1280] = img
        buf = devmem.DeviceBuffer(raw)
        got = det.detect_device(buf.ptr + off, stride)
        same_as_oracle(got, img)
        assert [d["id"] for d in got] == [d["id"] for d in ref]
        for p, q in zip(got, ref):
            assert np.array_equal(p["corners"], q["corners"])
            assert np.array_equal(p["homography"], q["homography"])
            assert np.float32(p["decision_margin"]) == np.float32(q["decision_margin"])


def test_stats_and_thread_safety():
    # detector/tests/test_detector.cpp:316-340: 4 threads x 5 frames, one handle
    import threading
    img = scenes.render_tag(0, 0, 640, 480, scenes.make_pose(320, 240, 170, 0.0))
    det = bdcudatag.Detector(0, 640, 480, **KA)
    bad = []

    def work():
        for _ in range(5):
            d = det.detect(img)
            if len(d) != 1 or d[0]["id"] != 0:
                bad.append(d)

    threads = [threading.Thread(target=work) for _ in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    st = det.stats()
    assert not bad
    assert st["frames"] == 20 and st["detections"] == 1 and st["candidates"] >= 1
    assert st["last_gpu_ms"] > 0 and st["average_gpu_ms"] > 0


def test_batch_of_uneven_frames_matches_single():
    sizes = [(640, 480), (1280, 800), (1920, 1200), (1001, 777)]
    imgs = [scenes.scene(2, w, h, 5 + i)[0] for i, (w, h) in enumerate(sizes)]
    batch = bdcudatag.Batch(0, sizes)
    got = batch.detect(imgs)
    for i, (w, h) in enumerate(sizes):
        single = bdcudatag.Detector(0, w, h).detect(imgs[i])
        assert [d["id"] for d in got[i]] == [d["id"] for d in single]
        for a, b in zip(got[i], single):
            assert np.array_equal(a["corners"], b["corners"])
        o = pyoracle.Oracle(pyoracle.default_config(0, w, h)).run(imgs[i])
        assert [d["id"] for d in o["detections"]] == [d["id"] for d in got[i]]


def test_reference_harness_scoring(tmp_path):
    # The reference's scene sets (detector/bench/scene_gen.cpp), exported in
    # its file format and scored by its harness rules
    # (compare_detectors.py:284-314): every tag found, no wrong id (README
    # accuracy table, README.md:140-155).
    scenes.export_reference_sets(str(tmp_path), 4)
    for s in scenes.read_scene_sets(str(tmp_path)):
        det = bdcudatag.Detector(0, s["width"], s["height"], max_detections=32)
        found = total = wrong = 0
        errs = []
        for sc in s["scenes"]:
            f, n, w, e = scenes.score_scene(sc["tags"], det.detect(sc["image"]))
            found, total, wrong = found + f, total + n, wrong + w
            errs += e
        assert found == total and wrong == 0, s["label"]
        assert np.median(errs) < 0.3, (s["label"], np.median(errs))
