"""Parity with the oracle on the paths that ordinary eight-tag frames never
reach, each test asserting from the GPU's own buffers that its path ran:

  - a batch holding a 4096x3072 frame (the largest frames of the uneven
    batch workload) beside a 640x480 frame, and a batch of eight 1280x800
    frames (k_binarize_runs with 256-thread workgroups; smaller launches
    use 1024);
  - bands with more runs than the LDS band capacity of k_binarize_runs
    (joined row by row in global memory by k_merge), rows with more runs
    than k_merge stages in LDS (8192 px wide), bands with more band roots than k_stats'
    walk list, and bands with more component roots than k_stats' root table
    (the global-atomic fallbacks), from 1-px stripes, isolated dots, a 1-px
    checkerboard and salt-and-pepper noise;
  - refine_edges = 0 (hull quads only) and use_cuda_graph = 0 (direct
    launches), on the eight-tag workload and on a batch of device frames;
  - tag16h5 on the dense 3840x2160 frame.

reference gpu_detector.cu:362-586 (binarize, runs, union-find, region
stats), detector.cpp:52-53 (refine_edges, use_cuda_graph).
"""
import numpy as np
import pytest

import bdcudatag
import pyoracle
import scenes
from parity import compare_frame, final_outputs, gpu_root_of, gpu_stage_dump, same_detections, same_final_outputs

pytestmark = pytest.mark.gpu

ROOT_FLAG = 0x80000000
MERGE_STAGE_MAX = 768   # runs per row k_merge stages in LDS at most (256 in its 128-thread form)
WALK_CAP = 512         # band roots k_stats walks from its LDS list
ROOT_SLOTS = 256       # component roots per band in k_stats' LDS table


def branch_counts(g: dict) -> dict:
    """What the GPU buffers of one frame show about the paths taken."""
    roi, rc = g["row_roi"], g["row_count"].astype(np.int64)
    runs, labels = g["runs"], g["labels"]
    H = g["height"]
    th = roi.shape[0]
    out = {"nonlocal_bands": 0, "local_bands": 0, "max_row_runs": int(rc.max()) if H else 0,
           "max_band_roots_local": 0, "max_band_roots": 0, "max_band_components": 0}
    for b in range(th):
        if not roi[b, 0]:
            continue
        if roi[b, 3]:
            out["local_bands"] += 1
        else:
            out["nonlocal_bands"] += 1
        flagged, comps = 0, set()
        for y in range(8 * b, min(8 * b + 8, H)):
            n = int(rc[y])
            flagged += int(((runs[y, :n] & ROOT_FLAG) != 0).sum())
            comps.update(gpu_root_of(labels, y * g["row_cap"] + i) for i in range(n))
        out["max_band_roots"] = max(out["max_band_roots"], flagged)
        if roi[b, 3]:
            out["max_band_roots_local"] = max(out["max_band_roots_local"], flagged)
        out["max_band_components"] = max(out["max_band_components"], len(comps))
    return out


def check_batch(imgs, family=0, **cfg):
    # A run as the pipeline runs it, then one under stage capture for the
    # stage-by-stage comparison: the two give the same candidates, quads,
    # decode records and detections.
    sizes = [(im.shape[1], im.shape[0]) for im in imgs]
    batch = bdcudatag.Batch(family, sizes, **cfg)
    plain = batch.detect(imgs)
    plain_out = [final_outputs(batch, i) for i in range(len(imgs))]
    batch.capture_stages(True)
    dets = batch.detect(imgs)
    dumps = []
    for i, img in enumerate(imgs):
        W, H = sizes[i]
        print(f"  frame {i}: {W}x{H} family {family} cfg {cfg}", flush=True)
        assert same_final_outputs(plain_out[i], final_outputs(batch, i)), i
        assert same_detections(plain[i], dets[i]), i
        g = gpu_stage_dump(batch, i)
        ocfg = {k: v for k, v in cfg.items() if k != "use_cuda_graph"}
        o = pyoracle.Oracle(pyoracle.default_config(family, W, H, **ocfg)).run(img)
        bad = compare_frame(g, o, dets[i], batch.stats(i))
        assert not bad, (i, bad)
        dumps.append((g, o, dets[i]))
    return dumps


def tag_frame(W, H, tag_id, cx, cy, size):
    return scenes.render_tag(0, tag_id, W, H, scenes.make_pose(cx, cy, size, 0.3))


def test_batch_4096x3072_beside_640x480():
    big, tags_big = scenes.scene(2, 4096, 3072, 21)
    small, _ = scenes.scene(2, 640, 480, 22)
    (gb, ob, db), (gs, os_, ds) = check_batch([big, small])
    assert gb["width"] == 4096 and int(ob["counts"][0]) > 0
    assert len(db) == len(ob["detections"]) and len(db) >= 1


def test_batch_of_eight_1280x800():
    # 800 bands in one launch: more than one round of 1024-thread
    # workgroups, so k_binarize_runs runs with 256 threads per workgroup
    # (a single frame's 100 bands run with 1024).
    imgs = [scenes.scene(0, 1280, 800, 60 + i)[0] for i in range(8)]
    dumps = check_batch(imgs)
    assert sum(len(d) for _, _, d in dumps) >= 40


def test_batch_of_six_merge_rows_beyond_narrow_staging():
    # Six frames: each lane holds three, so k_merge runs its 128-thread form
    # (rows of up to 256 runs staged in LDS).  One frame has 1-px dots on
    # every band's first row (about 450 runs there) while its bands stay
    # joined in LDS, so the boundary rows exceed the staging and are read
    # from global memory.
    imgs = [scenes.scene(0, 1280, 800, 80 + i)[0] for i in range(5)]
    dots = imgs[0].copy()
    for y in range(80, 720, 8):
        dots[y, 200:1100:2] = 20
    imgs.append(dots)
    dumps = check_batch(imgs)
    g = dumps[-1][0]
    rc, roi = g["row_count"], g["row_roi"]
    wide_local = [b for b in range(roi.shape[0]) if roi[b, 0] and roi[b, 3] and int(rc[8 * b]) > 256]
    assert len(wide_local) > 10, [int(rc[8 * b]) for b in range(roi.shape[0])]
    assert sum(len(d) for _, _, d in dumps) >= 30


def test_stripes_1280x800_band_overflow():
    # 1-px stripes: 450 runs per row, 3600 per band, above the 2048 runs a
    # 1280-px band joins in LDS.
    img = tag_frame(1280, 800, 3, 160, 400, 160)
    img[100:700, 300:1200:2] = 20
    (g, o, dets), = check_batch([img])
    c = branch_counts(g)
    assert c["nonlocal_bands"] > 0 and c["max_band_roots"] > WALK_CAP, c
    assert [d["id"] for d in dets] == [3]


def test_stripes_1920x1080_merge_and_stats_overflow():
    # 850 runs per row over 1700 px: every fallback at once.
    img = tag_frame(1920, 1080, 17, 960, 860, 200)
    img[300:500, 200:1900:2] = 20
    (g, o, dets), = check_batch([img])
    c = branch_counts(g)
    assert c["nonlocal_bands"] > 0, c
    assert c["max_row_runs"] > 760, c
    assert c["max_band_roots"] > WALK_CAP and c["max_band_components"] > ROOT_SLOTS, c
    assert [d["id"] for d in dets] == [17]


def test_stripes_8192_rows_beyond_merge_staging():
    # 1-px stripes over 6000 px: 3000 runs per row, more than k_merge stages
    # in LDS, so the boundary rows are joined from global memory.
    img = tag_frame(8192, 256, 21, 600, 128, 120)
    img[20:236, 1500:7500:2] = 20
    (g, o, dets), = check_batch([img])
    c = branch_counts(g)
    assert c["max_row_runs"] > MERGE_STAGE_MAX and c["nonlocal_bands"] > 0, c
    assert [d["id"] for d in dets] == [21]


def test_dots_local_band_many_roots():
    # Isolated pixels two apart: 1800 runs per band (joined in LDS), each
    # its own component: more band roots than the walk list and more roots
    # than the root table.
    img = tag_frame(1280, 800, 5, 160, 400, 160)
    img[100:700:2, 300:1200:2] = 10
    (g, o, dets), = check_batch([img])
    c = branch_counts(g)
    assert c["max_band_roots_local"] > WALK_CAP and c["max_band_components"] > ROOT_SLOTS, c
    assert [d["id"] for d in dets] == [5]


def test_checkerboard_one_component():
    # A 1-px checkerboard is one 8-connected component: 3200 runs per band,
    # joined in global memory, all unions into one root.
    img = tag_frame(1280, 800, 9, 160, 400, 160)
    yy, xx = np.mgrid[200:600, 400:1200]
    sub = img[200:600, 400:1200]
    sub[((xx + yy) & 1) == 0] = 15
    (g, o, dets), = check_batch([img])
    c = branch_counts(g)
    assert c["nonlocal_bands"] > 0, c
    assert [d["id"] for d in dets] == [9]


@pytest.mark.parametrize("W,H,rows,black", [(1280, 800, (300, 400), 0.2), (4096, 3072, (1000, 1200), 0.5)])
def test_salt_and_pepper(W, H, rows, black):
    # 20 % black pixels stay below 8-connected percolation: about 800 small
    # components per band, more than the root table holds.  At 50 % they
    # percolate; at 4096 px that gives 1000 runs per row and 8000 per band.
    img, _ = scenes.scene(2, W, H, 1)
    rng = np.random.default_rng(W)
    band = img[rows[0]:rows[1]]
    m = rng.random(band.shape) < black
    band[m] = 10
    band[~m] = 240
    (g, o, dets), = check_batch([img])
    c = branch_counts(g)
    if W == 4096:
        assert c["nonlocal_bands"] > 0 and c["max_row_runs"] > 760, c
    else:
        assert c["max_band_components"] > ROOT_SLOTS, c


@pytest.mark.parametrize("seed", [0, 3])
@pytest.mark.parametrize("cfg", [dict(refine_edges=0), dict(use_cuda_graph=0),
                                 dict(refine_edges=0, use_cuda_graph=0)])
def test_config_flags_eight_tags(seed, cfg):
    img, tags = scenes.scene(0, 1280, 800, seed)
    (g, o, dets), = check_batch([img], **cfg)
    assert len(dets) >= 6


def test_direct_launch_device_batch():
    # use_cuda_graph = 0 on the device path: the frame descriptors are
    # uploaded before the direct launches whenever a pointer changes.
    import devmem
    W, H, n = 640, 480, 4
    a = [scenes.scene(0, W, H, 400 + i)[0] for i in range(n)]
    b = [scenes.scene(0, W, H, 500 + i)[0] for i in range(n)]
    abuf, bbuf = devmem.DeviceBuffer(np.stack(a)), devmem.DeviceBuffer(np.stack(b))
    pa = [(abuf.ptr + i * W * H, W) for i in range(n)]
    pb = [(bbuf.ptr + i * W * H, W) for i in range(n)]
    want = {}
    for key, imgs in (("a", a), ("b", b)):
        want[key] = [pyoracle.Oracle(pyoracle.default_config(0, W, H)).run(im)["detections"] for im in imgs]
    batch = bdcudatag.Batch(0, [(W, H)] * n, use_cuda_graph=0)
    for ptrs, key in ((pa, "a"), (pb, "b"), (pa, "a")):
        got = batch.detect_device(ptrs)
        for f in range(n):
            assert [d["id"] for d in got[f]] == [d["id"] for d in want[key][f]]
            for x, y in zip(got[f], want[key][f]):
                assert np.array_equal(x["corners"], y["corners"])
                assert np.array_equal(x["homography"], y["homography"])


def test_tag16h5_dense_3840x2160():
    img, tags = scenes.scene(1, 3840, 2160, 3)
    (g, o, dets), = check_batch([img], family=1)
    assert int(o["counts"][0]) > 0
