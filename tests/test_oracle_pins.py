"""Independent checks of each oracle stage (what the mathematics fixes).

Each stage of the oracle is recomputed here by a different route -- numpy
reductions, scipy's 8-connected labelling, brute-force searches, closed forms
in float64 -- so that a dropped term, a wrong index or a transposed operand in
the oracle fails one of these.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import pyoracle
import scenes
from parity import bits_from_u32, oracle_components, oracle_root_name


def frames():
    out = []
    img, _ = scenes.scene(0, 640, 480, 3)
    out.append(img)
    img, _ = scenes.scene(1, 1920, 1080, 5)
    out.append(img[:600, :900].copy())
    img, _ = scenes.scene(3, 1280, 800, 2)
    out.append(img[:401, :517].copy())  # ragged tile edges
    rng = np.random.default_rng(7)
    out.append(rng.integers(0, 256, size=(67, 131), dtype=np.uint8))
    return out


FRAMES = frames()


def run(img, **kw):
    H, W = img.shape
    return pyoracle.Oracle(pyoracle.default_config(0, W, H, **kw)).run(img)


def tile_view(img, fill):
    H, W = img.shape
    th, tw = (H + 7) // 8, (W + 7) // 8
    pad = np.full((th * 8, tw * 8), fill, dtype=np.float64)
    pad[:H, :W] = img
    return pad.reshape(th, 8, tw, 8).transpose(0, 2, 1, 3).reshape(th, tw, 64)


@pytest.mark.parametrize("k", range(len(FRAMES)))
def test_tile_stats_and_histograms(k):
    img = FRAMES[k]
    o = run(img)
    H, W = img.shape
    mn = tile_view(img, 255).min(-1)
    mx = tile_view(img, 0).max(-1)
    s = tile_view(img, 0).sum(-1)
    ones = tile_view(np.ones_like(img), 0).sum(-1)
    mean = (s + ones // 2) // ones
    assert np.array_equal(o["tile_min"], mn.astype(np.uint8).ravel())
    assert np.array_equal(o["tile_max"], mx.astype(np.uint8).ravel())
    assert np.array_equal(o["tile_mean"], mean.astype(np.uint8).ravel())
    span = (mx - mn).astype(int).ravel()
    assert np.array_equal(o["span_hist"], np.bincount(span, minlength=256))
    assert np.array_equal(o["mean_hist"], np.bincount(mean.astype(int).ravel() >> 2, minlength=64))
    # lower quartile: first bin whose cumulative count exceeds total / 4
    c = np.cumsum(o["span_hist"])
    assert o["ref"][0] == int(np.argmax(c > c[-1] // 4))


@pytest.mark.parametrize("k", range(len(FRAMES)))
def test_otsu_is_the_argmax_of_between_class_variance(k):
    o = run(FRAMES[k])
    h = o["mean_hist"].astype(np.float64)
    i = np.arange(64)
    wb = np.cumsum(h)
    sb = np.cumsum(i * h)
    wa = wb[-1] - wb
    ok = (wb > 0) & (wa > 0)
    between = np.where(ok, wb * wa * (sb / np.maximum(wb, 1) - (sb[-1] - sb) / np.maximum(wa, 1)) ** 2, -1)
    best = between.max()
    chosen = (int(o["ref"][1]) - 2) // 4
    assert between[chosen] >= best * (1 - 1e-5)
    assert (int(o["ref"][1]) - 2) % 4 == 0


@pytest.mark.parametrize("k", range(len(FRAMES)))
def test_threshold_roi_and_binary_image(k):
    img = FRAMES[k]
    o = run(img)
    H, W = img.shape
    tw, th = o["tile_w"], o["tile_h"]
    tmin = o["tile_min"].reshape(th, tw).astype(int)
    tmax = o["tile_max"].reshape(th, tw).astype(int)
    pmn = ndimage.minimum_filter(tmin, size=3, mode="nearest")
    pmx = ndimage.maximum_filter(tmax, size=3, mode="nearest")
    floor = max(12, int(o["ref"][0]) * 3 // 2)
    thr = np.where(pmx - pmn >= floor, (pmn + pmx + 1) >> 1, int(o["ref"][1]))
    assert np.array_equal(o["tile_thresh"].reshape(th, tw), thr)
    active = (tmax - tmin) >= 12
    assert np.array_equal(o["tile_active"].reshape(th, tw) != 0, active)
    rows = np.nonzero(active.any(1))[0]
    for ty in range(th):
        on = rows.size > 0 and rows[0] - 1 <= ty <= rows[-1] + 1
        assert bool(o["row_active"][ty]) == on
        if on:
            cols = np.nonzero(active[ty])[0]
            lo, hi = (0, tw - 1) if cols.size == 0 else (max(0, cols[0] - 4), min(tw - 1, cols[-1] + 4))
            assert (o["row_start"][ty], o["row_end"][ty]) == (lo, hi)
    # binary image: per pixel, threshold of its own tile, inside the 32-px ROI groups
    white = np.ones((H, W), dtype=bool)
    ys, xs = np.mgrid[0:H, 0:W]
    t = thr[ys >> 3, xs >> 3]
    g = xs >> 5
    ty = ys >> 3
    rs = o["row_start"][ty]
    re_ = o["row_end"][ty]
    inside = (o["row_active"][ty] != 0) & (4 * g <= re_) & (4 * g + 3 >= rs)
    white[inside] = img[inside] > t[inside]
    assert np.array_equal(bits_from_u32(o["bits"], W), white)


@pytest.mark.parametrize("k", range(len(FRAMES)))
def test_runs_and_partition_against_flood_fill(k):
    img = FRAMES[k]
    o = run(img)
    H, W = img.shape
    black = ~bits_from_u32(o["bits"], W)
    # maximal runs by differencing
    comps = oracle_components(o)
    for y in range(H):
        d = np.diff(np.concatenate(([0], black[y].astype(np.int8), [0])))
        starts = np.nonzero(d == 1)[0]
        ends = np.nonzero(d == -1)[0] - 1
        assert [(r[0], r[1]) for r in comps[y]] == list(zip(starts.tolist(), ends.tolist()))
    # 8-connected labelling of the black pixels
    lab, n = ndimage.label(black, structure=np.ones((3, 3), dtype=int))
    first = {}
    for y in range(H):
        for x0, x1, name in comps[y]:
            lbl = lab[y, x0]
            assert (lab[y, x0:x1 + 1] == lbl).all()
            first.setdefault(lbl, name)
            assert first[lbl] == name, "oracle component name differs inside one 8-connected component"
    assert len(set(first.values())) == n
    # component statistics at the roots
    objs = ndimage.find_objects(lab)
    areas = ndimage.sum_labels(np.ones_like(lab), lab, index=np.arange(1, n + 1))
    row_cap = o["row_cap"]
    for lbl, name in first.items():
        ry, rx = name
        sl = objs[lbl - 1]
        # root slot: the first fragment at (ry, rx)
        slot = None
        for i in range(int(o["row_count"][ry])):
            if int(o["runs"][ry, i]) & 0xFFFF == rx:
                slot = ry * row_cap + i
                break
        assert slot is not None
        assert int(o["area"][slot]) == int(areas[lbl - 1])
        x0, y0, x1, y1 = (int(v) for v in o["bbox"][slot])
        assert (y0, y1 + 1, x0, x1 + 1) == (sl[0].start, sl[0].stop, sl[1].start, sl[1].stop)


@pytest.mark.parametrize("k", range(len(FRAMES)))
def test_candidate_filter_and_order(k):
    img = FRAMES[k]
    o = run(img)
    H, W = img.shape
    black = ~bits_from_u32(o["bits"], W)
    lab, n = ndimage.label(black, structure=np.ones((3, 3), dtype=int))
    objs = ndimage.find_objects(lab)
    keys = []
    for lbl in range(1, n + 1):
        sl = objs[lbl - 1]
        a = int((lab[sl] == lbl).sum())
        y0, y1, x0, x1 = sl[0].start, sl[0].stop - 1, sl[1].start, sl[1].stop - 1
        bw, bh = x1 - x0 + 1, y1 - y0 + 1
        if a < 24 or bw < 8 or bh < 8 or x0 <= 0 or y0 <= 0 or x1 >= W - 1 or y1 >= H - 1:
            continue
        if bw > 6 * bh or bh > 6 * bw:
            continue
        fill = np.float32(a) / np.float32(bw * bh)
        if fill < np.float32(0.18) or fill > np.float32(0.95):
            continue
        ys, xs = np.nonzero(lab[sl] == lbl)
        first = (y0 + ys[0], x0 + xs[ys == ys[0]].min())
        keys.append((-a, first))
    keys.sort()
    assert int(o["counts"][0]) == len(keys)
    cap = o["max_candidates"]
    got = [oracle_root_name(o, s) for s in o["candidates"]]
    assert got == [k_[1] for k_ in keys[:cap]]


def test_candidate_cap_keeps_the_largest():
    img, _ = scenes.scene(1, 1920, 1080, 11)
    o = run(img, max_candidates=16)
    full = run(img)
    assert len(o["candidates"]) == 16
    assert [oracle_root_name(o, s) for s in o["candidates"]] == \
        [oracle_root_name(full, s) for s in full["candidates"][:16]]


def test_hull_direction_table_is_correctly_rounded():
    two_pi = np.float32(6.28318530718)
    for d in range(32):
        a = np.float32(np.float32(two_pi * np.float32(d)) / np.float32(32.0))
        ux, uy = pyoracle.hull_direction(d)
        assert ux == float(np.float32(math.cos(float(a))))
        assert uy == float(np.float32(math.sin(float(a))))


def test_axis_normal_matches_half_angle_of_atan2():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        cxx, cyy = rng.uniform(0, 100, 2)
        cxy = rng.uniform(-50, 50)
        if rng.random() < 0.1:
            cxy = 0.0
        cxx, cxy, cyy = (float(np.float32(v)) for v in (cxx, cxy, cyy))
        nx, ny = pyoracle.axis_normal(cxx, cxy, cyy, fp64=True)
        t = 0.5 * math.atan2(2 * cxy, cxx - cyy)
        assert abs(nx + math.sin(t)) < 1e-12 and abs(ny - math.cos(t)) < 1e-12
        fx, fy = pyoracle.axis_normal(cxx, cxy, cyy)
        assert abs(fx + math.sin(t)) < 2e-6 and abs(fy - math.cos(t)) < 2e-6
    # signed zero b with negative a selects -pi/2 as atan2 does
    nx, ny = pyoracle.axis_normal(1.0, -0.0, 2.0, fp64=True)
    assert nx == 1.0 and abs(ny) < 1e-15


def test_fill_homography_reprojects_corners():
    rng = np.random.default_rng(5)
    for _ in range(200):
        c = (np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]]) * rng.uniform(20, 80)
             + rng.normal(0, 6, (4, 2)) + rng.uniform(100, 500, 2)).astype(np.float32)
        h32, h64 = pyoracle.fill_homography(c)
        assert h64[8] == 1.0 or abs(h64[8] - 1.0) < 1e-15
        for i, (tx, ty) in enumerate(((-1, -1), (1, -1), (1, 1), (-1, 1))):
            w = h64[6] * tx + h64[7] * ty + h64[8]
            assert abs((h64[0] * tx + h64[1] * ty + h64[2]) / w - c[i, 0]) < 1e-9
            assert abs((h64[3] * tx + h64[4] * ty + h64[5]) / w - c[i, 1]) < 1e-9
        assert np.array_equal(h32, h64.astype(np.float32))


@pytest.mark.parametrize("family", [0, 1])
def test_match_code_brute_force(family):
    codes, bits = pyoracle.family_codes(family)
    rng = np.random.default_rng(family)
    n = bits * bits
    tests = [int(c) for c in codes[:40]]
    tests += [int(c) ^ (1 << int(b)) for c, b in zip(codes[:40], rng.integers(0, n, 40))]
    tests += [int(c) ^ (1 << int(b)) ^ (1 << int((b + 7) % n)) for c, b in zip(codes[:40], rng.integers(0, n, 40))]
    tests += [int(v) for v in rng.integers(0, 1 << n, 200, dtype=np.uint64)]
    for code in tests:
        for mh in (0, 1, 2, 3):
            d = np.array([bin(int(c) ^ code).count("1") for c in codes])
            exp = int(np.argmin(d)) if d.min() <= mh else -1
            got, h = pyoracle.match_code(family, code, mh)
            assert got == exp
            if exp >= 0:
                assert h == int(d.min())


def test_fp64_twin_agrees_on_clean_frames():
    for seed in range(3):
        img, tags = scenes.scene(0, 1280, 800, seed)
        a = run(img)
        cfg = pyoracle.default_config(0, 1280, 800)
        b = pyoracle.Oracle(cfg, fp64=True).run(img)
        ida = [d["id"] for d in a["detections"]]
        idb = [d["id"] for d in b["detections"]]
        assert ida == idb
        assert np.abs(a["det_corners"] - b["det_corners"]).max() < 0.05
