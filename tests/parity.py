"""Canonical forms shared by the parity tests.

The HIP detector stores maximal black runs in 64-bit words; the oracle keeps
the reference's fragments split at 32-pixel words.  Both are reduced here to
the same canonical objects:
  - the binary image (first `width` bits of every row, 1 = white);
  - the maximal black runs of every row, as (x0, x1);
  - the component of every maximal run, named by the (y, x0) of the
    component's first run in raster order (the reference's min-slot root);
  - per component: area and bounding box;
  - the ordered candidate list, as component names.
"""
from __future__ import annotations

import numpy as np

S_REF0, S_REF1, S_COUNTS, S_NCAND, S_MAXRUNS, S_DECODED = 2, 3, 4, 10, 11, 12


def bits_from_u32(words: np.ndarray, width: int) -> np.ndarray:
    b = np.unpackbits(words.astype("<u4").view(np.uint8).reshape(words.shape[0], -1),
                      axis=1, bitorder="little")
    return b[:, :width].astype(bool)


def bits_from_u64(words: np.ndarray, width: int) -> np.ndarray:
    b = np.unpackbits(words.astype("<u8").view(np.uint8).reshape(words.shape[0], -1),
                      axis=1, bitorder="little")
    return b[:, :width].astype(bool)


def merge_touching(row):
    """Joins consecutive runs that touch and belong to one component.  Runs
    that touch but stay in different components (the reference never joins
    the 32-pixel fragments of image row 0) are kept apart."""
    out = []
    for x0, x1, name in row:
        if out and out[-1][1] + 1 == x0 and out[-1][2] == name:
            out[-1] = (out[-1][0], x1, name)
        else:
            out.append((x0, x1, name))
    return out


def oracle_components(o: dict):
    """Runs per row with component names, from the oracle's fragments."""
    runs, parent, rc = o["runs"], o["parent"], o["row_count"]
    row_cap = o["row_cap"]
    H = runs.shape[0]
    rows = []
    for y in range(H):
        out = []
        for i in range(int(rc[y])):
            r = int(runs[y, i])
            root = int(parent[y, i])
            ry, ri = divmod(root, row_cap)
            out.append((r & 0xFFFF, r >> 16, (ry, int(runs[ry, ri]) & 0xFFFF)))
        rows.append(merge_touching(out))
    return rows


def oracle_root_name(o: dict, slot: int):
    ry, ri = divmod(int(slot), o["row_cap"])
    return (ry, int(o["runs"][ry, ri]) & 0xFFFF)


def gpu_root_of(labels: np.ndarray, slot: int) -> int:
    """The component root of a run: its label names its band root, whose
    label names the component root (a band root that is the component root
    names itself)."""
    flat = labels.reshape(-1)
    first = int(flat[slot])
    return int(flat[first])


def gpu_components(runs: np.ndarray, labels: np.ndarray, row_count: np.ndarray, row_cap: int):
    """Needs every run and label: a dump of a batch under stage capture."""
    H = runs.shape[0]
    rows = []
    for y in range(H):
        out = []
        for i in range(int(row_count[y])):
            r = int(runs[y, i])
            root = gpu_root_of(labels, y * row_cap + i)
            ry, ri = divmod(root, row_cap)
            out.append((r & 0xFFFF, (r >> 16) & 0x7FFF, (ry, int(runs[ry, ri]) & 0xFFFF)))
        rows.append(merge_touching(out))
    return rows


def gpu_root_name(runs: np.ndarray, row_cap: int, slot: int):
    ry, ri = divmod(int(slot), row_cap)
    return (ry, int(runs[ry, ri]) & 0xFFFF)


def component_stats_oracle(o: dict):
    out = {}
    area, bbox, parent, rc = o["area"], o["bbox"], o["parent"], o["row_count"]
    row_cap = o["row_cap"]
    for y in range(parent.shape[0]):
        for i in range(int(rc[y])):
            slot = y * row_cap + i
            if int(parent[y, i]) == slot:
                out[oracle_root_name(o, slot)] = (int(area[slot]), tuple(int(v) for v in bbox[slot]))
    return out


def component_stats_gpu(runs, labels, row_count, row_cap, area, bbox):
    out = {}
    for y in range(runs.shape[0]):
        for i in range(int(row_count[y])):
            slot = y * row_cap + i
            if int(labels[y, i]) == slot:
                out[gpu_root_name(runs, row_cap, slot)] = (int(area[slot]), tuple(int(v) for v in bbox[slot]))
    return out


def final_outputs(batch, index: int = 0) -> dict:
    """The buffers every run writes in full, stage capture or not: the
    candidate list, refined and hull quads, per-candidate decode records and
    the counts (n candidates), for comparing a run without capture with one
    under capture."""
    sc = batch.debug("scalars", index).view(np.uint32)
    n = int(sc[S_NCAND])
    mc = batch.frames[index]["max_candidates"]
    out = {"counts": sc[S_COUNTS:S_COUNTS + 4].copy(), "n": n,
           "candidates": batch.debug("candidates", index).view(np.uint32)[:n].copy()}
    out["quads"] = batch.debug("quads", index).view(np.uint32).reshape(mc, 8)[:n].copy()
    out["hull_quads"] = batch.debug("hull_quads", index).view(np.uint32).reshape(mc, 8)[:n].copy()
    out["decode"] = batch.debug("decode", index).reshape(mc, -1)[:n].copy()
    return out


def same_final_outputs(a: dict, b: dict) -> bool:
    return a["n"] == b["n"] and all(np.array_equal(a[k], b[k]) for k in
                                    ("counts", "candidates", "quads", "hull_quads", "decode"))


def same_detections(a: list, b: list) -> bool:
    return len(a) == len(b) and all(
        x["id"] == y["id"] and x["hamming"] == y["hamming"]
        and np.array_equal(np.asarray(x["corners"], np.float32).view(np.uint32),
                           np.asarray(y["corners"], np.float32).view(np.uint32))
        and np.array_equal(np.asarray(x["homography"], np.float32).view(np.uint32),
                           np.asarray(y["homography"], np.float32).view(np.uint32))
        for x, y in zip(a, b))


def gpu_stage_dump(batch, index: int = 0) -> dict:
    """All stage buffers of frame `index` of a bdcudatag.Batch as numpy
    arrays.  Runs and labels are whole only after a call under stage capture
    (Batch.capture_stages)."""
    f = batch.frames[index]
    W, H, tw, th = f["width"], f["height"], f["tile_w"], f["tile_h"]
    w64, row_cap, mc = f["words64"], f["row_cap"], f["max_candidates"]
    g = {k: batch.debug(k, index) for k in (
        "tile_min", "tile_max", "tile_mean", "tile_thresh", "hist", "scalars", "row_roi", "bits",
        "row_count", "runs", "labels", "area", "bbox", "candidates", "quads", "hull_quads", "decode")}
    out = {
        "width": W, "height": H, "row_cap": row_cap, "max_candidates": mc,
        "tile_min": g["tile_min"], "tile_max": g["tile_max"], "tile_mean": g["tile_mean"],
        "tile_thresh": g["tile_thresh"],
        "scalars": g["scalars"].view(np.uint32),
        "row_roi": g["row_roi"].view(np.int32).reshape(th, 4),
        "bits": g["bits"].view(np.uint64).reshape(H, w64),
        "row_count": g["row_count"].view(np.uint32),
        "runs": g["runs"].view(np.uint32).reshape(H, row_cap),
        "labels": g["labels"].view(np.uint32).reshape(H, row_cap),
        "area": g["area"].view(np.uint32),
        "bbox": g["bbox"].view(np.uint32).reshape(-1, 4),
        "candidates": g["candidates"].view(np.uint32),
        "quads": g["quads"].view(np.float32).reshape(mc, 4, 2),
        "hull_quads": g["hull_quads"].view(np.float32).reshape(mc, 4, 2),
    }
    dec = g["decode"].reshape(mc, 84)
    out["dec_id"] = dec[:, 0:4].copy().view(np.int32)[:, 0]
    out["dec_hamming"] = dec[:, 4:8].copy().view(np.int32)[:, 0]
    out["dec_margin"] = dec[:, 8:12].copy().view(np.float32)[:, 0]
    out["dec_corners"] = dec[:, 12:44].copy().view(np.float32).reshape(mc, 4, 2)
    return out


def same_f32(a, b) -> bool:
    """Bit-identical float32 arrays (NaN == NaN)."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def compare_frame(g: dict, o: dict, dets_gpu: list, stats_gpu: dict | None = None) -> list[str]:
    """Stage-by-stage comparison; returns a list of mismatch descriptions."""
    bad = []
    W = g["width"]
    for k in ("tile_min", "tile_max", "tile_mean", "tile_thresh"):
        if not np.array_equal(g[k], o[k]):
            bad.append(f"{k}: {int((g[k] != o[k]).sum())} tiles differ")
    sc = g["scalars"]
    if (int(sc[S_REF0]), int(sc[S_REF1])) != (int(o["ref"][0]), int(o["ref"][1])):
        bad.append(f"refs {sc[S_REF0]},{sc[S_REF1]} vs {o['ref']}")
    roi = g["row_roi"]
    if not (np.array_equal(roi[:, 0] != 0, o["row_active"] != 0)):
        bad.append("row_active differs")
    act = o["row_active"] != 0
    if not (np.array_equal(roi[act, 1], o["row_start"][act]) and np.array_equal(roi[act, 2], o["row_end"][act])):
        bad.append("row ROI extents differ")
    gb = bits_from_u64(g["bits"], W)
    ob = bits_from_u32(o["bits"], W)
    if not np.array_equal(gb, ob):
        bad.append(f"binary image: {int((gb != ob).sum())} pixels differ")
        return bad
    gc = gpu_components(g["runs"], g["labels"], g["row_count"], g["row_cap"])
    oc = oracle_components(o)
    if [[r[:2] for r in row] for row in gc] != [[r[:2] for r in row] for row in oc]:
        bad.append("maximal runs differ")
        return bad
    if gc != oc:
        bad.append("component partition differs")
        return bad
    gs = component_stats_gpu(g["runs"], g["labels"], g["row_count"], g["row_cap"], g["area"], g["bbox"])
    os_ = component_stats_oracle(o)
    if gs != os_:
        bad.append("component area/bbox differ")
    counts = sc[S_COUNTS:S_COUNTS + 5]
    if int(counts[0]) != int(o["counts"][0]):
        bad.append(f"candidate count {counts[0]} vs {o['counts'][0]}")
    n = int(sc[S_NCAND])
    gnames = [gpu_root_name(g["runs"], g["row_cap"], s) for s in g["candidates"][:n]]
    onames = [oracle_root_name(o, s) for s in o["candidates"]]
    if gnames != onames:
        bad.append("candidate order differs")
        return bad
    if not same_f32(g["quads"][:n], o["quads"][:n]):
        bad.append("refined quads differ")
    if not same_f32(g["hull_quads"][:n], o["hull_quads"][:n]):
        bad.append("hull quads differ")
    gid = g["dec_id"][:n]
    if not np.array_equal(gid, o["dec_id"][:n]):
        bad.append(f"decoded ids differ: {gid.tolist()} vs {o['dec_id'][:n].tolist()}")
    else:
        ok = gid >= 0
        if not np.array_equal(g["dec_hamming"][:n][ok], o["dec_hamming"][:n][ok]):
            bad.append("decoded hamming differs")
        if not same_f32(g["dec_margin"][:n][ok], o["dec_margin"][:n][ok]):
            bad.append("decoded margins differ")
        if not same_f32(g["dec_corners"][:n][ok], o["dec_corners"][:n][ok]):
            bad.append("decoded corners differ")
    od = o["detections"]
    if len(dets_gpu) != len(od):
        bad.append(f"detection count {len(dets_gpu)} vs {len(od)}")
    else:
        for a, b in zip(dets_gpu, od):
            if a["id"] != b["id"] or a["hamming"] != b["hamming"]:
                bad.append(f"detection id/hamming {a['id']}/{a['hamming']} vs {b['id']}/{b['hamming']}")
            elif not (same_f32(a["corners"], b["corners"]) and same_f32([a["decision_margin"]], [b["decision_margin"]])):
                bad.append(f"detection {a['id']} corners/margin differ")
            elif not same_f32(a["homography"], b["homography"]):
                bad.append(f"detection {a['id']} homography differs")
    if stats_gpu is not None:
        if stats_gpu["runs"] != int(o["counts"][2]):
            bad.append(f"runs statistic {stats_gpu['runs']} vs {o['counts'][2]}")
        if stats_gpu["quads"] != int(o["counts"][3]):
            bad.append(f"quads statistic {stats_gpu['quads']} vs {o['counts'][3]}")
        if stats_gpu["roi_rows"] != min(int(o["counts"][4]) * 8, g["height"]):
            bad.append("roi_rows statistic differs")
    return bad
