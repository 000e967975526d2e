"""ctypes binding of libbdscenes.so, the synthetic frame generator.

Frames come back as numpy uint8 arrays (height, width); tags as dicts with the
family, id (-1 for decoys), pose (tag square [-1,1]^2 -> continuous pixels) and
ground-truth corners in pixel-index coordinates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbdscenes.so")


class Tag(ctypes.Structure):
    _fields_ = [
        ("family", ctypes.c_int32),
        ("id", ctypes.c_int32),
        ("decoy", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("pose", ctypes.c_double * 9),
        ("corners", ctypes.c_double * 8),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.bds_make_pose.argtypes = [ctypes.c_double] * 5 + [ctypes.POINTER(ctypes.c_double)]
        L.bds_render_tag.argtypes = [
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
            ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
            ctypes.c_int32, ctypes.c_void_p,
        ]
        L.bds_render_tag.restype = ctypes.c_int32
        L.bds_rotate_code.argtypes = [ctypes.c_uint64, ctypes.c_int32]
        L.bds_rotate_code.restype = ctypes.c_uint64
        L.bds_bench_scene.argtypes = [
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p,
            ctypes.POINTER(Tag),
        ]
        L.bds_bench_scene.restype = ctypes.c_int32
        L.bds_scene.argtypes = [
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_void_p,
            ctypes.POINTER(Tag), ctypes.c_int32,
        ]
        L.bds_scene.restype = ctypes.c_int32
        L.bds_scene_batch.argtypes = [
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
            ctypes.c_void_p, ctypes.c_int32,
        ]
        L.bds_scene_batch.restype = ctypes.c_int32
        _lib = L
    return _lib


def make_pose(cx, cy, size, angle, perspective=0.0):
    out = (ctypes.c_double * 9)()
    lib().bds_make_pose(cx, cy, size, angle, perspective, out)
    return np.array(out[:], dtype=np.float64)


def project(pose, x, y):
    p = np.asarray(pose, dtype=np.float64)
    w = p[6] * x + p[7] * y + p[8]
    return (p[0] * x + p[1] * y + p[2]) / w, (p[3] * x + p[4] * y + p[5]) / w


def square_corners(pose, index_shift=0.0):
    """Black-square corners TL, TR, BR, BL of a pose (continuous coordinates,
    minus index_shift)."""
    pts = [project(pose, x, y) for x, y in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    return np.array(pts, dtype=np.float64) - index_shift


def render_tag(family, tag_id, width, height, pose, supersample=2, code=None, flip_one_bit=False):
    img = np.empty((height, width), dtype=np.uint8)
    p = (ctypes.c_double * 9)(*[float(v) for v in pose])
    rc = lib().bds_render_tag(family, tag_id, width, height, p, supersample,
                              1 if code is not None else 0, int(code or 0),
                              1 if flip_one_bit else 0, img.ctypes.data)
    if rc != 0:
        raise ValueError("bds_render_tag failed")
    return img


def rotate_code(code, bits):
    return int(lib().bds_rotate_code(code, bits))


def _tags(buf, n):
    out = []
    for i in range(n):
        t = buf[i]
        out.append({
            "family": int(t.family), "id": int(t.id), "decoy": bool(t.decoy),
            "pose": np.array(t.pose[:], dtype=np.float64),
            "corners": np.array(t.corners[:], dtype=np.float64).reshape(4, 2),
        })
    return out


def bench_scene(width, height, index, seed=6324):
    """Scene `index` of the reference benchmark set at width x height."""
    img = np.empty((height, width), dtype=np.uint8)
    buf = (Tag * 8)()
    n = lib().bds_bench_scene(width, height, index, seed, img.ctypes.data, buf)
    if n < 0:
        raise ValueError("bds_bench_scene failed")
    return img, _tags(buf, n)


def scene(config, width, height, seed, max_tags=256):
    img = np.empty((height, width), dtype=np.uint8)
    buf = (Tag * max_tags)()
    n = lib().bds_scene(config, width, height, seed, img.ctypes.data, buf, max_tags)
    if n < 0:
        raise ValueError("bds_scene failed")
    return img, _tags(buf, n)


def scene_batch(config, width, height, first, count, threads=8):
    imgs = np.empty((count, height, width), dtype=np.uint8)
    if lib().bds_scene_batch(config, width, height, first, count, imgs.ctypes.data, threads) != count:
        raise ValueError("bds_scene_batch failed")
    return imgs


# ---------------------------------------------------------------------------
# Fixture formats of the reference's harnesses
# ---------------------------------------------------------------------------

def write_pgm(path: str, img: np.ndarray) -> None:
    """Binary 8-bit PGM (reference scenes.hpp:212-220)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (img.shape[1], img.shape[0]))
        f.write(img.tobytes())


def read_pgm(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        data = f.read()
    fields, pos = [], 0
    while len(fields) < 4:  # magic, width, height, maxval (comments not used)
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        fields.append(data[pos:end])
        pos = end
    if fields[0] != b"P5" or int(fields[3]) != 255:
        raise ValueError(f"{path}: not an 8-bit binary PGM")
    w, h = int(fields[1]), int(fields[2])
    return np.frombuffer(data[pos + 1:pos + 1 + w * h], dtype=np.uint8).reshape(h, w).copy()


def write_scene_set(scene_list, directory: str, prefix: str, family: str = "tag36h11") -> str:
    """`<directory>/scenes_<prefix>.json` plus one PGM per scene, in the
    reference exporter's layout (scenes.hpp:222-261), which
    compare_detectors.py reads.  scene_list: (image, tags) pairs as returned
    by bench_scene / scene.  Returns the manifest path."""
    os.makedirs(directory, exist_ok=True)
    manifest = os.path.join(directory, f"scenes_{prefix}.json")
    parts = [f'{{\n  "family": "{family}",\n  "scenes": [\n']
    for i, (img, tags) in enumerate(scene_list):
        name = f"{prefix}_{i:03d}.pgm"
        write_pgm(os.path.join(directory, name), img)
        parts.append(f'    {{"file": "{name}", "width": {img.shape[1]}, "height": {img.shape[0]}, "tags": [')
        for k, t in enumerate(tags):
            corners = ", ".join(f"[{c[0]:.3f}, {c[1]:.3f}]" for c in t["corners"])
            parts.append(f'{"," if k else ""}\n      {{"id": {t["id"]}, "corners": [{corners}]}}')
        parts.append(f'\n    ]}}{"," if i + 1 < len(scene_list) else ""}\n')
    parts.append("  ]\n}\n")
    with open(manifest, "w") as f:
        f.write("".join(parts))
    return manifest


def read_scene_sets(directory: str) -> list:
    """Every scenes_*.json of a directory as compare_detectors.py loads them
    (:317-354), images read without OpenCV."""
    import glob
    import json
    sets = []
    for path in sorted(glob.glob(os.path.join(directory, "scenes_*.json"))):
        with open(path) as f:
            manifest = json.load(f)
        items = [{"image": read_pgm(os.path.join(directory, e["file"])), "tags": e["tags"],
                  "width": e["width"], "height": e["height"]} for e in manifest["scenes"]]
        if items:
            sets.append({"label": f"{items[0]['width']}x{items[0]['height']}", "width": items[0]["width"],
                         "height": items[0]["height"], "family": manifest.get("family"), "scenes": items})
    sets.sort(key=lambda s: (s["width"], s["height"]))
    return sets


def export_reference_sets(directory: str, count: int = 32, seed: int = 6324) -> list:
    """What the reference's bdcudatag_scenes exporter writes (scene_gen.cpp):
    `count` benchmark scenes at 640x480, 1280x720, 1280x800 and 1920x1080."""
    out = []
    for w, h in ((640, 480), (1280, 720), (1280, 800), (1920, 1080)):
        out.append(write_scene_set([bench_scene(w, h, i, seed) for i in range(count)], directory, f"{w}x{h}"))
    return out


def write_napa_group(root: str, group: str, scene_list, scenario: str = "synthetic") -> str:
    """Images plus `<root>/data/<group>/metadata.jsonl` in the layout
    eval_napa_dataset.py reads (:90-95): one record per image with
    file_name, width, height, scenario, tag_ids and corners."""
    import json
    folder = os.path.join(root, "data", group)
    os.makedirs(folder, exist_ok=True)
    path = os.path.join(folder, "metadata.jsonl")
    with open(path, "w") as f:
        for i, (img, tags) in enumerate(scene_list):
            name = f"{group}_{i:05d}.pgm"
            write_pgm(os.path.join(folder, name), img)
            real = [t for t in tags if not t.get("decoy", False)]
            f.write(json.dumps({"file_name": name, "width": int(img.shape[1]), "height": int(img.shape[0]),
                                "scenario": scenario, "tag_ids": [int(t["id"]) for t in real],
                                "corners": [np.asarray(t["corners"]).round(3).tolist() for t in real]}) + "\n")
    return path


def score_scene(tags, detections):
    """The harness's scoring (compare_detectors.py:284-314): a tag is found
    when a detection of its id has its centre within 0.75 x the tag side of
    the true centre; corner error is the mean corner distance, corners
    matched as an unordered set; wrong ids are detections of ids absent from
    the scene.  detections: dicts with "id" and "corners" (4x2).
    Returns (found, total, wrong_ids, errors)."""
    import itertools
    orders = list(itertools.permutations(range(4)))
    truth = [(t["id"], np.asarray(t["corners"], dtype=np.float64)) for t in tags]
    known = {i for i, _ in truth}
    wrong = sum(1 for d in detections if d["id"] not in known)
    found, errors = 0, []
    for tag_id, corners in truth:
        centre = corners.mean(axis=0)
        side = 0.5 * (np.linalg.norm(corners[1] - corners[0]) + np.linalg.norm(corners[2] - corners[1]))
        hits = [np.asarray(d["corners"], dtype=np.float64) for d in detections if d["id"] == tag_id]
        hits = [c for c in hits if np.linalg.norm(c.mean(axis=0) - centre) <= 0.75 * side]
        if not hits:
            continue
        found += 1
        errors.append(min(min(float(np.linalg.norm(c[list(o)] - corners, axis=1).mean()) for o in orders)
                          for c in hits))
    return found, len(truth), wrong, errors


def protocol_score(tags, detections):
    """The reference benchmark's accuracy rules (detector/bench/
    bench_detector.cpp:52-85,153-182): a detection counts for a tag when it
    has the tag's id and its centre lies within 0.75 x the tag's mean side
    (of four) of the true centre; the tag's error is the least mean corner
    distance over the detections that count, corners matched as an
    unordered set; a wrong id is a detection whose id is in no tag of the
    scene.  Returns (found, total, wrong_ids, errors)."""
    import itertools
    orders = [list(o) for o in itertools.permutations(range(4))]
    ids = {t["id"] for t in tags}
    wrong = sum(1 for d in detections if d["id"] not in ids)
    found, errors = 0, []
    for t in tags:
        truth = np.asarray(t["corners"], dtype=np.float64)
        side = np.mean([np.hypot(*(truth[(i + 1) % 4] - truth[i])) for i in range(4)])
        best = None
        for d in detections:
            c = np.asarray(d["corners"], dtype=np.float64)
            if d["id"] != t["id"] or np.hypot(*(c.mean(0) - truth.mean(0))) > 0.75 * side:
                continue
            e = min(float(np.hypot(*(c - truth[o]).T).mean()) for o in orders)
            best = e if best is None else min(best, e)
        if best is not None:
            found += 1
            errors.append(best)
    return found, len(tags), wrong, errors


def protocol_percentile(values, fraction):
    """percentile() of bench_detector.cpp:43-48: the ceil(f * n)-th smallest."""
    v = sorted(values)
    if not v:
        return 0.0
    i = int(np.ceil(fraction * len(v)))
    return v[min(len(v) - 1, max(i, 1) - 1)]


# Published accuracy of the reference on its own scenes (README.md:140-155):
# resolution -> median corner error in px (100 % found, 0 wrong ids each).
REFERENCE_CORNER_MEDIAN = {(640, 480): 0.13, (1280, 720): 0.17, (1280, 800): 0.17, (1920, 1080): 0.23}


if __name__ == "__main__":
    import sys
    # python scenes.py [output-dir] [scene-count]: the reference exporter's sets
    target = sys.argv[1] if len(sys.argv) > 1 else "scenes"
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 32
    for m in export_reference_sets(target, n):
        print(f"wrote {m}")
