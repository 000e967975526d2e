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
