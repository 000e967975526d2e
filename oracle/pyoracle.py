"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by the smoke check and by the CPU-baseline leg of
bench.py, never by the detector itself.  `Oracle(config, fp64=False)` runs the
scalar restatement of the reference pipeline on one frame and exposes every
stage as numpy arrays.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(_HERE, "build")


class Config(ctypes.Structure):
    _fields_ = [
        ("family", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("max_hamming", ctypes.c_int32), ("min_cluster_pixels", ctypes.c_int32),
        ("min_tag_pixels", ctypes.c_int32), ("max_detections", ctypes.c_int32),
        ("max_candidates", ctypes.c_int32), ("refine_edges", ctypes.c_int32),
        ("min_decision_margin", ctypes.c_float), ("decode_sharpening", ctypes.c_float),
    ]


class Detection(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int32), ("hamming", ctypes.c_int32),
        ("decision_margin", ctypes.c_float), ("corners", ctypes.c_float * 8),
        ("reserved", ctypes.c_int32), ("homography", ctypes.c_float * 9),
    ]


_P8 = ctypes.POINTER(ctypes.c_uint8)
_P32 = ctypes.POINTER(ctypes.c_uint32)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PD = ctypes.POINTER(ctypes.c_double)


class View(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("tile_w", ctypes.c_int32),
        ("tile_h", ctypes.c_int32), ("words", ctypes.c_int32), ("row_cap", ctypes.c_int32),
        ("max_candidates", ctypes.c_int32), ("max_detections", ctypes.c_int32),
        ("max_hamming", ctypes.c_int32), ("dim", ctypes.c_int32), ("code_count", ctypes.c_int32),
        ("min_cluster_pixels", ctypes.c_int32), ("min_tag_pixels", ctypes.c_int32),
        ("refine_edges", ctypes.c_int32),
        ("min_decision_margin", ctypes.c_float), ("decode_sharpening", ctypes.c_float),
        ("tile_min", _P8), ("tile_max", _P8), ("tile_mean", _P8), ("tile_thresh", _P8),
        ("tile_active", _P8), ("span_hist", _P32), ("mean_hist", _P32),
        ("ref", ctypes.c_uint32 * 2), ("row_active", _P8), ("row_start", _PI32),
        ("row_end", _PI32), ("bits", _P32), ("row_count", _P32), ("runs", _P32),
        ("parent", _P32), ("area", _P32), ("bbox", _P32), ("counts", ctypes.c_uint32 * 5),
        ("candidates", _P32), ("n_candidates", ctypes.c_int32), ("quads", _PD),
        ("hull_quads", _PD), ("dec_id", _PI32), ("dec_hamming", _PI32), ("dec_margin", _PD),
        ("dec_corners", _PD), ("detections", ctypes.POINTER(Detection)),
        ("n_detections", ctypes.c_int32), ("det_corners", _PD), ("det_homography", _PD),
    ]


class FitTrace(ctypes.Structure):
    _fields_ = [("fitted", ctypes.c_int32), ("passes", ctypes.c_int32), ("refined", ctypes.c_int32 * 2),
                ("have", (ctypes.c_int32 * 4) * 2), ("best", ((ctypes.c_int32 * 16) * 4) * 2),
                ("decode_variant", ctypes.c_int32)]


_libs = {}


def lib(fp64=False, literal=False):
    """liboracle32 (the parity build), liboracle64 (FP64 twin) or
    liboracle32lit (the reference's transcendental calls and raw edge
    moments restated literally)."""
    key = 64 if fp64 else (3 if literal else 32)
    if key not in _libs:
        name = {32: "liboracle32", 64: "liboracle64", 3: "liboracle32lit"}[key]
        path = os.path.join(BUILD, f"{name}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.or_create.argtypes = [ctypes.POINTER(Config)]
        L.or_create.restype = ctypes.c_void_p
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
        L.or_run.restype = ctypes.c_int32
        L.or_get_view.argtypes = [ctypes.c_void_p, ctypes.POINTER(View)]
        L.or_real_bits.restype = ctypes.c_int32
        L.or_literal.restype = ctypes.c_int32
        L.or_fill_homography.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _PD]
        L.or_hull_direction.argtypes = [ctypes.c_int32, _PD, _PD]
        L.or_hull_direction.restype = ctypes.c_int32
        L.or_axis_normal.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _PD, _PD]
        L.or_match_code.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, _PI32]
        L.or_match_code.restype = ctypes.c_int32
        L.or_family_codes.argtypes = [ctypes.c_int32, _PI32, _PI32]
        L.or_family_codes.restype = ctypes.POINTER(ctypes.c_uint64)
        L.or_fit_edge.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, _PD, ctypes.c_int32,
                                  ctypes.c_int32, _PD]
        L.or_fit_edge.restype = ctypes.c_int32
        L.or_decode_rotation.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, _PD, ctypes.c_int32,
                                         _PI32, _PI32, _PD]
        L.or_decode_rotation.restype = ctypes.c_int32
        L.or_hull_to_quad.argtypes = [_P32, ctypes.c_int32, _PD]
        L.or_hull_to_quad.restype = ctypes.c_int32
        L.or_get_fit_trace.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(FitTrace)]
        L.or_get_fit_trace.restype = ctypes.c_int32
        L.or_decode_fp64_count.argtypes = [ctypes.c_int32]
        L.or_decode_fp64_count.restype = ctypes.c_int64
        assert L.or_real_bits() == (64 if fp64 else 32)
        assert L.or_literal() == (1 if literal else 0)
        _libs[key] = L
    return _libs[key]


def family_codes(family):
    n = ctypes.c_int32()
    bits = ctypes.c_int32()
    p = lib().or_family_codes(family, ctypes.byref(n), ctypes.byref(bits))
    return np.ctypeslib.as_array(p, shape=(n.value,)).copy(), bits.value


def default_config(family, width, height, **kw):
    """bd_config_init defaults (detector/src/detector.cpp:97-110)."""
    c = dict(family=family, width=width, height=height, max_hamming=-1, min_cluster_pixels=24,
             min_tag_pixels=8, max_detections=64, max_candidates=0, refine_edges=1,
             min_decision_margin=3.0, decode_sharpening=0.25)
    c.update(kw)
    return c


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Oracle:
    def __init__(self, config: dict, fp64: bool = False):
        self._lib = lib(fp64)
        self.config = dict(config)
        c = Config(**self.config)
        self._h = self._lib.or_create(ctypes.byref(c))
        if not self._h:
            raise ValueError("or_create rejected the configuration")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.or_destroy(h)
            self._h = None

    def trace(self, candidate: int) -> dict | None:
        """Discrete choices of candidate `candidate`'s fit and decode in the
        last run: passes, corners moved per pass, lines fitted per pass and
        edge, probe kept per sample, decode variant."""
        t = FitTrace()
        if not self._lib.or_get_fit_trace(self._h, candidate, ctypes.byref(t)):
            return None
        return {"fitted": t.fitted, "passes": t.passes, "refined": list(t.refined),
                "have": [list(r) for r in t.have], "best": np.ctypeslib.as_array(t.best).copy(),
                "decode_variant": t.decode_variant}

    def run(self, image: np.ndarray, stride: int | None = None):
        """Runs the pipeline; returns a dict of per-stage numpy arrays."""
        img = np.ascontiguousarray(image, dtype=np.uint8)
        if stride is None:
            stride = img.shape[1] if img.ndim == 2 else self.config["width"]
        n = self._lib.or_run(self._h, img.ctypes.data, stride)
        if n < 0:
            raise ValueError("or_run failed")
        v = View()
        self._lib.or_get_view(self._h, ctypes.byref(v))
        tiles = v.tile_w * v.tile_h
        slots = v.height * v.row_cap
        mc = v.max_candidates
        out = {
            "tile_w": v.tile_w, "tile_h": v.tile_h, "words": v.words, "row_cap": v.row_cap,
            "max_candidates": mc, "max_detections": v.max_detections,
            "max_hamming": v.max_hamming, "dim": v.dim,
            "tile_min": _arr(v.tile_min, tiles, np.uint8),
            "tile_max": _arr(v.tile_max, tiles, np.uint8),
            "tile_mean": _arr(v.tile_mean, tiles, np.uint8),
            "tile_thresh": _arr(v.tile_thresh, tiles, np.uint8),
            "tile_active": _arr(v.tile_active, tiles, np.uint8),
            "span_hist": _arr(v.span_hist, 256, np.uint32),
            "mean_hist": _arr(v.mean_hist, 64, np.uint32),
            "ref": np.array(v.ref[:], dtype=np.uint32),
            "row_active": _arr(v.row_active, v.tile_h, np.uint8),
            "row_start": _arr(v.row_start, v.tile_h, np.int32),
            "row_end": _arr(v.row_end, v.tile_h, np.int32),
            "bits": _arr(v.bits, v.height * v.words, np.uint32).reshape(v.height, v.words),
            "row_count": _arr(v.row_count, v.height, np.uint32),
            "runs": _arr(v.runs, slots, np.uint32).reshape(v.height, v.row_cap),
            "parent": _arr(v.parent, slots, np.uint32).reshape(v.height, v.row_cap),
            "area": _arr(v.area, slots, np.uint32),
            "bbox": _arr(v.bbox, slots * 4, np.uint32).reshape(slots, 4),
            "counts": np.array(v.counts[:], dtype=np.uint32),
            "candidates": _arr(v.candidates, v.n_candidates, np.uint32),
            "quads": _arr(v.quads, mc * 8, np.float64).reshape(mc, 4, 2),
            "hull_quads": _arr(v.hull_quads, mc * 8, np.float64).reshape(mc, 4, 2),
            "dec_id": _arr(v.dec_id, mc, np.int32),
            "dec_hamming": _arr(v.dec_hamming, mc, np.int32),
            "dec_margin": _arr(v.dec_margin, mc, np.float64),
            "dec_corners": _arr(v.dec_corners, mc * 8, np.float64).reshape(mc, 4, 2),
            "n_detections": v.n_detections,
        }
        nd = v.n_detections
        dets = []
        for i in range(nd):
            d = v.detections[i]
            dets.append({
                "id": int(d.id), "hamming": int(d.hamming),
                "decision_margin": float(d.decision_margin),
                "corners": np.array(d.corners[:], dtype=np.float32).reshape(4, 2),
                "homography": np.array(d.homography[:], dtype=np.float32),
            })
        out["detections"] = dets
        out["det_corners"] = _arr(v.det_corners, nd * 8, np.float64).reshape(nd, 4, 2)
        out["det_homography"] = _arr(v.det_homography, nd * 9, np.float64).reshape(nd, 9)
        return out


def fill_homography(corners):
    c = (ctypes.c_float * 8)(*[float(v) for v in np.asarray(corners).reshape(-1)])
    h = (ctypes.c_float * 9)()
    h64 = (ctypes.c_double * 9)()
    lib().or_fill_homography(c, h, h64)
    return np.array(h[:], dtype=np.float32), np.array(h64[:], dtype=np.float64)


def hull_direction(d, fp64=False):
    x = ctypes.c_double()
    y = ctypes.c_double()
    lib(fp64).or_hull_direction(d, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


def axis_normal(cxx, cxy, cyy, fp64=False):
    x = ctypes.c_double()
    y = ctypes.c_double()
    lib(fp64).or_axis_normal(cxx, cxy, cyy, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


def match_code(family, code, max_hamming):
    h = ctypes.c_int32()
    i = lib().or_match_code(family, code, max_hamming, ctypes.byref(h))
    return int(i), int(h.value)
