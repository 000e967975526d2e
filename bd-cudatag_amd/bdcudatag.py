"""ctypes binding of libbdcudatag.so (the HIP detector), shaped like the
reference's Python caller (detector/bench/compare_detectors.py:87-206): the
same structures and the same create / detect_host / get_stats / destroy calls,
plus the version 3 batch, device-pointer and introspection entry points.

Loading fails loudly when the library is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbdcudatag.so")

BD_OK = 0
BD_ERROR_INVALID_ARGUMENT = -1
BD_ERROR_CUDA = -2
BD_ERROR_OUT_OF_MEMORY = -3

STAGES = {
    "tile_min": 0, "tile_max": 1, "tile_mean": 2, "hist": 3, "scalars": 4, "row_roi": 5,
    "bits": 6, "row_count": 7, "runs": 8, "labels": 9, "area": 10, "bbox": 11,
    "candidates": 12, "quads": 13, "hull_quads": 14, "decode": 15, "tile_thresh": 16,
}
KERNELS = ["tile_stats", "binarize_runs", "merge", "stats", "collect",
           "select", "fit", "decode", "finalize"]


class Config(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("max_hamming", ctypes.c_int32),
        ("min_cluster_pixels", ctypes.c_int32), ("min_tag_pixels", ctypes.c_int32),
        ("max_detections", ctypes.c_int32), ("max_candidates", ctypes.c_int32),
        ("refine_edges", ctypes.c_int32), ("use_cuda_graph", ctypes.c_int32),
        ("min_decision_margin", ctypes.c_float), ("decode_sharpening", ctypes.c_float),
    ]


class Detection(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int32), ("hamming", ctypes.c_int32),
        ("decision_margin", ctypes.c_float), ("corners", ctypes.c_float * 8),
        ("reserved", ctypes.c_int32), ("homography", ctypes.c_float * 9),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("frames", ctypes.c_uint64), ("last_gpu_ms", ctypes.c_float),
        ("average_gpu_ms", ctypes.c_float), ("last_total_ms", ctypes.c_float),
        ("runs", ctypes.c_int32), ("candidates", ctypes.c_int32),
        ("detections", ctypes.c_int32), ("quads", ctypes.c_int32), ("roi_rows", ctypes.c_int32),
    ]


class Frame(ctypes.Structure):
    _fields_ = [
        ("gray", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("stride", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class ServerClass(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("lanes", ctypes.c_int32),
                ("lane_frames", ctypes.c_int32), ("queue_depth", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("request", ctypes.c_int64), ("class_index", ctypes.c_int32),
                ("count", ctypes.c_int32), ("written", ctypes.c_int32),
                ("batch_frames", ctypes.c_int32), ("latency_ms", ctypes.c_double)]


class StreamSpec(ctypes.Structure):
    _fields_ = [("class_index", ctypes.c_int32), ("frame_count", ctypes.c_int32),
                ("fps", ctypes.c_double), ("frames", ctypes.POINTER(ctypes.c_void_p)),
                ("stride", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ServeReport(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("offered", "completed", "dropped", "launches")] + \
               [(k, ctypes.c_double) for k in ("seconds", "fps", "mpix_s", "p50_ms", "p99_ms",
                                                "p999_ms", "max_ms", "mean_ms", "mean_batch")]


# Every symbol the public header declares, with (restype, argtypes).
class Tune(ctypes.Structure):
    _fields_ = [("fit_grid", ctypes.c_int32), ("decode_grid", ctypes.c_int32),
                ("head_ranks", ctypes.c_int32), ("streams", ctypes.c_int32),
                ("decode_waves", ctypes.c_int32), ("chunk_px", ctypes.c_int32),
                ("head_area", ctypes.c_int32), ("band_rows", ctypes.c_int32)]


TUNE_KEYS = tuple(name for name, _ in Tune._fields_)


_P = ctypes.c_void_p
_I = ctypes.c_int32
SIGNATURES = {
    "bd_api_version": (ctypes.c_uint32, []),
    "bd_version_string": (ctypes.c_char_p, []),
    "bd_family_count": (_I, []),
    "bd_family_name": (ctypes.c_char_p, [_I]),
    "bd_family_bits_per_side": (_I, [_I]),
    "bd_family_min_hamming": (_I, [_I]),
    "bd_family_default_max_hamming": (_I, [_I]),
    "bd_config_init": (None, [ctypes.POINTER(Config)]),
    "bd_detector_create": (_P, [_I, _I, _I, ctypes.POINTER(Config), ctypes.POINTER(_I)]),
    "bd_detector_destroy": (None, [_P]),
    "bd_detector_detect_host": (_I, [_P, _P, _I, ctypes.POINTER(Detection), _I]),
    "bd_detector_detect_device": (_I, [_P, _P, _I, ctypes.POINTER(Detection), _I]),
    "bd_detector_get_stats": (None, [_P, ctypes.POINTER(Stats)]),
    "bd_last_error": (ctypes.c_char_p, []),
    "bd_batch_create": (_P, [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(Config),
                             ctypes.POINTER(_I)]),
    "bd_batch_destroy": (None, [_P]),
    "bd_batch_detect_host": (_I, [_P, ctypes.POINTER(Frame), _I, ctypes.POINTER(Detection), _I,
                                  ctypes.POINTER(_I)]),
    "bd_batch_detect_device": (_I, [_P, ctypes.POINTER(Frame), _I, ctypes.POINTER(Detection), _I,
                                    ctypes.POINTER(_I)]),
    "bd_batch_get_stats": (None, [_P, _I, ctypes.POINTER(Stats)]),
    "bd_batch_debug_copy": (ctypes.c_int64, [_P, _I, _I, _P, ctypes.c_int64]),
    "bd_batch_frame_info": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "bd_detector_batch": (_P, [_P]),
    "bd_batch_time_stages": (_I, [_P, _I, ctypes.POINTER(ctypes.c_float)]),
    "bd_batch_replay": (_I, [_P, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "bd_batch_get_tune": (_I, [_P, ctypes.POINTER(Tune)]),
    "bd_batch_set_tune": (_I, [_P, ctypes.POINTER(Tune)]),
    "bd_batch_set_capture": (_I, [_P, _I]),
    "bd_batch_autotune": (_I, [_P, _I, ctypes.POINTER(Tune), ctypes.POINTER(ctypes.c_float)]),
    "bd_server_create": (_P, [_I, _I, ctypes.POINTER(ServerClass), ctypes.POINTER(Config),
                              ctypes.POINTER(_I)]),
    "bd_server_create_resident": (_P, [_I, _I, ctypes.POINTER(ServerClass), ctypes.POINTER(Config),
                                       ctypes.c_double, ctypes.POINTER(_I)]),
    "bd_server_destroy": (None, [_P]),
    "bd_server_submit": (ctypes.c_int64, [_P, _I, _P, _I]),
    "bd_server_poll": (_I, [_P, ctypes.POINTER(ServerResult), _I, ctypes.POINTER(Detection), _I]),
    "bd_server_wait_idle": (_I, [_P, _I]),
    "bd_server_simulate": (_I, [_P, ctypes.POINTER(StreamSpec), _I, ctypes.c_double, ctypes.c_uint64,
                                ctypes.POINTER(ServeReport)]),
    "bd_server_last_error": (ctypes.c_char_p, []),
    "bd_detect_device_noalloc": (_I, [_I, _I, _I, ctypes.POINTER(Config), _P, _I, _P, _I, _P, _P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: the HIP detector is required (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().bd_last_error()
    return msg.decode() if msg else ""


def default_config(**overrides) -> Config:
    c = Config()
    lib().bd_config_init(ctypes.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def detection_to_dict(d: Detection) -> dict:
    return {
        "id": int(d.id), "hamming": int(d.hamming), "decision_margin": float(d.decision_margin),
        "corners": np.array(d.corners[:], dtype=np.float32).reshape(4, 2),
        "homography": np.array(d.homography[:], dtype=np.float32),
    }


# bd_detection as a numpy record (84 bytes, the C layout).
DETECTION_DTYPE = np.dtype([("id", "<i4"), ("hamming", "<i4"), ("decision_margin", "<f4"),
                            ("corners", "<f4", (4, 2)), ("reserved", "<i4"), ("homography", "<f4", (9,))])
assert DETECTION_DTYPE.itemsize == ctypes.sizeof(Detection)


def detections_from(buf, first: int, n: int) -> list:
    """Records first .. first + n - 1 of a ctypes Detection array as dicts,
    from one numpy copy of those records (the dicts' arrays are views of
    the copy, so the buffer may be reused at once)."""
    if n <= 0:
        return []
    recs = np.frombuffer(buf, dtype=DETECTION_DTYPE, count=first + n)[first:].copy()
    ids, hams, margins = recs["id"].tolist(), recs["hamming"].tolist(), recs["decision_margin"].tolist()
    corners, homs = recs["corners"], recs["homography"]
    return [{"id": ids[k], "hamming": hams[k], "decision_margin": margins[k],
             "corners": corners[k], "homography": homs[k]} for k in range(n)]


def _stats_dict(s: Stats) -> dict:
    return {k: getattr(s, k) for k, _ in Stats._fields_}


class Batch:
    """Several frames of fixed (possibly different) sizes per call."""

    def __init__(self, family: int, sizes, handle=None, owner=None, **config):
        self._owner = owner
        if handle is not None:
            self.handle = handle
            self._owned = False
        else:
            sizes = list(sizes)
            w = (ctypes.c_int32 * len(sizes))(*[s[0] for s in sizes])
            h = (ctypes.c_int32 * len(sizes))(*[s[1] for s in sizes])
            status = ctypes.c_int32(0)
            cfg = default_config(**config)
            self.handle = lib().bd_batch_create(family, len(sizes), w, h, ctypes.byref(cfg),
                                                ctypes.byref(status))
            if not self.handle:
                raise RuntimeError(f"bd_batch_create failed ({status.value}): {last_error()}")
            self._owned = True
        info = (ctypes.c_int32 * 8)()
        self.frames = []
        i = 0
        while lib().bd_batch_frame_info(self.handle, i, info) == BD_OK:
            self.frames.append({
                "width": info[0], "height": info[1], "tile_w": info[2], "tile_h": info[3],
                "words64": info[4], "row_cap": info[5], "max_candidates": info[6],
                "max_detections": info[7]})
            i += 1
        self.max_detections = self.frames[0]["max_detections"]
        self.family = family
        self.sizes = [(f["width"], f["height"]) for f in self.frames]

    def close(self):
        if getattr(self, "_owned", False) and self.handle:
            lib().bd_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        self.close()

    def _run(self, frames, device: bool):
        n = len(self.frames)
        arr = (Frame * n)()
        keep = []
        for i, f in enumerate(frames):
            if device:
                ptr, stride = f
            else:
                img = np.ascontiguousarray(f, dtype=np.uint8)
                keep.append(img)
                ptr, stride = img.ctypes.data, img.strides[0]
            arr[i].gray = ptr
            arr[i].width = self.frames[i]["width"]
            arr[i].height = self.frames[i]["height"]
            arr[i].stride = stride
        md = self.max_detections
        out = (Detection * (n * md))()
        counts = (ctypes.c_int32 * n)()
        fn = lib().bd_batch_detect_device if device else lib().bd_batch_detect_host
        rc = fn(self.handle, arr, n, out, md, counts)
        if rc < 0:
            raise RuntimeError(f"batch detect failed ({rc}): {last_error()}")
        return [detections_from(out, i * md, min(counts[i], md)) for i in range(n)]

    def detect(self, images):
        return self._run(images, False)

    def prepare_device(self, ptr_strides):
        """Pre-builds the frame table and output buffers for detect_raw()."""
        n = len(self.frames)
        arr = (Frame * n)()
        for i, (ptr, stride) in enumerate(ptr_strides):
            arr[i].gray = ptr
            arr[i].width = self.frames[i]["width"]
            arr[i].height = self.frames[i]["height"]
            arr[i].stride = stride
        md = self.max_detections
        self._raw = (arr, (Detection * (n * md))(), (ctypes.c_int32 * n)(), n, md)

    def detect_raw(self) -> int:
        """One bd_batch_detect_device call on the prepared frames; results stay
        in the prepared C buffers (no Python conversion).  Returns the total
        number of records written."""
        arr, out, counts, n, md = self._raw
        rc = lib().bd_batch_detect_device(self.handle, arr, n, out, md, counts)
        if rc < 0:
            raise RuntimeError(f"batch detect failed ({rc}): {last_error()}")
        return rc

    def detect_device(self, ptr_strides):
        return self._run(ptr_strides, True)

    def stats(self, index=0) -> dict:
        s = Stats()
        lib().bd_batch_get_stats(self.handle, index, ctypes.byref(s))
        return _stats_dict(s)

    def debug(self, stage: str, index: int = 0) -> np.ndarray:
        """Raw copy of one stage buffer of frame `index` of the last call."""
        code = STAGES[stage]
        n = lib().bd_batch_debug_copy(self.handle, index, code, None, 0)
        if n < 0:
            raise RuntimeError(f"bd_batch_debug_copy failed: {last_error()}")
        buf = np.empty(int(n), dtype=np.uint8)
        rc = lib().bd_batch_debug_copy(self.handle, index, code, buf.ctypes.data, int(n))
        if rc < 0:
            raise RuntimeError(f"bd_batch_debug_copy failed: {last_error()}")
        return buf

    def capture_stages(self, on: bool = True) -> None:
        """Stage capture: later calls also keep the runs and labels of the
        rows inside each band, so debug() shows every run (a run's label is
        its band root; a band root's label is its component root)."""
        if lib().bd_batch_set_capture(self.handle, 1 if on else 0) != BD_OK:
            raise RuntimeError(f"bd_batch_set_capture failed: {last_error()}")

    def time_stages(self, reps: int = 10) -> dict:
        ms = (ctypes.c_float * (len(KERNELS) + 1))()
        rc = lib().bd_batch_time_stages(self.handle, reps, ms)
        if rc != BD_OK:
            raise RuntimeError(f"bd_batch_time_stages failed: {last_error()}")
        out = {k: float(ms[i]) for i, k in enumerate(KERNELS)}
        out["pipeline"] = float(ms[len(KERNELS)])
        return out

    def replay(self, reps: int, device: bool = False) -> float:
        """Total device ms of `reps` back-to-back graph replays."""
        ms = ctypes.c_float()
        rc = lib().bd_batch_replay(self.handle, reps, 1 if device else 0, ctypes.byref(ms))
        if rc != BD_OK:
            raise RuntimeError(f"bd_batch_replay failed: {last_error()}")
        return float(ms.value)

    def get_tune(self) -> dict:
        t = Tune()
        if lib().bd_batch_get_tune(self.handle, ctypes.byref(t)) != BD_OK:
            raise RuntimeError(f"bd_batch_get_tune failed: {last_error()}")
        return {k: getattr(t, k) for k in TUNE_KEYS}

    def set_tune(self, fit_grid: int, decode_grid: int, head_ranks: int, streams: int = 2,
                 decode_waves: int = 8, chunk_px: int = 64, head_area: int = 1000, band_rows: int = 8) -> None:
        t = Tune(fit_grid, decode_grid, head_ranks, streams, decode_waves, chunk_px, head_area, band_rows)
        rc = lib().bd_batch_set_tune(self.handle, ctypes.byref(t))
        if rc != BD_OK:
            raise (ValueError if rc == BD_ERROR_INVALID_ARGUMENT else RuntimeError)(
                f"bd_batch_set_tune failed: {last_error()}")

    def signature(self) -> str:
        """Key of this batch's shape for a tuning cache."""
        return f"family{self.family}|" + ",".join(
            f"{f['width']}x{f['height']}c{f['max_candidates']}" for f in self.frames)

    def autotune(self, reps: int = 10, cache: str | None = None) -> dict:
        """Picks the fastest launch parameters for the frames of the last
        detect call (bd_batch_autotune).  With `cache` (a JSON file), a
        setting stored for this batch shape is applied instead of measuring,
        and a new measurement is stored."""
        import json
        key = self.signature()
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
            if key in table:
                t = table[key]
                self.set_tune(t["fit_grid"], t["decode_grid"], t["head_ranks"], t.get("streams", 2),
                              t.get("decode_waves", 8), t.get("chunk_px", 64), t.get("head_area", 1000),
                              t.get("band_rows", 8))
                return dict(t, cached=True)
        best = Tune()
        ms = ctypes.c_float()
        rc = lib().bd_batch_autotune(self.handle, reps, ctypes.byref(best), ctypes.byref(ms))
        if rc != BD_OK:
            raise RuntimeError(f"bd_batch_autotune failed: {last_error()}")
        out = {k: getattr(best, k) for k in TUNE_KEYS}
        out["pipeline_ms"] = round(float(ms.value), 4)
        if cache:
            table[key] = out
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        return dict(out, cached=False)


class Detector:
    """One family, one frame size (the reference's bd_detector)."""

    def __init__(self, family: int, width: int, height: int, **config):
        status = ctypes.c_int32(0)
        cfg = default_config(**config)
        self.width, self.height = width, height
        self.handle = lib().bd_detector_create(family, width, height, ctypes.byref(cfg),
                                               ctypes.byref(status))
        if not self.handle:
            raise RuntimeError(f"bd_detector_create failed ({status.value}): {last_error()}")
        self.status = status.value
        self.max_detections = max(1, min(cfg.max_detections, 256))
        self.out = (Detection * self.max_detections)()
        self.batch = Batch(family, None, handle=lib().bd_detector_batch(self.handle), owner=self)

    def close(self):
        if getattr(self, "handle", None):
            lib().bd_detector_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def detect(self, image: np.ndarray, stride: int | None = None):
        img = np.ascontiguousarray(image, dtype=np.uint8) if stride is None else image
        if stride is None:
            stride = img.strides[0]
        rc = lib().bd_detector_detect_host(self.handle, img.ctypes.data, stride, self.out,
                                           self.max_detections)
        if rc < 0:
            raise RuntimeError(f"bd_detector_detect_host failed ({rc}): {last_error()}")
        return detections_from(self.out, 0, min(rc, self.max_detections))

    def detect_device(self, ptr: int, stride: int):
        rc = lib().bd_detector_detect_device(self.handle, ptr, stride, self.out, self.max_detections)
        if rc < 0:
            raise RuntimeError(f"bd_detector_detect_device failed ({rc}): {last_error()}")
        return detections_from(self.out, 0, min(rc, self.max_detections))

    def stats(self) -> dict:
        s = Stats()
        lib().bd_detector_get_stats(self.handle, ctypes.byref(s))
        return _stats_dict(s)


def detect_device_noalloc(family: int, width: int, height: int, gray_device: int, stride: int,
                          out_device: int, out_capacity: int, count_device: int, stream: int | None = None,
                          **config) -> int:
    """bd_detect_device_noalloc: queues one device frame on `stream`; records
    and count land in the given device buffers.  Returns BD_OK or raises."""
    cfg = default_config(**config)
    rc = lib().bd_detect_device_noalloc(family, width, height, ctypes.byref(cfg), gray_device, stride,
                                        out_device, out_capacity, count_device, stream)
    if rc != BD_OK:
        raise (ValueError if rc == BD_ERROR_INVALID_ARGUMENT else RuntimeError)(
            f"bd_detect_device_noalloc failed ({rc}): {last_error()}")
    return rc


class Server:
    """Continuous-batching server over device-resident frames.

    classes: list of dicts {width, height, lanes, lane_frames, queue_depth}.
    resident=True: admission on the device (bd_server_create_resident), each
    lane launch waiting up to admit_wait_us for requests.
    """

    def __init__(self, family: int, classes, resident: bool = False, admit_wait_us: float = 500.0, **config):
        arr = (ServerClass * len(classes))()
        for i, c in enumerate(classes):
            arr[i] = ServerClass(c["width"], c["height"], c["lanes"], c["lane_frames"],
                                 c["queue_depth"], 0)
        self.classes = [dict(c) for c in classes]
        status = ctypes.c_int32(0)
        cfg = default_config(**config)
        self.max_detections = max(1, min(cfg.max_detections, 256))
        if resident:
            self.handle = lib().bd_server_create_resident(family, len(classes), arr, ctypes.byref(cfg),
                                                          admit_wait_us, ctypes.byref(status))
        else:
            self.handle = lib().bd_server_create(family, len(classes), arr, ctypes.byref(cfg),
                                                 ctypes.byref(status))
        self.resident = resident
        if not self.handle:
            msg = lib().bd_server_last_error()
            raise RuntimeError(f"bd_server_create failed ({status.value}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "handle", None):
            lib().bd_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def submit(self, class_index: int, ptr: int, stride: int) -> int:
        """Request id (>= 0), or a negative status when dropped."""
        return int(lib().bd_server_submit(self.handle, class_index, ptr, stride))

    def poll(self, max_results: int = 256):
        res = (ServerResult * max_results)()
        md = self.max_detections
        dets = (Detection * (max_results * md))()
        n = lib().bd_server_poll(self.handle, res, max_results, dets, md)
        if n < 0:
            raise RuntimeError(f"bd_server_poll failed ({n})")
        out = []
        for i in range(n):
            r = res[i]
            out.append({"request": int(r.request), "class_index": int(r.class_index),
                        "count": int(r.count), "batch_frames": int(r.batch_frames),
                        "latency_ms": float(r.latency_ms),
                        "detections": [detection_to_dict(dets[i * md + k]) for k in range(r.written)]})
        return out

    def wait_idle(self, timeout_ms: int = 10000) -> bool:
        return lib().bd_server_wait_idle(self.handle, timeout_ms) == BD_OK

    def simulate(self, streams, seconds: float, seed: int = 0) -> dict:
        """streams: list of dicts {class_index, fps, frames: [device ptrs], stride}."""
        arr = (StreamSpec * len(streams))()
        keep = []
        for i, st in enumerate(streams):
            ptrs = (ctypes.c_void_p * len(st["frames"]))(*st["frames"])
            keep.append(ptrs)
            arr[i] = StreamSpec(st["class_index"], len(st["frames"]), float(st["fps"]),
                                ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p)), st["stride"], 0)
        rep = ServeReport()
        rc = lib().bd_server_simulate(self.handle, arr, len(streams), seconds, seed, ctypes.byref(rep))
        if rc != BD_OK:
            raise RuntimeError(f"bd_server_simulate failed ({rc})")
        return {k: getattr(rep, k) for k, _ in ServeReport._fields_}
