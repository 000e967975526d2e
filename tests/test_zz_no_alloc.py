"""bd_detect_device_noalloc: the entry point that allocates nothing.

Its records and count, read back from the caller's device buffers, must be
bit-identical to bd_detector_detect_device's on the same frame, on two calls
in a row (the second with a smaller frame, so the array keeps the first
call's contents), with an output capacity below the count, and with
tag16h5.  Oversized frames and bad arguments are rejected on the host
before any GPU work.
"""
import ctypes

import numpy as np
import pytest

import bdcudatag
import scenes

REC = ctypes.sizeof(bdcudatag.Detection)


def records(raw: np.ndarray, n: int):
    arr = (bdcudatag.Detection * max(n, 1)).from_buffer_copy(raw.tobytes()[:max(n, 1) * REC])
    return [bdcudatag.detection_to_dict(arr[i]) for i in range(n)]


def same(a, b):
    assert [d["id"] for d in a] == [d["id"] for d in b]
    for x, y in zip(a, b):
        assert x["hamming"] == y["hamming"]
        assert np.float32(x["decision_margin"]) == np.float32(y["decision_margin"])
        assert np.array_equal(x["corners"].view(np.uint32), y["corners"].view(np.uint32))
        assert np.array_equal(x["homography"].view(np.uint32), y["homography"].view(np.uint32))


def test_rejections_on_the_host():
    # No GPU work: argument checks and the working-memory limit come first.
    L = bdcudatag.lib()
    count = ctypes.c_int32()
    out = (bdcudatag.Detection * 4)()
    cfg = bdcudatag.default_config()
    args = (ctypes.byref(cfg), 0x1000, 4096, out, 4, ctypes.byref(count), None)
    assert L.bd_detect_device_noalloc(0, 4096, 3072, *args) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert "working memory" in bdcudatag.last_error() and "1920x1200" in bdcudatag.last_error()
    assert L.bd_detect_device_noalloc(5, 640, 480, *args) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert bdcudatag.last_error() == "unknown tag family"
    assert L.bd_detect_device_noalloc(0, 640, 480, ctypes.byref(cfg), None, 640, out, 4, ctypes.byref(count),
                                      None) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert bdcudatag.last_error() == "null device image pointer"
    assert L.bd_detect_device_noalloc(0, 640, 480, ctypes.byref(cfg), 0x1000, 600, out, 4, ctypes.byref(count),
                                      None) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert L.bd_detect_device_noalloc(0, 640, 480, ctypes.byref(cfg), 0x1000, 640, out, 4, None,
                                      None) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert bdcudatag.last_error() == "null device count pointer"


@pytest.mark.gpu
def test_matches_detect_device_two_calls_then_smaller():
    import devmem
    cases = [(0, 1280, 800, scenes.scene(0, 1280, 800, 3)[0], 64),
             (0, 640, 480, scenes.scene(0, 640, 480, 5)[0], 64),       # smaller second call
             (0, 640, 480, scenes.scene(0, 640, 480, 6)[0], 3),        # capacity below the count
             (1, 1920, 1080, scenes.scene(1, 1920, 1080, 7)[0], 64)]   # tag16h5, dense frame
    for family, W, H, img, cap in cases:
        print(f"  noalloc family {family} {W}x{H} capacity {cap}", flush=True)
        frame = devmem.DeviceBuffer(img)
        out = devmem.DeviceBuffer(np.full(64 * REC, 0xAB, dtype=np.uint8))
        cnt = devmem.DeviceBuffer(np.full(4, 0xFF, dtype=np.uint8))
        bdcudatag.detect_device_noalloc(family, W, H, frame.ptr, W, out.ptr, cap, cnt.ptr)
        devmem.sync()
        n = int(cnt.download().view(np.int32)[0])
        raw = out.download()
        want = bdcudatag.Detector(family, W, H).detect_device(frame.ptr, W)
        assert n == len(want) and n > 0
        written = min(n, cap)
        same(records(raw, written), want[:written])
        # Nothing written past the capacity.
        assert (raw[written * REC:] == 0xAB).all()
