"""The C-ABI library loads and exports every symbol the public header
declares, with the reference's structure layouts and argument validation.
None of these calls reaches the GPU."""
import ctypes
import os
import re

import pytest

import bdcudatag

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "bdcudatag", "bdcudatag.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"BD_API\s+[\w\s\*]+?\b(bd_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 24
    lib = bdcudatag.lib()
    for s in syms:
        assert hasattr(lib, s), s
        assert s in bdcudatag.SIGNATURES, f"{s} missing from the Python binding"


def test_reference_v2_symbols_present():
    # detector/include/bdcudatag/bdcudatag.h:103-145 (14 symbols)
    v2 = ["bd_api_version", "bd_version_string", "bd_family_count", "bd_family_name",
          "bd_family_bits_per_side", "bd_family_min_hamming", "bd_family_default_max_hamming",
          "bd_config_init", "bd_detector_create", "bd_detector_destroy", "bd_detector_detect_host",
          "bd_detector_detect_device", "bd_detector_get_stats", "bd_last_error"]
    assert set(v2) <= set(declared_symbols())


def test_layouts():
    assert ctypes.sizeof(bdcudatag.Config) == 40
    assert ctypes.sizeof(bdcudatag.Detection) == 84
    assert ctypes.sizeof(bdcudatag.Stats) == 40


def test_family_queries_and_defaults():
    L = bdcudatag.lib()
    assert L.bd_api_version() == 3
    assert L.bd_family_count() == 2
    assert L.bd_family_name(0) == b"tag36h11" and L.bd_family_name(1) == b"tag16h5"
    assert L.bd_family_name(2) is None
    assert (L.bd_family_bits_per_side(0), L.bd_family_bits_per_side(1)) == (6, 4)
    assert (L.bd_family_min_hamming(0), L.bd_family_min_hamming(1)) == (11, 5)
    assert (L.bd_family_default_max_hamming(0), L.bd_family_default_max_hamming(1)) == (2, 1)
    assert L.bd_family_bits_per_side(-1) == -1
    c = bdcudatag.default_config()
    # detector/src/detector.cpp:97-110
    assert (c.struct_size, c.max_hamming, c.min_cluster_pixels, c.min_tag_pixels, c.max_detections,
            c.max_candidates, c.refine_edges, c.use_cuda_graph) == (40, -1, 24, 8, 64, 0, 1, 1)
    assert c.min_decision_margin == pytest.approx(3.0) and c.decode_sharpening == pytest.approx(0.25)


@pytest.mark.parametrize("family,w,h,size", [(99, 640, 480, 40), (0, 0, 480, 40), (0, 640, 8, 40),
                                             (0, 9000, 480, 40), (-1, 640, 480, 40), (0, 640, 480, 12)])
def test_create_rejects_invalid_arguments(family, w, h, size):
    L = bdcudatag.lib()
    cfg = bdcudatag.default_config()
    cfg.struct_size = size
    status = ctypes.c_int32(0)
    assert not L.bd_detector_create(family, w, h, ctypes.byref(cfg), ctypes.byref(status))
    assert status.value == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert bdcudatag.last_error() != ""


def test_null_handles_are_rejected():
    L = bdcudatag.lib()
    out = (bdcudatag.Detection * 4)()
    assert L.bd_detector_detect_host(None, None, 640, out, 4) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    assert bdcudatag.last_error() == "null detector"
    assert L.bd_detector_detect_device(None, None, 640, out, 4) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    L.bd_detector_destroy(None)
    L.bd_detector_get_stats(None, None)
    frames = (bdcudatag.Frame * 1)()
    counts = (ctypes.c_int32 * 1)()
    assert L.bd_batch_detect_host(None, frames, 1, out, 4, counts) == bdcudatag.BD_ERROR_INVALID_ARGUMENT
    w = (ctypes.c_int32 * 1)(640)
    h = (ctypes.c_int32 * 1)(480)
    status = ctypes.c_int32(0)
    assert not L.bd_batch_create(0, 0, w, h, None, ctypes.byref(status))
    assert status.value == bdcudatag.BD_ERROR_INVALID_ARGUMENT


def test_resident_server_rejects_invalid_arguments():
    # Checked before any GPU work.
    L = bdcudatag.lib()
    cls = (bdcudatag.ServerClass * 1)(bdcudatag.ServerClass(1280, 800, 2, 4, 64, 0))
    many = (bdcudatag.ServerClass * 1)(bdcudatag.ServerClass(1280, 800, 17, 4, 64, 0))
    status = ctypes.c_int32(0)
    for classes, n, wait in [(cls, 1, 0.0), (cls, 1, -5.0), (cls, 1, 2e6), (many, 1, 500.0), (cls, 0, 500.0)]:
        assert not L.bd_server_create_resident(0, n, classes, None, wait, ctypes.byref(status))
        assert status.value == bdcudatag.BD_ERROR_INVALID_ARGUMENT
