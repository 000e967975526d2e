"""Serving: continuous batching over the preallocated arena, with the
scheduling on the host (bd_server_create) or the admission on the device
(bd_server_create_resident).

Each request's result must be the same alone (single-frame detector, device
pointer path) and in the mix (admitted while other requests are on the GPU,
batched with frames of other streams): bit-exact detections, and both equal
the oracle's.
"""
import numpy as np
import pytest

import bdcudatag
import pyoracle
import scenes

pytestmark = pytest.mark.gpu

CLASSES = [
    {"width": 1280, "height": 800, "lanes": 2, "lane_frames": 4, "queue_depth": 256},
    {"width": 1600, "height": 1200, "lanes": 2, "lane_frames": 2, "queue_depth": 256},
]


@pytest.fixture(scope="module")
def frames():
    import devmem
    out = []
    for ci, c in enumerate(CLASSES):
        imgs = scenes.scene_batch(4, c["width"], c["height"], 100 * ci, 10, threads=8)
        dev = devmem.DeviceBuffer(imgs)
        px = c["width"] * c["height"]
        out.append((imgs, dev, [dev.ptr + i * px for i in range(len(imgs))]))
    return out


@pytest.mark.parametrize("resident", [False, True])
def test_same_alone_and_in_the_mix(frames, resident):
    alone = []
    for ci, c in enumerate(CLASSES):
        det = bdcudatag.Detector(0, c["width"], c["height"])
        alone.append([det.detect_device(p, c["width"]) for p in frames[ci][2]])
        det.close()
        for img, got in zip(frames[ci][0], alone[ci]):
            want = pyoracle.Oracle(pyoracle.default_config(0, c["width"], c["height"])).run(img)["detections"]
            assert [d["id"] for d in got] == [d["id"] for d in want]
            for a, b in zip(got, want):
                assert np.array_equal(a["corners"], b["corners"])
                assert np.array_equal(a["homography"], b["homography"])
    srv = bdcudatag.Server(0, CLASSES, resident=resident)
    order = []
    for rep in range(3):
        for i in range(10):
            for ci, c in enumerate(CLASSES):
                rid = srv.submit(ci, frames[ci][2][i], c["width"])
                assert rid >= 0
                order.append((rid, ci, i))
    assert srv.wait_idle(20000)
    results = {}
    while True:
        got = srv.poll(64)
        if not got:
            break
        for r in got:
            results[r["request"]] = r
    assert len(results) == len(order)
    batched = 0
    for rid, ci, i in order:
        r = results[rid]
        want = alone[ci][i]
        assert r["count"] == len(want)
        batched += r["batch_frames"] > 1
        for a, b in zip(r["detections"], want):
            assert a["id"] == b["id"] and a["hamming"] == b["hamming"]
            assert np.array_equal(a["corners"], b["corners"])
            assert np.array_equal(a["homography"], b["homography"])
            assert np.float32(a["decision_margin"]) == np.float32(b["decision_margin"])
    assert batched > 0, "no request was batched with another"
    srv.close()


@pytest.mark.parametrize("resident", [False, True])
def test_simulated_streams_short(frames, resident):
    srv = bdcudatag.Server(0, CLASSES, resident=resident)
    streams = [{"class_index": 0, "fps": 120.0, "frames": frames[0][2], "stride": 1280} for _ in range(4)]
    streams += [{"class_index": 1, "fps": 60.0, "frames": frames[1][2], "stride": 1600} for _ in range(2)]
    rep = srv.simulate(streams, seconds=2.0, seed=1)
    assert rep["completed"] + rep["dropped"] == rep["offered"]
    assert rep["offered"] > 800
    assert rep["p50_ms"] > 0 and rep["p99_ms"] >= rep["p50_ms"]
    srv.close()


def test_resident_server_stops_when_idle():
    # Lanes cycle through bounded admission waits while nothing arrives;
    # destroying the server sets the stop flag, which ends every wait.
    import time
    srv = bdcudatag.Server(0, CLASSES, resident=True, admit_wait_us=2000.0)
    time.sleep(0.05)
    t0 = time.perf_counter()
    srv.close()
    assert time.perf_counter() - t0 < 1.0
