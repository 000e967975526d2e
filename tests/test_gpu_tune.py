"""Launch parameters (persistent grid sizes, largest-first scheduling)
change speed only: every setting must give bit-identical detections, and
the autotuner must leave a valid setting in place and honour its cache."""
import numpy as np
import pytest

import bdcudatag
import scenes

pytestmark = pytest.mark.gpu


def same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert [d["id"] for d in x] == [d["id"] for d in y]
        for p, q in zip(x, y):
            assert np.array_equal(p["corners"], q["corners"])
            assert np.array_equal(p["homography"], q["homography"])
            assert np.float32(p["decision_margin"]) == np.float32(q["decision_margin"])


def test_launch_parameters_keep_results(tmp_path):
    sizes = [(1280, 800)] * 4 + [(640, 480)] * 2 + [(1920, 1200)]
    imgs = [scenes.scene(2, w, h, 40 + i)[0] for i, (w, h) in enumerate(sizes)]
    batch = bdcudatag.Batch(0, sizes)
    # Default persistent grids are whole rounds of resident workgroups.
    t0 = batch.get_tune()
    assert t0["fit_grid"] > 0 and t0["decode_grid"] > 0
    assert {k: t0[k] for k in ("head_ranks", "streams", "decode_waves", "chunk_px", "head_area", "band_rows")} == \
        {"head_ranks": 2, "streams": 2, "decode_waves": 8, "chunk_px": 64, "head_area": 1000, "band_rows": 8}
    ref = batch.detect(imgs)
    assert sum(len(r) for r in ref) > 0
    # head_area 1 puts every candidate (up to 32 per frame) in the head
    # region, 2**31 - 1 none beyond the head ranks.
    for fg, dg, hr, ns, dw, cp, ha, br in [(64, 64, 0, 1, 4, 128, 1, 16), (256, 2048, 4, 2, 16, 64, 2**31 - 1, 8),
                                           (4096, 512, 1, 1, 8, 128, 500, 16), (1280, 1024, 2, 2, 4, 64, 1, 16),
                                           (1280, 100, 2, 1, 16, 128, 8000, 8)]:
        batch.set_tune(fg, dg, hr, ns, dw, cp, ha, br)
        assert batch.get_tune() == {"fit_grid": fg, "decode_grid": dg, "head_ranks": hr, "streams": ns,
                                    "decode_waves": dw, "chunk_px": cp, "head_area": ha, "band_rows": br}
        same(batch.detect(imgs), ref)
    for bad in [(10, 1024, 2), (1024, 1 << 20, 2), (1024, 1024, 5), (1024, 1024, 2, 3), (1024, 1024, 2, 2, 12),
                (1024, 1024, 2, 2, 8, 96), (1024, 1024, 2, 2, 8, 64, -5), (1024, 1024, 2, 2, 8, 64, 1000, 12)]:
        with pytest.raises(ValueError):
            batch.set_tune(*bad)
    cache = str(tmp_path / "tune.json")
    first = batch.autotune(reps=2, cache=cache)
    assert not first["cached"] and first["pipeline_ms"] > 0
    assert batch.get_tune() == {k: first[k] for k in bdcudatag.TUNE_KEYS}
    same(batch.detect(imgs), ref)
    again = batch.autotune(reps=2, cache=cache)
    assert again["cached"] and {k: again[k] for k in bdcudatag.TUNE_KEYS} == {k: first[k] for k in bdcudatag.TUNE_KEYS}


def test_one_frame_batch_runs_one_lane():
    batch = bdcudatag.Batch(0, [(640, 480)])
    assert batch.get_tune()["streams"] == 1
    img = scenes.scene(2, 640, 480, 7)[0]
    ref = batch.detect([img])
    batch.set_tune(1280, 1024, 2, 2)  # accepted; a single frame still runs as one lane
    assert batch.get_tune()["streams"] == 1
    same(batch.detect([img]), ref)
