"""The scene generator: the port of the reference's benchmark scenes is
pixel-identical to the reference's own exporter (built from its sources into
oracle/_ref when /root/reference is present), and the workload generator is
deterministic and its ground truth matches what the oracle finds."""
import json
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXPORTER = os.path.join(ROOT, "oracle", "_ref", "bdcudatag_scenes")


def read_pgm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], dtype=np.uint8).reshape(h, w)


@pytest.mark.skipif(not os.path.exists(REF_EXPORTER), reason="reference exporter not built here")
def test_bench_scenes_match_reference_exporter(tmp_path):
    subprocess.run([REF_EXPORTER, str(tmp_path), "4"], check=True, capture_output=True)
    for res in ("640x480", "1280x800"):
        w, h = map(int, res.split("x"))
        manifest = json.load(open(tmp_path / f"scenes_{res}.json"))
        for i, sc in enumerate(manifest["scenes"]):
            ref = read_pgm(tmp_path / sc["file"])
            img, tags = scenes.bench_scene(w, h, i)
            assert np.array_equal(ref, img)
            assert [t["id"] for t in sc["tags"]] == [t["id"] for t in tags]
            gt = np.array([t["corners"] for t in sc["tags"]])
            assert np.abs(gt - np.array([t["corners"] for t in tags])).max() < 6e-4  # %.3f


def test_workload_scenes_are_deterministic():
    a, ta = scenes.scene(0, 1280, 800, 17)
    b, tb = scenes.scene(0, 1280, 800, 17)
    c, _ = scenes.scene(0, 1280, 800, 18)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert [t["id"] for t in ta] == [t["id"] for t in tb]
    batch = scenes.scene_batch(0, 1280, 800, 17, 2, threads=2)
    assert np.array_equal(batch[0], a) and np.array_equal(batch[1], c)


def test_workload_eight_tags_ground_truth():
    found = total = wrong = 0
    errs = []
    for seed in range(1, 4):
        img, tags = scenes.scene(0, 1280, 800, seed)
        real = [t for t in tags if not t["decoy"]]
        assert len(real) == 8
        d = pyoracle.Oracle(pyoracle.default_config(0, 1280, 800)).run(img)["detections"]
        ids = [t["id"] for t in real]
        wrong += sum(1 for x in d if x["id"] not in ids)
        for t in real:
            total += 1
            for x in d:
                if x["id"] == t["id"]:
                    e = np.abs(x["corners"] - t["corners"]).max()
                    if e < 2.0:
                        found += 1
                        errs.append(e)
                        break
    assert wrong == 0
    assert found >= total - 1
    assert np.median(errs) < 0.5


def test_scene_set_files_match_reference_exporter_bytes(tmp_path):
    """The Python writer produces the reference exporter's files byte for
    byte (manifest text and PGMs) when that exporter is built here; either
    way the files round-trip through the harness-style reader."""
    ours = tmp_path / "ours"
    scenes.export_reference_sets(str(ours), 2)
    sets = scenes.read_scene_sets(str(ours))
    assert [s["label"] for s in sets] == ["640x480", "1280x720", "1280x800", "1920x1080"]
    for s in sets:
        for i, sc in enumerate(s["scenes"]):
            img, tags = scenes.bench_scene(s["width"], s["height"], i)
            assert np.array_equal(sc["image"], img)
            assert [t["id"] for t in sc["tags"]] == [t["id"] for t in tags]
    if os.path.exists(REF_EXPORTER):
        ref = tmp_path / "ref"
        subprocess.run([REF_EXPORTER, str(ref), "2"], check=True, capture_output=True)
        names = sorted(os.listdir(ref))
        assert names == sorted(os.listdir(ours))
        for name in names:
            assert (ref / name).read_bytes() == (ours / name).read_bytes(), name


def test_napa_metadata_layout(tmp_path):
    items = [scenes.scene(0, 640, 480, s) for s in range(2)]
    path = scenes.write_napa_group(str(tmp_path), "synthetic_small", items, scenario="eight_tags")
    lines = [json.loads(x) for x in open(path)]
    assert len(lines) == 2
    for rec, (img, tags) in zip(lines, items):
        assert set(rec) == {"file_name", "width", "height", "scenario", "tag_ids", "corners"}
        assert (rec["width"], rec["height"]) == (640, 480)
        real = [t for t in tags if not t["decoy"]]
        assert rec["tag_ids"] == [t["id"] for t in real]
        assert np.abs(np.array(rec["corners"]) - np.array([t["corners"] for t in real])).max() <= 5e-4
        back = scenes.read_pgm(str(tmp_path / "data" / "synthetic_small" / rec["file_name"]))
        assert np.array_equal(back, img)


def test_harness_scoring_rules():
    img, tags = scenes.bench_scene(640, 480, 0)
    truth = [{"id": t["id"], "corners": t["corners"]} for t in tags]
    # exact answers, corners in another order: found, zero error
    dets = [{"id": t["id"], "corners": np.roll(t["corners"], 1, axis=0)} for t in tags]
    found, total, wrong, errs = scenes.score_scene(truth, dets)
    assert (found, total, wrong) == (len(tags), len(tags), 0) and max(errs) < 1e-9
    # a 1 px shift shows up as 1 px mean corner error; an absent id is wrong;
    # a detection far from its tag does not count
    shifted = [{"id": tags[0]["id"], "corners": tags[0]["corners"] + [1.0, 0.0]},
               {"id": 9999, "corners": tags[1]["corners"]},
               {"id": tags[2]["id"], "corners": tags[2]["corners"] + 500.0}]
    found, total, wrong, errs = scenes.score_scene(truth, shifted)
    assert (found, wrong) == (1, 1) and abs(errs[0] - 1.0) < 1e-9
    # the oracle on the reference's own scene: every tag, no wrong id
    o = pyoracle.Oracle(pyoracle.default_config(0, 640, 480, max_detections=32)).run(img)
    found, total, wrong, errs = scenes.score_scene(truth, o["detections"])
    assert found == total and wrong == 0 and np.median(errs) < 0.5
