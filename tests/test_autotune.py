"""The autotuner's host logic: size classes, the candidate set and reading
the persisted MI355X table back (no GPU)."""
import json

from fluid_amd import autotune


def test_size_classes():
    assert autotune.size_class((384, 384, 384)) == 384
    assert autotune.size_class((64, 64, 64)) == 64
    assert autotune.size_class((63, 64, 64)) == 48
    assert autotune.size_class((16, 16, 16)) == 32
    assert autotune.size_class((256, 128, 64)) == 128  # 2^21 cells = 128^3
    assert autotune.size_class((1024, 1024, 1024)) == 512


def test_candidates_respect_the_tile_width():
    assert len(autotune.candidates((64, 64, 64))) == 36
    assert len(autotune.candidates((64, 64, 64), "fp32")) == 18
    narrow = autotune.candidates((64, 64, 24))
    assert len(narrow) == 18 and all(c["tile_z"] == 16 for c in narrow)


def test_persisted_table(tmp_path):
    table = autotune.load()
    for key in ("64", "96", "128", "192", "384"):
        rec = table[key]
        assert rec["best"]["sweeps"] in autotune.SWEEPS
        assert rec["best"]["tile_z"] in autotune.TILES
        assert rec["best"]["groups"] in autotune.GROUPS
        assert rec["best_ms"] <= rec["default_ms"]
        assert all(t["ok"] for t in rec["trials"])
    assert autotune.choice_for((384, 384, 384)) == dict(table["384"]["best"], pairs=table["384"]["best"].get("pairs", 0))
    # a 32-wide tile falls back to 16 on a grid thinner than 32 along z
    c = autotune.choice_for((128, 128, 24 * 8))
    assert c is not None
    thin = autotune.choice_for((512, 512, 16))
    assert thin is None or thin["tile_z"] == 16
    # save merges by size class
    p = tmp_path / "t.json"
    autotune.save([{"dims": [40, 40, 40], "best": {"sweeps": 2, "tile_z": 32, "groups": 8192}}], str(p))
    autotune.save([{"dims": [100, 100, 100], "best": {"sweeps": 1, "tile_z": 16, "groups": 4096}}], str(p))
    assert sorted(json.loads(p.read_text())) == ["32", "96"]
    assert autotune.choice_for((33, 33, 33), str(p))["groups"] == 8192
