"""The autotuner's table: keys, nearest-signature lookup per precision."""
import json

from olcao_amd import autotune, configs


def test_lookup_takes_nearest_signature_of_the_same_precision(tmp_path):
    s = configs.small_test()
    sig = autotune.signature(s)
    table = {
        "a": {"signature": dict(sig, n=sig["n"] * 10), "precision": 0,
              "shapes": {"chunk_items": 768, "seg_pairs": 64, "bloch": 1}},
        "b": {"signature": dict(sig), "precision": 1,
              "shapes": {"chunk_items": 64, "seg_pairs": 8, "bloch": 2}},
        "c": {"signature": dict(sig, nk=sig["nk"] * 2), "precision": 0,
              "shapes": {"chunk_items": 128, "seg_pairs": 16, "bloch": 2}},
    }
    p = tmp_path / "t.json"
    p.write_text(json.dumps(table))
    assert autotune.lookup(s, 0, str(p)) == table["c"]["shapes"]
    assert autotune.lookup(s, 1, str(p)) == table["b"]["shapes"]
    assert autotune.lookup(s, 2, str(p)) is None
    assert autotune.lookup(s, 0, str(tmp_path / "missing.json")) is None


def test_committed_table_is_well_formed():
    for key, e in autotune.load().items():
        assert key == autotune._key(e["signature"], e["precision"])
        sh = e["shapes"]
        assert sh["chunk_items"] in autotune.CANDIDATES["chunk_items"]
        assert sh["seg_pairs"] in autotune.CANDIDATES["seg_pairs"]
        assert sh["bloch"] in autotune.CANDIDATES["bloch"]
        assert e["build_ms"] <= e["default_ms"]
