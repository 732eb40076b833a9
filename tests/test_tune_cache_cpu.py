"""The tuner's cache of ring-pass shapes (pfsht_tune_*): host functions, no GPU needed.

Entries are keyed by (architecture, lmax, rings, largest order, ntrans); a cache file written by
pfsht_tune_save merges back with pfsht_tune_load; malformed shapes and files are refused.
"""
import os

import pytest


@pytest.fixture
def P():
    import pfsht
    pfsht.lib()
    pfsht.tune_clear()
    yield pfsht
    pfsht.tune_clear()


def test_put_get_and_keys(P):
    assert P.tune_size() == 0
    assert P.tune_get("gfx950:256", 2048, 360, 1640, 1) is None
    P.tune_put("gfx950:256", 2048, 360, 1640, 1, seg_blocks=4, adj_tiles=1, mfma_tiles=0, ms=0.25)
    assert P.tune_get("gfx950:256", 2048, 360, 1640, 1) == {"seg_blocks": 4, "adj_tiles": 1, "mfma_tiles": 0,
                                                            "ms": 0.25}
    # every key field matters
    for k in [("gfx942:304", 2048, 360, 1640, 1), ("gfx950:256", 1024, 360, 1640, 1),
              ("gfx950:256", 2048, 359, 1640, 1), ("gfx950:256", 2048, 360, 1639, 1),
              ("gfx950:256", 2048, 360, 1640, 2)]:
        assert P.tune_get(*k) is None
    # a later entry replaces an earlier one
    P.tune_put("gfx950:256", 2048, 360, 1640, 1, seg_blocks=2, adj_tiles=2, ms=0.2)
    assert P.tune_get("gfx950:256", 2048, 360, 1640, 1)["seg_blocks"] == 2
    assert P.tune_size() == 1


def test_save_load_round_trip(P, tmp_path):
    keys = [("gfx950:256", 256, 45, 230, 1, 3, 2, 0), ("gfx950:256", 1024, 180, 820, 16, 2, 2, 2),
            ("gfx950:256", 2048, 360, 1640, 1, 6, 1, 0)]
    for a, L, n, mm, B, per, atg, mtg in keys:
        P.tune_put(a, L, n, mm, B, seg_blocks=per, adj_tiles=atg, mfma_tiles=mtg, ms=0.125 * per)
    path = tmp_path / "tune.txt"
    P.tune_save(str(path))
    assert len(path.read_text().splitlines()) == 3
    P.tune_clear()
    assert P.tune_size() == 0
    P.tune_load(str(path))
    assert P.tune_size() == 3
    for a, L, n, mm, B, per, atg, mtg in keys:
        assert P.tune_get(a, L, n, mm, B) == {"seg_blocks": per, "adj_tiles": atg, "mfma_tiles": mtg,
                                              "ms": 0.125 * per}


def test_bad_shapes_and_files_refused(P, tmp_path):
    with pytest.raises(P.ArgumentError):
        P.tune_put("gfx950:256", 256, 45, 230, 1, seg_blocks=65)
    with pytest.raises(P.ArgumentError):
        P.tune_put("gfx950:256", 256, 45, 230, 1, adj_tiles=3)
    with pytest.raises(P.ArgumentError):
        P.tune_put("", 256, 45, 230, 1)
    with pytest.raises(P.ArgumentError):
        P.tune_load(str(tmp_path / "missing.txt"))
    bad = tmp_path / "bad.txt"
    bad.write_text("gfx950:256 256 45 230 1 3 2\n")  # a field short
    with pytest.raises(P.ArgumentError):
        P.tune_load(str(bad))
    assert not os.path.exists(str(tmp_path / "x.tmp"))
