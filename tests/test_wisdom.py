"""The autotuner's JSON wisdom file through the C ABI (no GPU): store, look up, replace,
keys, malformed files."""
import json

import pytest

import mxr


def _t(**kw):
    t = mxr.Tuning()
    base = dict(zdirect_kmax=16, zdirect_rpb=4, live_tiles=1, col_tc_x=16, col_tc_y=16)
    base.update(kw)
    for k, v in base.items():
        setattr(t, k, v)
    return t


def test_missing_file_has_no_entry(tmp_path):
    assert mxr.wisdom_lookup(str(tmp_path / "none.json"), "gfx950", 256, 512, 1) is None


def test_store_lookup_replace(tmp_path):
    p = str(tmp_path / "w.json")
    mxr.wisdom_store(p, "gfx950", 256, 512, 1, 0, _t(), 41.1)
    mxr.wisdom_store(p, "gfx950", 256, 256, 1, 0, _t(zdirect_kmax=8, col_tc_y=8), 6.0)
    mxr.wisdom_store(p, "gfx950", 256, 512, 0, 0, _t(zdirect_kmax=0, col_tc_x=4), 80.0)
    got = mxr.wisdom_lookup(p, "gfx950", 256, 256, 1)
    assert got.as_dict() == _t(zdirect_kmax=8, col_tc_y=8).as_dict()
    assert mxr.wisdom_lookup(p, "gfx950", 256, 512, 0).col_tc_x == 4
    # every key field distinguishes entries
    assert mxr.wisdom_lookup(p, "gfx950", 256, 512, 1, spin=1) is None
    assert mxr.wisdom_lookup(p, "gfx950", 304, 512, 1) is None
    assert mxr.wisdom_lookup(p, "gfx942", 256, 512, 1) is None
    assert mxr.wisdom_lookup(p, "gfx950", 256, 128, 1) is None
    # replacing keeps one entry per key
    mxr.wisdom_store(p, "gfx950", 256, 512, 1, 0, _t(zdirect_rpb=8), 40.0)
    doc = json.load(open(p))
    assert doc["format"] == "mxr-wisdom-1"
    keys = [(e["arch"], e["cus"], e["N"], e["filter"], e["spin"]) for e in doc["entries"]]
    assert len(keys) == len(set(keys)) == 3
    assert mxr.wisdom_lookup(p, "gfx950", 256, 512, 1).zdirect_rpb == 8


def test_file_written_by_hand_is_read(tmp_path):
    p = tmp_path / "w.json"
    p.write_text(json.dumps({"format": "mxr-wisdom-1", "entries": [
        {"arch": "gfx950", "cus": 256, "N": 160, "filter": 1, "spin": 0, "zdirect_kmax": 0, "zdirect_rpb": 4,
         "live_tiles": 1, "col_tc_x": 8, "col_tc_y": 4, "ms": 3.5, "note": "extra keys are ignored"},
        {"arch": "gfx950", "cus": 256, "N": 96, "filter": 1, "spin": 0, "zdirect_kmax": 999, "zdirect_rpb": 4,
         "live_tiles": 1, "col_tc_x": 8, "col_tc_y": 4}]}, indent=1))
    t = mxr.wisdom_lookup(str(p), "gfx950", 256, 160, 1)
    assert (t.col_tc_x, t.col_tc_y) == (8, 4)
    assert mxr.wisdom_lookup(str(p), "gfx950", 256, 96, 1) is None  # out-of-range entry dropped


@pytest.mark.parametrize("text", ["", "{", "[]", '{"format": "other", "entries": []}', '{"entries": {}}',
                                  '{"format": "mxr-wisdom-1", "entries": [1, 2'])
def test_malformed_file_is_an_error(tmp_path, text):
    p = tmp_path / "bad.json"
    p.write_text(text)
    with pytest.raises(mxr.MxrError) as e:
        mxr.wisdom_lookup(str(p), "gfx950", 256, 512, 1)
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
    with pytest.raises(mxr.MxrError):
        mxr.wisdom_store(str(p), "gfx950", 256, 512, 1, 0, _t())


def test_invalid_tuning_is_not_stored(tmp_path):
    p = str(tmp_path / "w.json")
    for bad in (dict(col_tc_x=5), dict(zdirect_kmax=65), dict(zdirect_rpb=0), dict(live_tiles=2)):
        with pytest.raises(mxr.MxrError) as e:
            mxr.wisdom_store(p, "gfx950", 256, 512, 1, 0, _t(**bad))
        assert e.value.code == mxr.MXR_ERR_INVALID_ARG
    assert mxr.wisdom_lookup(p, "gfx950", 256, 512, 1) is None


def test_shipped_mi355x_wisdom():
    """The package's MI355X entries (bench.py's default wisdom file) parse through the C ABI and
    hold valid launch shapes for the benchmark's 512^3 sharp-k call, 1024^3 and the batch's sizes."""
    import os

    path = os.path.join(os.path.dirname(mxr.__file__), "wisdom", "mi355x.json")
    for n in (512, 1024, 96, 128, 160):
        t = mxr.wisdom_lookup(path, "gfx950", 256, n, 1)
        assert t is not None, n
        d = t.as_dict()
        assert 0 <= d["zdirect_kmax"] <= 64 and 1 <= d["zdirect_rpb"] <= 4096, d
        assert d["col_tc_x"] in (4, 8, 16, 32) and d["col_tc_y"] in (4, 8, 16, 32), d
    assert mxr.wisdom_lookup(path, "gfx950", 256, 512, 1).zdirect_kmax == 32
