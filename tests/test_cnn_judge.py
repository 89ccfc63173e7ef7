"""The learned basic pass's judge (vszipcu.cnn.learned_allowed) on the CPU: the learned pass
replaces the plain one only where the measured PSNR delta at the measured sigmas around the
request is within the bound; outside the measured range or without a table the plain pass runs.
The shipped table (measured on MI355X by `python -m vszipcu.cnn --judge`) is checked for shape."""
from vszipcu import cnn

TABLE = {"bound_db": 0.05, "rows": [
    {"kind": "luma", "sigma": 2.0, "delta_db": -0.3}, {"kind": "luma", "sigma": 8.0, "delta_db": -0.04},
    {"kind": "luma", "sigma": 12.0, "delta_db": 0.01}, {"kind": "luma", "sigma": 25.0, "delta_db": 0.02},
    {"kind": "chroma", "sigma": 2.0, "delta_db": -0.06}, {"kind": "chroma", "sigma": 5.0, "delta_db": 0.2}]}


def test_rule():
    assert not cnn.learned_allowed(5.0, "luma", TABLE)          # bracket 2 (-0.3) .. 8
    assert cnn.learned_allowed(8.0, "luma", TABLE)              # measured point within the bound
    assert cnn.learned_allowed(10.0, "luma", TABLE)             # both brackets within
    assert not cnn.learned_allowed(30.0, "luma", TABLE)         # outside the measured range
    assert not cnn.learned_allowed(1.0, "luma", TABLE)
    assert not cnn.learned_allowed(3.0, "chroma", TABLE)        # lower bracket -0.06
    assert cnn.learned_allowed(5.0, "chroma", TABLE)
    assert not cnn.learned_allowed(5.0, "luma", None)           # no table: plain pass


def test_shipped_table():
    t = cnn.load_judge()
    assert t is not None and t["bound_db"] == cnn.JUDGE_BOUND_DB
    kinds = {r["kind"] for r in t["rows"]}
    assert kinds == {"luma", "chroma"}
    for k in kinds:
        sig = sorted(r["sigma"] for r in t["rows"] if r["kind"] == k)
        assert sig[0] <= 2.0 and sig[-1] >= 25.0 and len(sig) >= 6
