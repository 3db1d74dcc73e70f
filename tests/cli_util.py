"""Comparison of QCGD command-line run outputs (a JSON list of
utils::serialize objects, then utils::print lines) with the reference's
(tests/golden/qcgd_cli_*.txt)."""
import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def runs():
    with open(os.path.join(GOLDEN, "qcgd_cli.json")) as f:
        return json.load(f)


def golden(name):
    with open(os.path.join(GOLDEN, f"qcgd_cli_{name}.txt")) as f:
        return f.read()


def split(text):
    """(list of stats dicts, printed graph lines)."""
    end = text.index("\n]\n") + 3
    return json.loads(text[:end]), text[end:].splitlines()


def _norm(line):
    """|m|^2 from a printed line's 5-decimal magnitude."""
    re_, rest = line.strip().split(" ", 1)
    sign, im = rest.split(" ", 2)[:2]
    im = float(im.rstrip("i"))
    return float(re_) ** 2 + im * im


def assert_same_run(got, ref, what):
    """Statistics: counts exactly, the rest to the 6 digits they are printed
    with.  Graph lines: the same set of lines, each list ordered by
    non-increasing |m|^2.  Positions are not compared: graphs whose |m|^2
    agree to rounding may swap (the reference sorts with an unstable
    parallel sort, and magnitudes differ from it in the last bits).  When
    the print is cut, only the printed magnitudes are compared."""
    gs, gl = split(got)
    rs, rl = split(ref)
    assert len(gs) == len(rs), what
    for i, (a, b) in enumerate(zip(gs, rs)):
        assert a.keys() == b.keys(), (what, i)
        assert a["num_graphs"] == b["num_graphs"], (what, i)
        for k in a:
            assert abs(a[k] - b[k]) <= 1e-5 * max(1.0, abs(b[k])), (what, i, k, a[k], b[k])
    assert len(gl) == len(rl), what
    trailer = [ln for ln in rl if "other graphs" in ln]
    assert [ln for ln in gl if "other graphs" in ln] == trailer, what
    gg = [ln for ln in gl if "other graphs" not in ln]
    rg = [ln for ln in rl if "other graphs" not in ln]
    for lines in (gg, rg):  # |m| from 5-decimal parts: each within 7.1e-6
        mods = [_norm(ln) ** 0.5 for ln in lines]
        assert all(x >= y - 1.5e-5 for x, y in zip(mods, mods[1:])), what
    if not trailer:
        assert sorted(gg) == sorted(rg), what
    else:
        mag = lambda ln: ln.split("i  ")[0]  # noqa: E731
        assert sorted(mag(ln) for ln in gg) == sorted(mag(ln) for ln in rg), what
