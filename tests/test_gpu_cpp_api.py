"""The drop-in C++ API (quids_amd/quids.hpp) running the reference's two
example programs on the GPU, checked against the reference's golden dumps."""
import math
import os
import subprocess

import numpy as np
import pytest

from golden_util import load

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "api_example")


def run():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/api_example is not built (make -C tests/cpp)")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout.splitlines()


def checkpoints(lines):
    res = {}
    for ln in lines:
        p = ln.split()
        if p[0] in ("qc_final",):
            res.setdefault(p[0], []).append((int(p[1]), float(p[2]), float(p[3])))
        else:
            res[p[0]] = (int(p[1]), float(p[2]), float(p[3]), int(p[4], 16))
    return res


def fixture_checkpoint(f, d):
    hx = 0
    for h in f[f"d{d}_hash"]:
        hx ^= int(h)
    return int(f[f"d{d}_N"]), float(f[f"d{d}_total_proba"]), float(np.sum(np.abs(f[f"d{d}_mag"]) ** 2)), hx


def test_cpp_examples(ctx):
    cp = checkpoints(run())
    qc = load("qc_example")
    for label, d in (("qc_h1", 0), ("qc_h2", 1), ("qc_gates", 5), ("qc_reversed", 11)):
        n, tp, s, hx = cp[label]
        rn, rtp, rs, rhx = fixture_checkpoint(qc, d)
        assert (n, hx) == (rn, rhx), label
        assert abs(s - rs) < 1e-12, label
    # reversed circuit: the initial registers with magnitudes times -1
    fin = sorted(cp["qc_final"])
    init = sorted(zip([4, 5], [-(1 / math.sqrt(2)), 0.0], [0.0, -(1 / math.sqrt(2))]))
    for (sz, re, im), (sz0, re0, im0) in zip(fin, init):
        assert sz == sz0 and abs(re - re0) < 1e-12 and abs(im - im0) < 1e-12
    g = load("qcgd_example")
    for label, d in (("qcgd_coin_ec", 3), ("qcgd_sm", 6), ("qcgd_reversed", 11)):
        n, tp, s, hx = cp[label]
        rn, rtp, rs, rhx = fixture_checkpoint(g, d)
        assert (n, hx) == (rn, rhx), label
        assert abs(tp - rtp) < 1e-12 and abs(s - rs) < 1e-12, label
