"""The oracle's forward-mode derivative (the reference gradient of the
adjoint tests) against central finite differences of the oracle's own plain
steps, and against closed forms (CPU)."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from quids_amd.adjoint import OBS_NODES_EQ, OBS_NUM_NODES, OBS_PARTICLES, Program, RuleOp, qcgd_program
from quids_amd.rules import qcgd
from quids_amd.state import HostState

P0 = np.array([0.7, 0.3, 1.1, 0.4])


def _fd(hs, prog, p, obs, arg, h=1e-5):
    out = []
    for i in range(len(p)):
        e = np.zeros(len(p))
        e[i] = h
        out.append((orc.program_loss(hs, prog, p + e, obs, arg) - orc.program_loss(hs, prog, p - e, obs, arg)) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("nodes,seed,obs,arg,n_iter", [
    (3, 7, OBS_PARTICLES, 0, 3),
    (3, 7, OBS_NUM_NODES, 0, 3),
    (4, 1, OBS_NODES_EQ, 4, 3),
    (5, 3, OBS_PARTICLES, 0, 2),
])
def test_jvp_matches_central_differences(nodes, seed, obs, arg, n_iter):
    hs = HostState.from_objects([qcgd.seeded_graph(nodes, seed)])
    prog = qcgd_program(n_iter)
    v, g, _ = orc.program_jvp(hs, prog, P0, obs, arg)
    assert v == pytest.approx(orc.program_loss(hs, prog, P0, obs, arg), rel=1e-12)
    fd = _fd(hs, prog, P0, obs, arg)
    # central differences with h = 1e-5: O(h^2) truncation + O(eps/h) rounding
    assert np.abs(fd - g).max() <= 1e-8 * max(1.0, np.abs(g).max())


def test_jvp_single_erase_create_closed_form():
    """One erase_create on a 1-node graph with both particles set: children
    |erased> with do_conj = conj(polar(sin t, p)) and |kept> with
    -do_not_conj = -conj(polar(cos t, 0)); the expected particle count is
    2 cos^2 t, so d/dt = -2 sin 2t and d/dp = 0.  Catches a dropped or
    conjugated derivative term."""
    g1 = qcgd.make_graph(1)
    g1 = qcgd.with_particles(g1, [True], [True])
    hs = HostState.from_objects([g1])
    prog = Program([RuleOp(qcgd.erase_create, (0, 1))], 1)
    t, p = 0.37, 1.3
    v, g, _ = orc.program_jvp(hs, prog, [t, p], OBS_PARTICLES)
    assert v == pytest.approx(2 * math.cos(t) ** 2, rel=1e-14)
    assert g[0] == pytest.approx(-2 * math.sin(2 * t), rel=1e-13)
    assert abs(g[1]) < 1e-15


def test_jvp_cap_at_state_size_is_untruncated():
    """A cap that keeps every object leaves the derivative unchanged, and a
    cap below it changes the value (the mask is active)."""
    hs = HostState.from_objects([qcgd.seeded_graph(4, 2)])
    v0, g0, fin = orc.program_jvp(hs, qcgd_program(2), P0, OBS_PARTICLES)
    n = fin.num_object
    v1, g1, _ = orc.program_jvp(hs, qcgd_program(2, max_num_object=10 ** 9), P0, OBS_PARTICLES)
    assert v0 == v1 and np.array_equal(g0, g1)
    v2, g2, fin2 = orc.program_jvp(hs, qcgd_program(2, max_num_object=n // 2), P0, OBS_PARTICLES)
    assert fin2.num_object <= n // 2 and not np.array_equal(g0, g2)


def test_philox_known_answers():
    """Philox2x64-10, the generator of the probabilistic truncation, against
    the published known-answer vectors of Random123 (counter, key -> out)."""
    import ctypes
    L = orc._load()
    L.or_philox2x64_10.argtypes = [ctypes.c_uint64] * 3 + [ctypes.POINTER(ctypes.c_uint64)]
    for c0, c1, key, want in [
        (0, 0, 0, (0xca00a0459843d731, 0x66c24222c9a845b5)),
        (2 ** 64 - 1, 2 ** 64 - 1, 2 ** 64 - 1, (0x65b021d60cd8310f, 0x4d02f3222f86df20)),
    ]:
        out = (ctypes.c_uint64 * 2)()
        L.or_philox2x64_10(c0, c1, key, out)
        assert (out[0], out[1]) == want
