"""QCGD command-line flags, statistics and printing of the Python mirror
(quids_amd.rules.qcgd.flags / utils / run_simulation) against the
reference's output on the same runs (tests/golden/qcgd_cli_*.txt, written
by the reference's own flags::parse_simulation, utils::serialize and
utils::print).  The steps run on the oracle here (CPU); the GPU runs are in
test_gpu_qcgd_cli.py."""
import io
import math

import numpy as np
import pytest

import quids_amd as q
from cli_util import assert_same_run, golden, runs
from oracle import oracle as orc
from quids_amd.rules import qcgd
from quids_amd.state import HostState

U64MAX = (1 << 64) - 1


@pytest.fixture(autouse=True)
def _restore_globals():
    saved = (q.tolerance, q.align_byte_length, q.simple_truncation, q.safety_margin,
             q.load_balancing_bucket_per_thread, qcgd.utils.max_print_num_graphs)
    yield
    (q.tolerance, q.align_byte_length, q.simple_truncation, q.safety_margin,
     q.load_balancing_bucket_per_thread, qcgd.utils.max_print_num_graphs) = saved


def _oracle_run(arg):
    def step_rule(st, rule, max_num_object):
        out, stats = orc.simulate(st, rule, max_num_object or U64MAX, tolerance=q.tolerance,
                                  align=q.align_byte_length)
        return out, (stats["num_candidates"], stats["num_unique"])

    return qcgd.run_simulation(arg, step_rule, lambda st, mod: orc.apply_modifier(st, mod),
                               state=HostState.from_objects([]), file=io.StringIO())


@pytest.mark.parametrize("name", sorted(runs()))
def test_run_matches_reference(name):
    assert_same_run(_oracle_run(runs()[name]), golden(name), name)


def test_read_n_iter_flags():
    n, rev, mx = qcgd.flags.read_n_iter("12,seed=3,reversed_n_iter=4,max_print_num_graphs=7,tolerance=1e-12,"
                                        "safety_margin=0.3,align=4,simple_truncate=0,"
                                        "load_balancing_bucket_per_thread=9,max_num_object=1000")
    assert (n, rev, mx) == (12, 4, 1000)
    assert q.tolerance == float(np.float32(1e-12)) and q.align_byte_length == 4
    assert q.simple_truncation is False and q.load_balancing_bucket_per_thread == 9
    assert q.safety_margin == float(np.float32(0.3)) and qcgd.utils.max_print_num_graphs == 7
    # defaults; the tolerance passes through a C float even when not given
    q.tolerance = 1e-30
    assert qcgd.flags.read_n_iter("5") == (5, 0, 0)
    assert q.tolerance == float(np.float32(1e-30)) != 1e-30


def test_read_rule_angles_and_reverses():
    sim = qcgd.flags.read_rule("split_merge,theta=0.3,phi=0.1,xi=0.2,n_iter=3;step;erase_create;coin,xi=1")
    assert [(e[0], e[1]) for e in sim] == [(3, True), (1, False), (1, True), (1, True)]
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    th, ph, xi = f32(math.pi * f32(0.3)), f32(math.pi * f32(0.1)), f32(math.pi * f32(0.2))
    d, r = sim[0][3].desc(), sim[0][5].desc()
    ref, rref = qcgd.split_merge(th, ph, xi).desc(), qcgd.split_merge(th, ph, -xi).desc()
    for a, b in ((d, ref), (r, rref)):
        assert (a.do_re, a.do_im, a.do_not_re, a.do_not_im) == (b.do_re, b.do_im, b.do_not_re, b.do_not_im)
    assert sim[1][2] is qcgd.step and sim[1][4] is qcgd.reversed_step
    e = sim[2][3].desc()
    ref = qcgd.erase_create(f32(math.pi * f32(0.25))).desc()  # default theta 0.25 pi
    assert (e.do_re, e.do_not_re) == (ref.do_re, ref.do_not_re)


def test_read_state_magnitudes_and_particles():
    hs = HostState.from_objects([])
    qcgd.flags.read_n_iter("1,seed=42")
    qcgd.flags.read_state("12;3,n_graphs=4,real=0.5,imag=-0.25", hs)
    assert hs.num_object == 5
    # the first graph is the seeded 12-node graph (the same rand() sequence)
    assert hs.get_object(0)[0] == qcgd.seeded_graph(12, 42)
    m = np.asarray(hs.magnitude)
    root = np.sqrt(np.float32(4))
    assert m[0] == 1 and m[1] == complex(float(np.float32(0.5) / root), float(np.float32(-0.25) / root))
    assert all(qcgd.num_nodes(hs.get_object(i)[0]) == 3 for i in range(1, 5))


def test_value_parsing_quirks():
    # a key after the first separator reads to the end; values stop at a
    # separator like atoi / atof
    assert qcgd._value_of("seed=12,x=3", "x=", ",") == "3"
    assert qcgd._atoi(qcgd._value_of("seed=12,x=3", "seed=", ",")) == 12
    assert qcgd._atof("2.5e-3,rest") == 2.5e-3 and qcgd._atoi("  -7abc") == -7 and qcgd._atoi("abc") == 0
