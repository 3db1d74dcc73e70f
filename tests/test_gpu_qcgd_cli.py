"""QCGD command-line runs on the GPU, through both drop-in mirrors: the C++
API (tests/cpp/qcgd_cli, the reference's flags::parse_simulation /
utils::serialize / utils::print over quids_amd/rules/qcgd.hpp) and the
Python one (quids_amd.rules.qcgd.run_simulation), against the reference's
output on the same runs (tests/golden/qcgd_cli_*.txt)."""
import io
import os
import subprocess

import pytest

import quids_amd as q
from cli_util import assert_same_run, golden, runs
from quids_amd.rules import qcgd

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "qcgd_cli")


@pytest.fixture(autouse=True)
def _restore_globals():
    saved = (q.tolerance, q.align_byte_length, q.simple_truncation, q.safety_margin,
             q.load_balancing_bucket_per_thread, qcgd.utils.max_print_num_graphs)
    yield
    (q.tolerance, q.align_byte_length, q.simple_truncation, q.safety_margin,
     q.load_balancing_bucket_per_thread, qcgd.utils.max_print_num_graphs) = saved


@pytest.mark.parametrize("name", sorted(runs()))
def test_cpp_cli_matches_reference(name):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/qcgd_cli is not built (make -C tests/cpp)")
    print(f"qcgd_cli {name}", flush=True)
    out = subprocess.run([EXE, runs()[name]], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert_same_run(out.stdout, golden(name), name)


@pytest.mark.parametrize("name", sorted(runs()))
def test_python_cli_matches_reference(ctx, name):
    print(f"run_simulation {name}", flush=True)
    text = qcgd.run_simulation(runs()[name], file=io.StringIO())
    assert_same_run(text, golden(name), name)
