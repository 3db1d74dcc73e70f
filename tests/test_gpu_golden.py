"""The HIP step against the reference's golden vectors, step by step."""
import pytest

import quids_amd as q

from golden_util import fixtures, replay_stepwise

pytestmark = pytest.mark.gpu


def _gpu_rule(hs, rule, max_num_object, tol, align):
    q.tolerance, q.align_byte_length = tol, align
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    st = q.simulate(a, rule, b, q.sy_it_t(), max_num_object)
    return b.download(), st


def _gpu_mod(hs, mod):
    a = q.it_t()
    a.upload(hs)
    q.simulate(a, mod)
    return a.download()


@pytest.mark.parametrize("name", fixtures())
def test_gpu_matches_reference(ctx, name):
    assert replay_stepwise(name, _gpu_rule, _gpu_mod) > 0
