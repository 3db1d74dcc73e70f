"""Probabilistic truncation (simple_truncation False): keep the smallest
u / |m|^2 with u a Philox draw keyed by (object hash, seed).  The device
selection must match the oracle's bit for bit, in the plain step (including
parent pre-truncation), in hash-range passes and in batches."""
import math

import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd.rules import qcgd
from quids_amd.state import HostState

pytestmark = pytest.mark.gpu


@pytest.fixture
def probabilistic():
    q.simple_truncation = False
    q.truncation_seed = 1234
    yield
    q.simple_truncation = True
    q.truncation_seed = 0


def _chain(hs, cap, n_iter, seed):
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    ref = hs
    q.truncation_seed = seed
    for _ in range(n_iter):
        q.simulate(a, qcgd.step)
        ref = orc.apply_modifier(ref, qcgd.step)
        for r in (qcgd.split_merge(math.pi / 4), qcgd.erase_create(math.pi / 3, math.pi / 5)):
            st = q.simulate(a, r, b, None, cap)
            ref, rst = orc.simulate(ref, r, cap, simple_truncation=False, seed=seed)
            a, b = b, a
            got = a.download()
            assert st["num_unique"] == rst["num_unique"]
            assert np.array_equal(got.objects, ref.objects)
            assert np.array_equal(got.magnitude, ref.magnitude)
            assert got.total_proba == ref.total_proba
    return a.download()


def test_probabilistic_matches_oracle(probabilistic):
    hs = HostState.from_objects([qcgd.seeded_graph(7, 3)])
    x = _chain(hs, 300, 3, 1234)
    y = _chain(hs, 300, 3, 99)
    assert x.num_object == y.num_object == 300
    assert not np.array_equal(x.objects, y.objects)  # another seed, another sample


def test_probabilistic_parent_pretruncation(probabilistic):
    """A state larger than the cap is pre-truncated with the same keys."""
    objs = [qcgd.seeded_graph(5, s) for s in range(60)]
    mags = np.linspace(0.5, 1.5, len(objs)) * np.exp(1j * np.arange(len(objs)))
    hs = HostState.from_objects(objs, mags / np.linalg.norm(mags))
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    r = qcgd.erase_create(0.7, 0.2)
    q.simulate(a, r, b, None, 20)
    ref, _ = orc.simulate(hs, r, 20, simple_truncation=False, seed=1234)
    got = b.download()
    assert np.array_equal(got.objects, ref.objects)
    assert np.array_equal(got.magnitude, ref.magnitude)


def test_probabilistic_range_passes_and_batch(probabilistic):
    ctx = q.default_context()
    hs = HostState.from_objects([qcgd.seeded_graph(7, 8)])
    ctx.set_tuning(chunk_log2=10)
    try:
        _chain(hs, 500, 2, 1234)
        from quids_amd.batch import apply_modifier, batch_t, simulate_batched
        states = [HostState.from_objects([qcgd.seeded_graph(n, n)]) for n in (3, 4, 5)]
        a, b = batch_t.from_states(states), batch_t()
        refs = list(states)
        for _ in range(2):
            apply_modifier(a, qcgd.step)
            refs = [orc.apply_modifier(x, qcgd.step) for x in refs]
            for r in (qcgd.split_merge(0.6), qcgd.erase_create(0.9, 0.3)):
                simulate_batched(a, [r] * 3, b, 40)
                a, b = b, a
                refs = [orc.simulate(x, r, 40, simple_truncation=False, seed=1234)[0] for x in refs]
                for got, ref in zip(a.instances(), refs):
                    assert np.array_equal(got.objects, ref.objects)
                    assert np.array_equal(got.magnitude, ref.magnitude)
    finally:
        ctx.set_tuning(chunk_log2=0)
