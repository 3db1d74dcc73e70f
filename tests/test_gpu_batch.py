"""The batched step against separate plain steps of each instance: objects,
magnitudes, total probabilities and counters must agree bit for bit."""
import math

import numpy as np
import pytest

import quids_amd as q
from quids_amd.batch import apply_modifier, batch_t, simulate_batched
from quids_amd.rules import qcgd
from quids_amd.state import HostState

from helpers import sparse_graph

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert a.num_object == b.num_object, f"{what}: {a.num_object} vs {b.num_object} objects"
    assert np.array_equal(a.object_size, b.object_size), what
    assert np.array_equal(a.object_begin, b.object_begin), what
    assert np.array_equal(a.objects, b.objects), what
    assert np.array_equal(a.magnitude, b.magnitude), what


def _run(specs, n_iter, caps):
    """specs: (nodes, seed, theta1, phi1, theta2, phi2) per instance; an
    instance of 0 nodes starts with no object at all."""
    print(f"batch instances={len(specs)} iters={n_iter} caps={caps}", flush=True)
    states = [HostState.from_objects([qcgd.seeded_graph(n, s)] if n else []) for n, s, *_ in specs]
    a, b = batch_t.from_states(states), batch_t()
    plain = []
    for hs in states:
        it = q.it_t()
        it.upload(hs)
        plain.append(it)
    caps = list(caps) if not np.isscalar(caps) else [caps] * len(specs)
    for step in range(n_iter):
        apply_modifier(a, qcgd.step)
        for it in plain:
            q.simulate(it, qcgd.step)
        for which in (0, 1):
            rules = [qcgd.split_merge(t1, p1) if which == 0 else qcgd.erase_create(t2, p2)
                     for _, _, t1, p1, t2, p2 in specs]
            stats = simulate_batched(a, rules, b, caps)
            a, b = b, a
            got = a.instances()
            for i, it in enumerate(plain):
                nxt = q.it_t()
                st = q.simulate(it, rules[i], nxt, None, caps[i])
                plain[i] = nxt
                ref = nxt.download()
                what = f"iteration {step} rule {which} instance {i}"
                _same(got[i], ref, what)
                for k in ("num_candidates", "num_unique", "num_survivors", "output_bytes"):
                    assert stats[i][k] == st[k], (what, k, stats[i][k], st[k])
                assert stats[i]["total_proba"] == st["total_proba"], what


def test_batch_matches_plain_untruncated():
    specs = [(3, 1, 0.7, 0.3, 1.1, 0.4), (4, 2, math.pi / 4, 0.0, math.pi / 3, math.pi / 5),
             (5, 3, 0.2, 2.0, 0.9, 0.1), (2, 4, 1.2, 0.5, 0.3, 2.5)]
    _run(specs, 3, q.UINT64_MAX)


def test_batch_matches_plain_truncated_uneven():
    rng = np.random.default_rng(3)
    specs = []
    for i in range(24):
        specs.append((int(rng.integers(3, 9)), int(rng.integers(0, 1000)), float(rng.uniform(0, math.pi / 2)),
                      float(rng.uniform(0, math.pi)), float(rng.uniform(0, math.pi / 2)),
                      float(rng.uniform(0, math.pi))))
    caps = [int(c) for c in rng.choice([7, 50, 300, 2000, 10 ** 6], size=len(specs))]
    _run(specs, 4, caps)


def test_empty_instances():
    """Instances with no object first, between truncated ones and last, then
    a batch of empty instances only: every instance as its own plain step
    (an empty state steps to an empty state of total probability 0)."""
    specs = [(0, 0, 0.7, 0.3, 1.1, 0.4), (5, 1, 0.7, 0.3, 1.1, 0.4), (0, 0, 0.2, 1.3, 0.5, 0.9),
             (6, 2, math.pi / 4, 0.0, math.pi / 3, math.pi / 5), (4, 5, 1.2, 0.5, 0.3, 2.5),
             (0, 0, 0.2, 1.3, 0.5, 0.9)]
    _run(specs, 3, [10, 20, 10, 30, 10 ** 6, 10])
    _run(specs[:1] + specs[2:3], 1, 10)


def test_equal_objects_of_different_instances_do_not_merge():
    # the same graph twice with different angles, plus an identical pair
    specs = [(5, 9, 0.7, 0.3, 1.1, 0.4), (5, 9, 0.2, 1.3, 0.5, 0.9), (5, 9, 0.7, 0.3, 1.1, 0.4)]
    _run(specs, 2, 400)


def test_many_small_instances():
    rng = np.random.default_rng(11)
    specs = [(int(rng.integers(2, 7)), int(rng.integers(0, 10 ** 6)), float(rng.uniform(0, 1.5)),
              float(rng.uniform(0, 3)), float(rng.uniform(0, 1.5)), float(rng.uniform(0, 3))) for _ in range(256)]
    _run(specs, 2, 256)


def test_most_instances():
    """4096 instances, the most one call takes (the select's per-instance
    counters in LDS and its instance scans over chunks of 1024 all full),
    most of them truncated."""
    rng = np.random.default_rng(13)
    specs = [(int(rng.integers(2, 6)), int(rng.integers(0, 10 ** 6)), float(rng.uniform(0, 1.5)),
              float(rng.uniform(0, 3)), float(rng.uniform(0, 1.5)), float(rng.uniform(0, 3))) for _ in range(4096)]
    caps = [int(c) for c in rng.choice([2, 3, 5, 10 ** 6], size=len(specs))]
    _run(specs, 1, caps)


def test_wide_instance_steps_alone():
    """Instances holding graphs of more than 64 nodes (beyond the batched
    kernels) between small ones: they step alone through the plain path and
    the small ones batched, the outputs concatenated; every instance as its
    own plain step.  (split_merge and coin only: erase_create would give the
    sparse graphs 2^n children.)"""
    print("batch with wide instances", flush=True)
    states = [HostState.from_objects([g]) for g in (
        qcgd.seeded_graph(5, 1), sparse_graph(70, (3, 11, 20, 33, 47, 58, 66), (7, 25, 40)),
        qcgd.seeded_graph(6, 2), qcgd.seeded_graph(4, 3), sparse_graph(66, (2, 9, 30, 44, 51), (14, 36)))]
    a, b = batch_t.from_states(states), batch_t()
    plain = []
    for hs in states:
        it = q.it_t()
        it.upload(hs)
        plain.append(it)
    caps = [40, 600, 30, 10 ** 6, 600]
    for k, rule in enumerate([qcgd.split_merge(0.6, 0.3), qcgd.coin(0.9, 0.4), qcgd.split_merge(1.1, 0.2)]):
        stats = simulate_batched(a, [rule] * len(states), b, caps)
        a, b = b, a
        got = a.instances()
        for i, it in enumerate(plain):
            nxt = q.it_t()
            st = q.simulate(it, rule, nxt, None, caps[i])
            plain[i] = nxt
            what = f"step {k} instance {i}"
            _same(got[i], nxt.download(), what)
            for key in ("num_candidates", "num_unique", "num_survivors", "output_bytes"):
                assert stats[i][key] == st[key], (what, key, stats[i][key], st[key])
            assert stats[i]["total_proba"] == st["total_proba"], what


def test_batch_split_into_chunks():
    """A candidate budget of 2^10 per launch sequence splits the batch into
    many chunks (concatenated on the device): results are unchanged."""
    ctx = q.default_context()
    rng = np.random.default_rng(7)
    specs = [(int(rng.integers(3, 8)), int(rng.integers(0, 1000)), float(rng.uniform(0, 1.5)),
              float(rng.uniform(0, 3)), float(rng.uniform(0, 1.5)), float(rng.uniform(0, 3))) for _ in range(20)]
    ctx.set_tuning(chunk_log2=10)
    try:
        _run(specs, 3, 500)
    finally:
        ctx.set_tuning(chunk_log2=0)
