"""Steps run in hash-range passes (a candidate budget far below the step)
against the same steps in one pass, and against the oracle: identical
results."""
import math

import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd.rules import qcgd
from quids_amd.state import HostState

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert a.num_object == b.num_object, f"{what}: {a.num_object} vs {b.num_object}"
    assert np.array_equal(a.object_begin, b.object_begin), what
    assert np.array_equal(a.objects, b.objects), what
    assert np.array_equal(a.magnitude, b.magnitude), what
    assert a.total_proba == b.total_proba, what


def _chain(hs, n_iter, cap, chunk_log2, two_walk=0):
    ctx = q.default_context()
    ctx.set_tuning(chunk_log2=chunk_log2, split_generate=two_walk)
    try:
        a, b = q.it_t(), q.it_t()
        a.upload(hs)
        out = []
        for _ in range(n_iter):
            q.simulate(a, qcgd.step)
            for r in (qcgd.split_merge(math.pi / 4, 0.0), qcgd.erase_create(math.pi / 3, math.pi / 5)):
                st = q.simulate(a, r, b, None, cap)
                a, b = b, a
                out.append((a.download(), st))
        return out
    finally:
        ctx.set_tuning(chunk_log2=0, split_generate=0)


# The pass budget (2^chunk candidates) against a step's C candidates sets how
# the passes cover the 2^T - 1 hash slices the children are tagged with
# (T <= 8): several slices per pass, one, or (C > 255 * 2^chunk) a slice
# split over several passes that each keep the children of their bins.
@pytest.mark.parametrize("nodes,seed,n_iter,cap,chunk,min_c", [
    (6, 3, 3, 3000, 10, 0),
    (7, 5, 3, q.UINT64_MAX, 10, 0),
    (7, 5, 3, q.UINT64_MAX, 14, 0),
    (8, 42, 3, 700, 10, 0),
    (9, 2, 3, q.UINT64_MAX, 12, 255 * 4096 + 1),
    (40, 7, 2, 24, 16, 0),  # graphs of more than 32 nodes: the walks' 64-bit masks
])
@pytest.mark.parametrize("two_walk", [0, 1])
def test_range_passes_match_single_pass(nodes, seed, n_iter, cap, chunk, min_c, two_walk):
    """two_walk=1: each pass counts its children per bucket in one walk and
    writes them in a second (the path a region overflow takes)."""
    print(f"range nodes={nodes} seed={seed} iters={n_iter} cap={cap} chunk=2^{chunk} two_walk={two_walk}", flush=True)
    hs = HostState.from_objects([qcgd.seeded_graph(nodes, seed)])
    one = _chain(hs, n_iter, cap, 0)
    many = _chain(hs, n_iter, cap, chunk, two_walk)
    used = False
    for i, ((x, sx), (y, sy)) in enumerate(zip(one, many)):
        _same(x, y, f"step {i}")
        for k in ("num_candidates", "num_unique", "num_survivors"):
            assert sx[k] == sy[k], (i, k)
        used |= bool(sy["reserved"] & 4)
    assert used, "no step ran in hash-range passes"
    assert all(not (sy["reserved"] & 8) for _, sy in many) or two_walk, "a region overflowed"
    print("candidates per step:", [sx["num_candidates"] for _, sx in one], flush=True)
    assert max(sx["num_candidates"] for _, sx in one) >= min_c


def test_coin_in_range_passes_matches_single_pass():
    """coin (flip where left != right) through hash-range passes."""
    hs = HostState.from_objects([qcgd.seeded_graph(9, 11)])
    rules = (qcgd.coin(0.8, 0.3), qcgd.split_merge(0.5, 0.2), qcgd.erase_create(0.9, 1.3))

    def chain(chunk):
        ctx = q.default_context()
        ctx.set_tuning(chunk_log2=chunk)
        try:
            a, b = q.it_t(), q.it_t()
            a.upload(hs)
            out = []
            for _ in range(3):
                q.simulate(a, qcgd.step)
                for r in rules:
                    st = q.simulate(a, r, b, None, 2000)
                    a, b = b, a
                    out.append((a.download(), st))
            return out
        finally:
            ctx.set_tuning(chunk_log2=0)

    one, many = chain(0), chain(10)
    assert any(sy["reserved"] & 4 for _, sy in many), "no step ran in hash-range passes"
    for i, ((x, sx), (y, sy)) in enumerate(zip(one, many)):
        _same(x, y, f"step {i}")


def test_range_passes_match_oracle():
    hs = HostState.from_objects([qcgd.seeded_graph(6, 8)])
    ctx = q.default_context()
    ctx.set_tuning(chunk_log2=10)
    try:
        a, b = q.it_t(), q.it_t()
        a.upload(hs)
        ref = hs
        for _ in range(2):
            q.simulate(a, qcgd.step)
            ref = orc.apply_modifier(ref, qcgd.step)
            for r in (qcgd.split_merge(0.6, 0.2), qcgd.erase_create(1.0, 0.3)):
                st = q.simulate(a, r, b, None, 400)
                ref, rst = orc.simulate(ref, r, 400)
                a, b = b, a
                got = a.download()
                assert st["num_unique"] == rst["num_unique"]
                assert np.array_equal(got.objects, ref.objects)
                assert np.array_equal(got.magnitude, ref.magnitude)
                assert got.total_proba == ref.total_proba
    finally:
        ctx.set_tuning(chunk_log2=0)


def test_batch_with_instances_beyond_the_budget():
    """Instances whose step exceeds the budget run alone, in hash-range
    passes; the batch still matches each instance's plain step."""
    from quids_amd.batch import apply_modifier, batch_t, simulate_batched
    specs = [(3, 1), (7, 2), (2, 3), (8, 4), (5, 5)]
    states = [HostState.from_objects([qcgd.seeded_graph(n, s)]) for n, s in specs]
    ctx = q.default_context()
    ctx.set_tuning(chunk_log2=10)
    try:
        a, b = batch_t.from_states(states), batch_t()
        plain = []
        for hs in states:
            it = q.it_t()
            it.upload(hs)
            plain.append(it)
        for _ in range(3):
            apply_modifier(a, qcgd.step)
            for it in plain:
                q.simulate(it, qcgd.step)
            for r in (qcgd.split_merge(0.7, 0.1), qcgd.erase_create(0.4, 1.1)):
                stats = simulate_batched(a, [r] * len(specs), b, 500)
                a, b = b, a
                got = a.instances()
                for i in range(len(specs)):
                    nxt = q.it_t()
                    st = q.simulate(plain[i], r, nxt, None, 500)
                    plain[i] = nxt
                    ref = nxt.download()
                    assert np.array_equal(got[i].objects, ref.objects), i
                    assert np.array_equal(got[i].magnitude, ref.magnitude), i
                    assert stats[i]["num_unique"] == st["num_unique"], i
    finally:
        ctx.set_tuning(chunk_log2=0)
