"""The autotuner: picks the fastest of its grid and the starting tuning, sets
it on the context, and every tuning gives bit-identical steps."""
import math

import numpy as np
import pytest

import quids_amd as q
from quids_amd import tuning
from quids_amd.rules import qcgd
from quids_amd.state import HostState

pytestmark = pytest.mark.gpu


def _state():
    it, nxt = q.it_t(), q.it_t()
    it.upload(HostState.from_objects([qcgd.seeded_graph(8, 5)]))
    q.simulate(it, qcgd.split_merge(math.pi / 4), nxt, None, q.UINT64_MAX)
    return nxt


def test_autotune_sets_a_grid_point_and_keeps_results():
    ctx = q.default_context()
    before = ctx.get_tuning()
    st = _state()
    rules = [qcgd.erase_create(math.pi / 3, math.pi / 5)]
    grid = {"dedup_bucket_target": (256, 512), "first_pass_bits": (6, 12)}
    try:
        ref = q.it_t()
        q.simulate(st, rules[0], ref, None, 5000)
        res = tuning.autotune(ctx, st, rules, 5000, grid=grid, reps=1, key="test")
        assert res["tried"] == 4
        t = ctx.get_tuning()
        for k, v in res["tuning"].items():
            # a grid point, or the starting value when none was faster
            assert t[k] == v and (v in grid[k] or v == before[k])
        got = q.it_t()
        q.simulate(st, rules[0], got, None, 5000)
        a, b = ref.download(), got.download()
        assert np.array_equal(a.objects, b.objects) and np.array_equal(a.magnitude, b.magnitude)
        # memoised per (device, key)
        again = tuning.autotune(ctx, st, rules, 5000, grid=grid, reps=1, key="test")
        assert again is res
    finally:
        ctx.set_tuning(**{k: before[k] for k in ("dedup_bucket_target", "first_pass_bits", "generate_grid")})


@pytest.mark.parametrize("first_pass_bits", [8, 6])
def test_fused_and_split_generate_agree(first_pass_bits):
    """Generate fused with the first radix pass against generate followed by
    the first pass as its own kernel: identical steps (plain and batched)."""
    from quids_amd.batch import batch_t, simulate_batched
    ctx = q.default_context()
    before = ctx.get_tuning()
    hs = HostState.from_objects([qcgd.seeded_graph(9, 4)])
    rules = [qcgd.split_merge(0.7, 0.2), qcgd.erase_create(1.1, 0.5)]
    out = {}
    try:
        for split in (0, 1):
            ctx.set_tuning(split_generate=split, first_pass_bits=first_pass_bits, dedup_bucket_target=256)
            a, b = q.it_t(), q.it_t()
            a.upload(hs)
            res = []
            for _ in range(2):
                q.simulate(a, qcgd.step)
                for r in rules:
                    q.simulate(a, r, b, None, 20000)
                    a, b = b, a
                    res.append(a.download())
            bt, bb = batch_t.from_states([hs, HostState.from_objects([qcgd.seeded_graph(6, 1)])]), batch_t()
            for r in rules:
                simulate_batched(bt, [r, r], bb, 5000)
                bt, bb = bb, bt
            res.append(bt.state.download())
            out[split] = res
        for x, y in zip(out[0], out[1]):
            assert np.array_equal(x.objects, y.objects) and np.array_equal(x.magnitude, y.magnitude)
    finally:
        ctx.set_tuning(split_generate=0, first_pass_bits=before["first_pass_bits"],
                       dedup_bucket_target=before["dedup_bucket_target"])


@pytest.mark.parametrize("copies", [3, 6, 12, 1200])
@pytest.mark.parametrize("cap,target", [(256, 128), (512, 256), (1024, 512), (2048, 1024)])
def test_dedup_table_sizes_match_oracle(cap, target, copies):
    """Every dedup table size (256 .. 2048 entries per workgroup) against
    the oracle, on a state of `copies` copies of one graph (each child hash
    then has that many members: groups summed in registers up to 8 members,
    by the sort-based kernel past that, by the oversized-bucket kernel past
    the table size) and on a QCGD chain."""
    from oracle import oracle as orc
    ctx = q.default_context()
    before = ctx.get_tuning()
    rng = np.random.default_rng(cap + copies)
    g = qcgd.seeded_graph(5, 3)
    mags = rng.normal(size=copies) + 1j * rng.normal(size=copies)
    dup = HostState.from_objects([g] * copies, mags=mags / np.linalg.norm(mags))
    tol0 = q.tolerance
    try:
        q.tolerance = 1e-30
        ctx.set_tuning(dedup_cap=cap, dedup_bucket_target=target)
        for hs, rule in ((dup, qcgd.erase_create(0.8, 0.3)), (dup, qcgd.split_merge(0.5, 0.1)),
                         (HostState.from_objects([qcgd.seeded_graph(8, 7)]), qcgd.split_merge(0.9, 0.4))):
            ref = hs
            a = q.it_t()
            a.upload(hs)
            for _ in range(2):
                b = q.it_t()
                st = q.simulate(a, rule, b, None, q.UINT64_MAX)
                ref, rst = orc.simulate(ref, rule)
                a = b
                got = a.download()
                assert st["num_unique"] == rst["num_unique"]
                assert np.array_equal(got.objects, ref.objects)
                assert np.array_equal(got.magnitude, ref.magnitude)
                assert got.total_proba == ref.total_proba
    finally:
        q.tolerance = tol0
        ctx.set_tuning(dedup_cap=before["dedup_cap"], dedup_bucket_target=before["dedup_bucket_target"])


def test_fused_generate_overflow_falls_back():
    """6000 copies of one graph: each child hash has 6000 members, more than
    a first-pass bucket of the fused generate holds, so the step reports the
    unfused fallback (stats reserved bit 8) and still matches the oracle."""
    from oracle import oracle as orc
    rng = np.random.default_rng(6000)
    mags = rng.normal(size=6000) + 1j * rng.normal(size=6000)
    hs = HostState.from_objects([qcgd.seeded_graph(5, 3)] * 6000, mags=mags / np.linalg.norm(mags))
    tol0 = q.tolerance
    try:
        q.tolerance = 1e-30
        a, b = q.it_t(), q.it_t()
        a.upload(hs)
        rule = qcgd.erase_create(0.8, 0.3)
        st = q.simulate(a, rule, b, None, q.UINT64_MAX)
        ref, rst = orc.simulate(hs, rule)
        assert st["reserved"] & 8, st
        got = b.download()
        assert st["num_unique"] == rst["num_unique"]
        assert np.array_equal(got.objects, ref.objects)
        assert np.array_equal(got.magnitude, ref.magnitude)
        assert got.total_proba == ref.total_proba
    finally:
        q.tolerance = tol0


def test_select_graph_replays_match_eager():
    """The select's digit rounds replayed from the context's hipGraphs
    against the same rounds launched one by one (eager_select=1): a chain
    of truncated QCGD steps, each replaying the graph between the step's own
    kernels and host reads, gives identical states, and so do the parent
    pre-truncation (a graph of more rounds) and truncated batched steps
    (the per-instance rounds' graph), also after the context's scratch was
    trimmed (a graph is captured again when its state moved)."""
    from quids_amd.batch import apply_modifier, batch_t, simulate_batched
    ctx = q.default_context()
    before = ctx.get_tuning()
    rules = [qcgd.split_merge(0.7, 0.2), qcgd.erase_create(1.1, 0.5)]
    out = {}
    try:
        for eager in (1, 0, 1, 0):
            ctx.set_tuning(eager_select=eager)
            a, b = q.it_t(), q.it_t()
            a.upload(HostState.from_objects([qcgd.seeded_graph(8, 4)]))
            res = []
            for _ in range(3):
                q.simulate(a, qcgd.step)
                for cap in (300, 41):
                    for r in rules:
                        q.simulate(a, r, b, None, cap)
                        a, b = b, a
                        res.append(a.download())
            ctx.trim()  # the scratch, the loops' device state included, is allocated again
            big = res[-3]  # pre-truncation: a parent state larger than the cap
            c, d = q.it_t(), q.it_t()
            c.upload(big)
            q.simulate(c, rules[0], d, None, max(1, big.num_object // 3))
            res.append(d.download())
            # batched: per-instance rounds from the device
            bt = batch_t.from_states([HostState.from_objects([qcgd.seeded_graph(n, sd)])
                                      for n, sd in ((9, 4), (6, 1), (12, 3), (7, 7))])
            bb = batch_t()
            for _ in range(2):
                apply_modifier(bt, qcgd.step)
                for r in rules:
                    simulate_batched(bt, [r] * 4, bb, 250)
                    bt, bb = bb, bt
                    res.append(bt.state.download())
            out.setdefault(eager, []).append(res)
        ref = out[1][0]
        for runs in out.values():
            for res in runs:
                for x, y in zip(res, ref):
                    assert np.array_equal(x.objects, y.objects) and np.array_equal(x.magnitude, y.magnitude)
                    assert x.total_proba == y.total_proba
    finally:
        ctx.set_tuning(eager_select=before["eager_select"])
