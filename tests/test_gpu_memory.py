"""Memory-driven truncation (max_num_object == 0) on the device against the
reference's two binary searches (src/quids.hpp:459-486, 510-536) restated
here over the oracle's data, and the step itself against the oracle at the
counts the searches give.

The library reports the target memory of each side (its inputs -- free
memory, state and scratch sizes -- are the device's); the free memory it
sees is capped through qd_tuning.free_mem_mib so that the cuts bite.  The
memory model (src/quids.hpp:574-589, 850-861) with the library's footprints:
a parent weighs 28 + its padded bytes + 112 per child, times 9/8 for the
upsize policy; a unique weighs 28 + its child's padded size."""
import math

import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd.rules import qcgd
from quids_amd.state import HostState

from helpers import assert_states_identical, qcgd_initial, qcgd_iteration_rules

pytestmark = pytest.mark.gpu

PER_OBJECT = 16 + 8 + 4
PER_CANDIDATE = 112
U64MAX = (1 << 64) - 1


def _norm(m):
    return m.real * m.real + m.imag * m.imag  # as the device computes |m|^2 (no fused multiply-add)


def reference_search(weights_in_order, target, upsize):
    """src/quids.hpp:472-485: begin / end / middle over the objects in
    truncation order; returns the count truncate() leaves (n if the total
    fits)."""
    n = len(weights_in_order)
    S = np.concatenate([[0], np.cumsum(np.asarray(weights_in_order, dtype=np.uint64))]).astype(object)
    total = S[n] + (S[n] // 8 if upsize else 0)
    if total <= target:
        return n
    begin, end, kept = 0, n, n
    while end > begin + 1:
        middle = (begin + end) // 2
        kept = middle
        used = S[middle] - S[begin]
        if upsize:
            used += used // 8
        if used < target:
            target -= used
            begin = middle
        else:
            end = middle
    return kept


def expected_parent_cap(hs, rule, target):
    N = hs.num_object
    w = []
    for i in range(N):
        o, _ = hs.get_object(i)
        nc, _ = orc.num_child(rule, o)
        w.append(PER_OBJECT + int(hs.object_begin[i + 1] - hs.object_begin[i]) + nc * PER_CANDIDATE)
    order = sorted(range(N), key=lambda i: (-_norm(hs.magnitude[i]), i))
    return reference_search([w[i] for i in order], target, True)


def expected_survivor_cap(uniq, rule, target, align=8):
    """uniq: every unique of the step with its summed (not normalised)
    magnitude (the oracle's output without a survivor cap); ordered here by
    the truncation key, |m|^2 desc then hash asc."""
    keyed = []
    for i in range(uniq.num_object):
        o, m = uniq.get_object(i)
        sz = len(o)
        pad = (align - sz % align) % align if align > 1 else 0
        keyed.append((-_norm(m), orc.rule_hash(rule, o), PER_OBJECT + sz + pad))
    keyed.sort()
    return reference_search([k[2] for k in keyed], target, False)


@pytest.mark.parametrize("free_mib", [2, 8, 32])
def test_memory_driven_truncation_matches_reference_search(ctx, free_mib):
    print(f"memory-driven truncation free_mem_mib={free_mib}", flush=True)
    old = ctx.get_tuning()
    hs = qcgd_initial(8, 5)
    rules = qcgd_iteration_rules()
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    ref = hs
    cut_parents = cut_survivors = 0
    try:
        ctx.set_tuning(free_mem_mib=free_mib)
        for it in range(3):
            for r in rules:
                if r.is_modifier:
                    q.simulate(a, r)
                    ref = orc.apply_modifier(ref, r)
                    continue
                ctx.trim()  # the scratch held by the context counts as available memory
                st = q.simulate(a, r, b, None, 0)
                a, b = b, a
                what = f"free {free_mib} MiB iteration {it + 1} {type(r).__name__}"
                Tp, Ts = st["mem_target_parents"], st["mem_target_survivors"]
                assert Tp > 0 or st["num_parent_kept"] <= 1, what
                k1 = expected_parent_cap(ref, r, Tp)
                assert st["num_parent_kept"] == min(k1, ref.num_object), (what, st["num_parent_kept"], k1)
                uniq, ust = orc.simulate_caps(ref, r, k1, U64MAX, normalize=False)
                assert ust["num_unique"] == st["num_unique"], what
                K = expected_survivor_cap(uniq, r, Ts)
                assert st["num_survivors"] == K, (what, st["num_survivors"], K)
                ref, ost = orc.simulate_caps(ref, r, k1, K)
                assert_states_identical(a.download(), ref, what)
                cut_parents += k1 < st["num_parent"]
                cut_survivors += K < st["num_unique"]
                if ref.num_object == 0:
                    break
    finally:
        ctx.set_tuning(free_mem_mib=old["free_mem_mib"])
        a.close()
        b.close()
    if free_mib <= 8:
        assert cut_parents >= 1 and cut_survivors >= 1, (cut_parents, cut_survivors)


def test_explicit_cap_reports_no_memory_target(ctx):
    hs = qcgd_initial(6, 2)
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    st = q.simulate(a, qcgd.split_merge(math.pi / 4), b, None, 1000)
    assert st["mem_target_parents"] == 0 and st["mem_target_survivors"] == 0
    a.close()
    b.close()
