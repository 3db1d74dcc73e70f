"""simulate()'s debug hook (mid_step_function, src/quids.hpp:90, 448): the
library reports the reference's phase names, in the reference's order and
once each (src/quids.hpp:454-1016), on every path of the step."""
import math

import pytest

import quids_amd as q
from quids_amd.rules import qcgd
from quids_amd.state import HostState

from helpers import qcgd_initial

pytestmark = pytest.mark.gpu

# the order in which one reference simulate() call reaches its hook calls
REFERENCE_PHASES = ["num_child", "truncate_symbolic - prepare", "truncate_symbolic", "prepare_index",
                    "symbolic_iteration", "compute_collisions - prepare", "compute_collisions - insert",
                    "compute_collisions - finalize", "truncate - prepare", "truncate", "prepare_final", "final",
                    "normalize", "end"]


def _phases(hs, rule, cap, **tuning):
    ctx = q.default_context()
    old = ctx.get_tuning()
    if tuning:
        ctx.set_tuning(**tuning)
    try:
        a, b = q.it_t(), q.it_t()
        a.upload(hs)
        seen = []
        st = q.simulate(a, rule, b, None, cap, mid_step_function=seen.append)
        a.close()
        b.close()
        return seen, st
    finally:
        if tuning:
            ctx.set_tuning(**{k: old[k] for k in tuning})


def _grown(nodes, seed, n_iter=1):
    from oracle import oracle as orc
    hs = qcgd_initial(nodes, seed)
    for _ in range(n_iter):
        for r in (qcgd.step, qcgd.split_merge(math.pi / 4), qcgd.erase_create(math.pi / 3, math.pi / 5)):
            hs = orc.apply_modifier(hs, r) if r.is_modifier else orc.simulate(hs, r, 2000)[0]
    return orc.apply_modifier(hs, qcgd.step)


@pytest.mark.parametrize("case", ["single pass", "truncated", "pre-truncated", "hash-range passes",
                                  "memory-driven", "empty"])
def test_phase_names_follow_the_reference(case):
    print(f"phase hook: {case}", flush=True)
    rule = qcgd.erase_create(math.pi / 3, math.pi / 5)
    if case == "empty":
        seen, st = _phases(HostState(), rule, 100)
    elif case == "hash-range passes":
        seen, st = _phases(_grown(8, 3, 2), rule, 500, chunk_log2=10)
        assert st["reserved"] & 4
    else:
        hs = _grown(7, 3)
        cap = {"single pass": q.UINT64_MAX, "truncated": 50, "pre-truncated": hs.num_object // 2,
               "memory-driven": 0}[case]
        seen, st = _phases(hs, rule, cap)
    assert seen == REFERENCE_PHASES, seen
