"""qd_average_value_noalloc (a static device scratch array, no device
allocation) against qd_average_value, both through the C ABI: two calls in
a row, the second on a smaller state, for every QCGD observable; value and
cotangent bit for bit, and the context's scratch pool stays empty.  A state
beyond the scratch's 2^22 objects is rejected before any launch."""
import ctypes

import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd._abi import QD_ERR_CAPACITY, QuidsError, check, lib
from quids_amd.state import HostState

from helpers import qcgd_initial, qcgd_iteration_rules

pytestmark = pytest.mark.gpu


def _state(nodes, seed, n_iter, cap):
    hs = qcgd_initial(nodes, seed)
    for _ in range(n_iter):
        for r in qcgd_iteration_rules():
            hs = orc.apply_modifier(hs, r) if r.is_modifier else orc.simulate(hs, r, cap)[0]
    it = q.it_t()
    it.upload(hs)
    return it


def _call(name, it, obs, arg):
    import torch
    n = it.num_object
    cot = torch.zeros((max(n, 1), 2), dtype=torch.float64, device="cuda")
    v = ctypes.c_double()
    torch.cuda.synchronize()
    check(getattr(lib(), name)(it.ctx.handle, it.handle, int(obs), int(arg), ctypes.byref(v),
                               ctypes.c_void_p(cot.data_ptr())), name)
    return v.value, cot[:n].cpu().numpy()


def test_noalloc_average_value_matches_allocating_one(ctx):
    big = _state(10, 42, 3, 1 << 16)
    small = _state(6, 3, 2, 1 << 10)
    assert big.num_object > 4 * small.num_object > 0
    print(f"noalloc average_value: {big.num_object} then {small.num_object} objects", flush=True)
    for it in (big, small):
        for obs, arg in ((0, 0), (1, 0), (2, 10), (2, 6)):
            ctx.trim()
            v1, c1 = _call("qd_average_value_noalloc", it, obs, arg)
            assert ctx.scratch_bytes() == 0, "the no-allocation entry point grew the scratch pool"
            v0, c0 = _call("qd_average_value", it, obs, arg)
            assert v1 == v0, (obs, arg, v1, v0)
            assert np.array_equal(c1, c0), (obs, arg)
    big.close()
    small.close()


def test_noalloc_rejects_states_beyond_its_scratch(ctx):
    n = (1 << 22) + 1
    hs = HostState(np.zeros(8 * n, np.uint8), np.arange(n + 1, dtype=np.uint64) * 8, np.full(n, 8, np.uint32),
                   np.zeros(n, np.complex128))
    it = q.it_t()
    it.upload(hs)
    with pytest.raises(QuidsError) as e:
        _call("qd_average_value_noalloc", it, 1, 0)
    assert e.value.code == QD_ERR_CAPACITY
    it.close()
