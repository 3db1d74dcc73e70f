"""Batched simulation: many independent QCGD simulations advanced together.

A ``batch_t`` keeps its instances back to back in one device state (the
reference's it_t layout, src/quids.hpp:266-276) with the instances' object
offsets on the host.  ``simulate_batched`` advances every instance by one
rule application in one launch sequence of the library
(``qd_simulate_batched``): each instance has its own rule coefficients and
``max_num_object``, and its result is bit for bit what its own
``simulate()`` call (src/quids.hpp:448-543) would return.  Modifiers act on
objects one at a time, so ``simulate(batch.state, modifier)`` applies one to
every instance.
"""
import ctypes

import numpy as np

from . import _params, it_t, simulate
from ._abi import QdRule, QdStepStats, check, lib
from .state import HostState


class batch_t:
    """Instances back to back in one device state, with their offsets."""

    def __init__(self, ctx=None):
        self.state = it_t(ctx=ctx)
        self.seg = np.zeros(1, np.uint64)

    @classmethod
    def from_states(cls, states, ctx=None):
        """A batch of HostStates (one per instance)."""
        b = cls(ctx)
        objs, mags, seg = [], [], [0]
        for hs in states:
            for i in range(hs.num_object):
                o, m = hs.get_object(i)
                objs.append(o)
                mags.append(m)
            seg.append(len(objs))
        from . import align_byte_length
        b.state.upload(HostState.from_objects(objs, mags, int(align_byte_length)))
        b.seg = np.array(seg, np.uint64)
        return b

    @property
    def num_instances(self):
        return int(self.seg.shape[0]) - 1

    def sizes(self):
        return np.diff(self.seg.astype(np.int64))

    def instance(self, i, hs=None):
        """Instance i as a HostState (offsets rebased to 0)."""
        hs = hs if hs is not None else self.state.download()
        a, b = int(self.seg[i]), int(self.seg[i + 1])
        o0, o1 = int(hs.object_begin[a]), int(hs.object_begin[b])
        return HostState(hs.objects[o0:o1].copy(), (hs.object_begin[a:b + 1] - hs.object_begin[a]).copy(),
                         hs.object_size[a:b].copy(), hs.magnitude[a:b].copy())

    def instances(self):
        hs = self.state.download()
        return [self.instance(i, hs) for i in range(self.num_instances)]

    def close(self):
        self.state.close()


def simulate_batched(batch, rules, next_batch, max_num_object):
    """One step of every instance: rules[i] (QCGD rules of one kind) on
    instance i with cap max_num_object[i] (an int applies to all).  Returns
    the per-instance counters (dicts, as simulate() returns), with
    'total_proba' the instance's probability before normalisation.
    1 to 4096 instances; an empty instance steps to an empty one, and an
    instance holding a graph of more than 64 nodes steps on its own inside
    the call, with the same results."""
    n = batch.num_instances
    if len(rules) != n:
        raise ValueError(f"{len(rules)} rules for {n} instances")
    caps = np.full(n, max_num_object, np.uint64) if np.isscalar(max_num_object) else np.asarray(max_num_object,
                                                                                                  np.uint64)
    batch.state._sync_to_device()
    descs = (QdRule * n)(*[r.desc() for r in rules])
    seg = np.ascontiguousarray(batch.seg, np.uint64)
    out_seg = np.zeros(n + 1, np.uint64)
    stats = (QdStepStats * n)()
    prm = _params()
    p64 = ctypes.POINTER(ctypes.c_uint64)
    check(lib().qd_simulate_batched(batch.state.ctx.handle, batch.state.handle, seg.ctypes.data_as(p64), n, descs,
                                    next_batch.state.handle, out_seg.ctypes.data_as(p64), caps.ctypes.data_as(p64),
                                    ctypes.byref(prm), stats), "qd_simulate_batched")
    next_batch.state._device_changed()
    next_batch.seg = out_seg
    return [stats[i].as_dict() for i in range(n)]


def apply_modifier(batch, mod):
    """simulate(it, modifier) on every instance."""
    simulate(batch.state, mod)
