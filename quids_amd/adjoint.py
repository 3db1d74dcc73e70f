"""Gradients of QCGD expectation values with respect to the rules' angles.

A program is the reference's driver loop (src/rules/qcgd.hpp:1168-1176):
``n_iter`` iterations, each applying a fixed sequence of modifiers and rules
with ``simulate`` (src/quids.hpp:436-543); the rules' real parameters (theta,
phi of split_merge and erase_create, qcgd.hpp:607-619, 459-471) come from
one parameter vector.  The loss is ``average_value`` of an observable of the
final state (src/quids.hpp:208-234).

The forward pass runs on the device and keeps a copy of the state every
``checkpoint_every`` iterations (the last segment keeps all its states).
The backward pass recomputes each earlier segment from its checkpoint and
runs the library's adjoint step
(``qd_simulate_adjoint``: the transpose of the interference scatter-add as a
hash-keyed gather, the truncation held fixed) and modifier adjoint through it
in reverse.  Cotangent vectors live in torch device tensors; the library
reads and writes them through plain device pointers.
"""
import ctypes

import numpy as np
import torch

from . import UINT64_MAX, it_t, simulate
from ._abi import QdRule, check, lib

OBS_NUM_NODES, OBS_PARTICLES, OBS_NODES_EQ = 0, 1, 2


class RuleOp:
    """A rule of a program whose real parameters are params[index[0]],
    params[index[1]], ... followed by the constants ``fixed``."""

    def __init__(self, cls, index, fixed=()):
        self.cls, self.index, self.fixed = cls, tuple(index), tuple(fixed)

    def make(self, params):
        return self.cls(*[float(params[i]) for i in self.index], *self.fixed)


class Program:
    """``ops``: modifiers and RuleOps applied in order, ``n_iter`` times, with
    ``max_num_object`` passed to every simulate() call."""

    def __init__(self, ops, n_iter, max_num_object=UINT64_MAX):
        self.ops, self.n_iter, self.max_num_object = list(ops), int(n_iter), max_num_object
        self.num_params = 1 + max((i for op in self.ops if isinstance(op, RuleOp) for i in op.index), default=-1)


def qcgd_program(n_iter, max_num_object=UINT64_MAX):
    """step -> split_merge(p0, p1) -> erase_create(p2, p3), n_iter times."""
    from .rules import qcgd
    return Program([qcgd.step, RuleOp(qcgd.split_merge, (0, 1)), RuleOp(qcgd.erase_create, (2, 3))], n_iter,
                   max_num_object)


def _device(it):
    return torch.device("cuda", it.ctx.device)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def average_value(it, observable, arg=0, with_cot=False):
    """Device average_value of a QCGD observable (OBS_*); with_cot also
    returns the (N, 2) float64 device tensor 2 obs m (the cotangent of the
    magnitudes)."""
    it._sync_to_device()
    n = it.num_object
    cot = torch.empty((max(n, 1), 2), dtype=torch.float64, device=_device(it)) if with_cot else None
    v = ctypes.c_double()
    torch.cuda.synchronize(_device(it))
    check(lib().qd_average_value(it.ctx.handle, it.handle, int(observable), int(arg), ctypes.byref(v), _ptr(cot)),
          "qd_average_value")
    return (v.value, cot[:n]) if with_cot else v.value


def average_value_of(it, obs, with_cot=False):
    """Device average_value of any observable (the reference's observable_t,
    src/quids.hpp:208-234): obs is a float64 device tensor holding the
    observable of every object, computed by the caller's own kernel
    (quids_amd/user_rule.hpp, eval_observable); same summation order and
    cotangent as average_value."""
    it._sync_to_device()
    n = it.num_object
    cot = torch.empty((max(n, 1), 2), dtype=torch.float64, device=_device(it)) if with_cot else None
    v = ctypes.c_double()
    torch.cuda.synchronize(_device(it))
    check(lib().qd_average_value_of(it.ctx.handle, it.handle, _ptr(obs), ctypes.byref(v), _ptr(cot)),
          "qd_average_value_of")
    return (v.value, cot[:n]) if with_cot else v.value


def step_adjoint(it_in, rule, it_out, out_cot, ndir, max_num_object=UINT64_MAX):
    """Adjoint of simulate(it_in, rule, it_out): (cotangent of it_in's
    magnitudes as an (N, 2) device tensor, [dL/dparam_i for i < ndir])."""
    n = it_in.num_object
    in_cot = torch.empty((max(n, 1), 2), dtype=torch.float64, device=_device(it_in))
    descs = rule.tangent_descs(ndir) if ndir else []
    arr = (QdRule * max(ndir, 1))(*descs)
    grad = (ctypes.c_double * 3)()
    r = rule.desc()
    out_cot = out_cot.contiguous()
    torch.cuda.synchronize(_device(it_in))
    check(lib().qd_simulate_adjoint(it_in.ctx.handle, it_in.handle, ctypes.byref(r), arr, int(ndir), it_out.handle,
                                    _ptr(out_cot), _ptr(in_cot), grad, int(max_num_object)), "qd_simulate_adjoint")
    return in_cot[:n], [grad[i] for i in range(ndir)]


def modifier_adjoint(it_after, mod, cot):
    """Cotangent through simulate(it, modifier), in place; it_after is the
    state after the modifier."""
    d = mod.desc()
    torch.cuda.synchronize(_device(it_after))
    check(lib().qd_modifier_adjoint(it_after.ctx.handle, it_after.handle, ctypes.byref(d), _ptr(cot)),
          "qd_modifier_adjoint")
    return cot


def _iterate(program, params, cur, records=None):
    """One iteration from `cur` (owned by the call).  Without records the
    consumed states are freed; with records every (op, input, output) is
    kept for the adjoint, and modifiers act on a copy so that a recorded rule
    output stays as the rule produced it."""
    for op in program.ops:
        if isinstance(op, RuleOp):
            rule = op.make(params)
            nxt = it_t(ctx=cur.ctx)
            simulate(cur, rule, nxt, None, program.max_num_object)
            if records is None:
                cur.close()
            else:
                records.append(("rule", op, rule, cur, nxt))
            cur = nxt
        else:
            if records is not None:
                cur = cur.copy()
                simulate(cur, op)
                records.append(("mod", op, cur))
            else:
                simulate(cur, op)
    return cur


def run(program, params, initial):
    """Forward only: the final state (a new it_t)."""
    cur = initial.copy()
    for _ in range(program.n_iter):
        cur = _iterate(program, params, cur)
    return cur


def loss(program, params, initial, observable, arg=0):
    fin = run(program, params, initial)
    v = average_value(fin, observable, arg)
    fin.close()
    return v


def forward_checkpointed(program, params, initial, observable, arg=0, checkpoint_every=2):
    """Forward pass keeping a copy of the state before every
    checkpoint_every-th iteration: (loss, checkpoints, cotangent of the final
    state's magnitudes as an (N, 2) device tensor).  The last segment keeps
    its intermediate states (its records) instead of a copy of its first
    state, so the backward pass does not recompute it."""
    params = np.asarray(params, np.float64)
    k = max(int(checkpoint_every), 1)
    last = ((program.n_iter - 1) // k) * k if program.n_iter else 0
    checkpoints = {}
    cur = initial.copy()
    tail = []
    for i in range(program.n_iter):
        if i >= last:
            cur = _iterate(program, params, cur, tail)
            continue
        if i % k == 0:
            checkpoints[i] = cur.copy()
        cur = _iterate(program, params, cur)
    value, g = average_value(cur, observable, arg, with_cot=True)
    if tail:
        checkpoints[last] = tail  # the final state is the last record's output
    else:
        cur.close()
    return value, checkpoints, g


def backward_checkpointed(program, params, checkpoints, g):
    """Runs the adjoint steps through each segment in reverse, starting from
    the cotangent g of the final state; a segment stored as a state is first
    recomputed from it, one stored as records is used as it is (both are
    consumed).  Returns dloss/dparams."""
    params = np.asarray(params, np.float64)
    grads = np.zeros(max(program.num_params, len(params)), np.float64)
    starts = sorted(checkpoints)
    for si in reversed(range(len(starts))):
        i0 = starts[si]
        i1 = starts[si + 1] if si + 1 < len(starts) else program.n_iter
        saved = checkpoints.pop(i0)
        if isinstance(saved, list):
            records = saved
        else:
            records = []
            cur = saved
            for _ in range(i0, i1):
                cur = _iterate(program, params, cur, records)
        for rec in reversed(records):
            if rec[0] == "rule":
                _, op, rule, it_in, it_out = rec
                g, gr = step_adjoint(it_in, rule, it_out, g, len(op.index), program.max_num_object)
                for t, pi in enumerate(op.index):
                    grads[pi] += gr[t]
            else:
                _, mod, it_after = rec
                modifier_adjoint(it_after, mod, g)
        seen = set()
        for rec in records:
            for st in (rec[3:] if rec[0] == "rule" else rec[2:]):
                if id(st) not in seen:
                    seen.add(id(st))
                    st.close()
    return grads[:len(params)]


def gradient(program, params, initial, observable, arg=0, checkpoint_every=2, timings=None):
    """(loss, dloss/dparams).  `timings`, a dict, receives 'forward' and
    'backward' wall seconds (device work included)."""
    import time
    dev = _device(initial)
    t0 = time.perf_counter()
    value, cps, g = forward_checkpointed(program, params, initial, observable, arg, checkpoint_every)
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    grads = backward_checkpointed(program, params, cps, g)
    torch.cuda.synchronize(dev)
    t2 = time.perf_counter()
    if timings is not None:
        timings["forward"] = t1 - t0
        timings["backward"] = t2 - t1
    return value, grads


class _Expectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, program, initial, observable, arg, checkpoint_every):
        p = params.detach().cpu().double().numpy()
        value, cps, g = forward_checkpointed(program, p, initial, observable, arg, checkpoint_every)
        ctx.program, ctx.p, ctx.cps, ctx.g = program, p, cps, g
        return torch.tensor(value, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.g * grad_out.to(ctx.g.device)
        grads = backward_checkpointed(ctx.program, ctx.p, ctx.cps, g)
        ctx.cps, ctx.g = None, None
        return torch.from_numpy(grads), None, None, None, None, None


def expectation(params, program, initial, observable, arg=0, checkpoint_every=2):
    """average_value(observable) of the program's final state as a
    differentiable function of the float64 parameter tensor `params`
    (torch.autograd): the forward pass keeps checkpoints, backward()
    recomputes each segment and runs the device adjoint through it."""
    return _Expectation.apply(params, program, initial, observable, arg, checkpoint_every)
