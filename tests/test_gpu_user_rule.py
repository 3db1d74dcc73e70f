"""User rules and observables compiled as device functors (the reference's
plugin API: a quids::rule subclass, src/quids.hpp:105-146, and any
observable_t, 208-234), through qd_simulate_user / qd_average_value_of.

The rules live in tests/cpp/user_rules.hpp, built by tests/cpp/Makefile into
libuser_rules.so the way a user's own code would be: a restatement of the
reference's hadamard (checked against the oracle and, bit for bit, against
the built-in hadamard under every truncation mode) and an RY rotation with
no built-in counterpart and its own hasher (checked against a dense state
vector)."""
import ctypes
import math
import os
import random
import subprocess

import numpy as np
import pytest
import torch

import quids_amd as q
from quids_amd import _abi
from quids_amd.adjoint import OBS_NUM_NODES, average_value, average_value_of
from oracle import oracle as orc
from quids_amd.rules import qcgd
from quids_amd.rules import quantum_computer as qc
from quids_amd.state import HostState

from helpers import assert_states_identical, qcgd_initial, qcgd_iteration_rules, random_circuit

pytestmark = pytest.mark.gpu

U64MAX = (1 << 64) - 1
CPP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")


@pytest.fixture(scope="module")
def plugin():
    path = os.path.join(CPP, "libuser_rules.so")
    if not os.path.exists(path):
        pytest.fail("tests/cpp/libuser_rules.so is not built (make -C tests/cpp)")
    _abi.lib()  # the library first (one HIP runtime, torch's)
    so = ctypes.CDLL(path)
    so.qdt_user_rule.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                 ctypes.POINTER(_abi.QdUserRule)]
    so.qdt_observable.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return so


def user(plugin, kind, bit, a, b=0.0):
    t = _abi.QdUserRule()
    assert plugin.qdt_user_rule(kind, bit, a, b, ctypes.byref(t)) == 0
    return q.user_rule(t, keepalive=plugin)


def user_hadamard(plugin, bit):
    return user(plugin, 0, bit, 1 / math.sqrt(2))


def step(hs, rule, cap=U64MAX, trim=False):
    if trim:  # (memory-driven truncation counts the context's cached scratch as available)
        q.default_context().trim()
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    sy = q.sy_it_t()
    st = q.simulate(a, rule, b, sy, cap)
    return b.download(), st


def modifier(hs, mod):
    a = q.it_t()
    a.upload(hs)
    q.simulate(a, mod)
    return a.download()


@pytest.mark.parametrize("seed", [1, 2])
def test_user_hadamard_matches_oracle(ctx, plugin, seed):
    """The 16-qubit, 24-step circuits of the plain parity test with every
    hadamard run as the user rule: every state bit for bit as the oracle's."""
    q.tolerance, q.align_byte_length = 1e-30, 8
    hs_g = hs_o = HostState.from_objects([bytes(16)])
    for k, op in enumerate(random_circuit(16, 24, seed)):
        if op.is_modifier:
            hs_g = modifier(hs_g, op)
            hs_o = orc.apply_modifier(hs_o, op)
        else:
            hs_g, st = step(hs_g, user_hadamard(plugin, op.bit))
            hs_o, rst = orc.simulate(hs_o, op)
            assert st["num_candidates"] == rst["num_candidates"]
        assert_states_identical(hs_g, hs_o, f"seed {seed} op {k}")


def hadamard_bits(n_qubits, n_steps, seed):
    """Every qubit once (the state doubles), then random qubits."""
    rng = random.Random(seed)
    bits = list(range(n_qubits))
    rng.shuffle(bits)
    return bits + [rng.randrange(n_qubits) for _ in range(n_steps - n_qubits)]


@pytest.mark.parametrize("mode", ["simple", "probabilistic"])
def test_user_hadamard_matches_builtin_under_truncation(ctx, plugin, mode):
    """An explicit cap with simple and with probabilistic keys (the parents'
    hashes then come from the user rule's hasher): the user path cuts where
    the built-in one does, bit for bit."""
    q.tolerance, q.align_byte_length = 1e-30, 8
    prm = (q.simple_truncation, q.truncation_seed)
    try:
        if mode == "probabilistic":
            q.simple_truncation, q.truncation_seed = False, 99
        rng = random.Random(5)
        hs = HostState.from_objects([bytes(12)])
        cut = 0
        for k, bit in enumerate(hadamard_bits(12, 20, 5)):
            if k % 4 == 3:
                a, b = rng.sample(range(12), 2)
                hs = modifier(hs, qc.cnot(a, b))
            g, st = step(hs, qc.hadamard(bit), 90)
            u, ust = step(hs, user_hadamard(plugin, bit), 90)
            for key in ("num_candidates", "num_unique", "num_survivors", "num_parent_kept"):
                assert st[key] == ust[key], (mode, k, key, st[key], ust[key])
            assert_states_identical(u, g, f"{mode} step {k}")
            cut += st["num_survivors"] < st["num_unique"] or st["num_parent_kept"] < st["num_parent"]
            hs = g
        assert cut > 0, "the truncation never cut"
    finally:
        q.simple_truncation, q.truncation_seed = prm


def test_user_hadamard_memory_driven_matches_reference_search(ctx, plugin):
    """Memory-driven truncation (max_num_object 0) with the free memory
    capped so that it cuts: the parent and survivor counts are the
    reference's binary searches over the oracle's data at the targets the
    step reports, and the state is the oracle's at those counts."""
    from test_gpu_memory import expected_parent_cap, expected_survivor_cap
    q.tolerance, q.align_byte_length = 1e-30, 8
    old = ctx.get_tuning()
    # (14 qubits in 128-byte objects: survivors weigh enough against the
    # step's scratch that both sides cut)
    ref = HostState.from_objects([bytes(128)])
    cut_parents = cut_survivors = 0
    try:
        ctx.set_tuning(free_mem_mib=1)
        for k, bit in enumerate(hadamard_bits(14, 16, 9)):
            r = qc.hadamard(bit)
            got, st = step(ref, user_hadamard(plugin, bit), 0, trim=True)
            Tp, Ts = st["mem_target_parents"], st["mem_target_survivors"]
            k1 = expected_parent_cap(ref, r, Tp)
            assert st["num_parent_kept"] == min(k1, ref.num_object), (k, st["num_parent_kept"], k1)
            uniq, ust = orc.simulate_caps(ref, r, k1, U64MAX, normalize=False)
            assert ust["num_unique"] == st["num_unique"], k
            K = expected_survivor_cap(uniq, r, Ts)
            assert st["num_survivors"] == K, (k, st["num_survivors"], K)
            ref, _ = orc.simulate_caps(ref, r, k1, K)
            assert_states_identical(got, ref, f"step {k}")
            cut_parents += k1 < st["num_parent"]
            cut_survivors += K < st["num_unique"]
    finally:
        ctx.set_tuning(free_mem_mib=old["free_mem_mib"])
    assert cut_parents >= 1 and cut_survivors >= 1, (cut_parents, cut_survivors)


def dense_from_state(hs, nq):
    v = np.zeros(1 << nq, np.complex128)
    for i in range(hs.num_object):
        o, m = hs.get_object(i)
        x = sum(1 << j for j in range(nq) if o[j])
        v[x] += m
    return v


def test_user_ry_matches_dense_state_vector(ctx, plugin):
    """A rule with no built-in counterpart: RY rotations (and CNOT
    modifiers) on 8 one-byte qubits against a dense state vector, amplitudes
    within 1e-12; interference cancels amplitudes to zero along the way."""
    q.tolerance, q.align_byte_length = 1e-30, 8
    nq = 8
    rng = random.Random(3)
    hs = HostState.from_objects([bytes(nq)])
    vec = np.zeros(1 << nq, np.complex128)
    vec[0] = 1
    merged = 0
    for k in range(30):
        if k % 5 == 4:
            a, b = rng.sample(range(nq), 2)
            hs = modifier(hs, qc.cnot(a, b))
            nv = np.zeros_like(vec)
            for x in range(1 << nq):
                y = x ^ (((x >> a) & 1) << b)
                nv[y] += vec[x]
            vec = nv
            continue
        bit = rng.randrange(nq)
        th = rng.uniform(0, math.pi)
        c, s = math.cos(th / 2), math.sin(th / 2)
        hs, st = step(hs, user(plugin, 1, bit, c, s))
        merged += st["num_candidates"] - st["num_unique"]
        nv = vec.copy()
        for x in range(1 << nq):
            if not (x >> bit) & 1:
                x1 = x | (1 << bit)
                nv[x] = c * vec[x] - s * vec[x1]
                nv[x1] = s * vec[x] + c * vec[x1]
        vec = nv
        got = dense_from_state(hs, nq)
        assert np.max(np.abs(got - vec)) < 1e-12, k
        assert abs(hs.total_proba - 1) < 1e-12
    assert merged > 0, "no interference happened"


def test_user_observables(ctx, plugin):
    """qd_average_value_of with a user kernel's values: the QCGD node count
    bit for bit as the built-in observable 0, the set-qubit count within
    1e-12 of the host sum, and the cotangent 2 obs m."""
    q.tolerance, q.align_byte_length = 1e-30, 8
    hs = qcgd_initial(6, 4)
    for r in qcgd_iteration_rules():
        hs = modifier(hs, r) if r.is_modifier else step(hs, r)[0]
    it = q.it_t()
    it.upload(hs)
    n = hs.num_object
    obs = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert plugin.qdt_observable(0, ctx.handle, it.handle, ctypes.c_void_p(obs.data_ptr())) == 0
    v, cot = average_value_of(it, obs, with_cot=True)
    ref, rcot = average_value(it, OBS_NUM_NODES, with_cot=True)
    assert v == ref
    assert torch.equal(cot, rcot)

    hq = HostState.from_objects([bytes(10)])
    for bit in (1, 4, 7):
        hq = step(hq, user_hadamard(plugin, bit))[0]
    it2 = q.it_t()
    it2.upload(hq)
    obs2 = torch.empty(hq.num_object, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert plugin.qdt_observable(1, ctx.handle, it2.handle, ctypes.c_void_p(obs2.data_ptr())) == 0
    host = sum(sum(1 for x in hq.get_object(i)[0] if x) * abs(hq.magnitude[i]) ** 2 for i in range(hq.num_object))
    assert abs(average_value_of(it2, obs2) - host) < 1e-12


def test_user_rule_missing_callback_is_rejected(ctx, plugin):
    t = _abi.QdUserRule()
    assert plugin.qdt_user_rule(0, 1, 1 / math.sqrt(2), 0.0, ctypes.byref(t)) == 0
    t.populate = None
    with pytest.raises(q.QuidsError, match="callback is missing"):
        step(HostState.from_objects([bytes(4)]), q.user_rule(t))


def test_cpp_user_rule_example(ctx):
    """The C++ API (quids::device_rule<R>, device_observable<F>): the user
    hadamard against the built-in over a seeded circuit with truncation,
    device observables against host observable_t lambdas, host-side virtual
    methods against the built-in rule's."""
    exe = os.path.join(CPP, "user_rule_example")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/user_rule_example is not built (make -C tests/cpp)")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "user_rule_example ok" in out.stdout
