"""The benchmark workloads on the HIP path against the oracle, at caps the
oracle finishes in seconds:

* 12-node seeded QCGD (cycle graph, srand(42) particles), the benchmark's
  iteration step -> split_merge(pi/4, 0) -> erase_create(pi/3, pi/5);
* 16-node seeded QCGD, in one pass and with the large steps forced into
  hash-range passes (the path that steps beyond device memory take);
* 256 independent instances of 6..16 nodes with per-instance angles,
  advanced by the batched step, each compared with its own oracle run;
* the 10-node, 6-iteration adjoint program (checkpoint every 2 iterations)
  against the oracle's forward-mode gradient;
* the adjoint step's parameter gradients (theta, phi, xi) under random
  complex cotangents, where the phi components are not zero.

Every state comparison is bit for bit (objects, offsets, magnitudes,
total_proba); gradients within 1e-10 (programs) or 1e-12 (one step) of the
largest component."""
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd import adjoint as adj
from quids_amd.rules import qcgd
from quids_amd.state import HostState

from helpers import assert_states_identical, qcgd_initial, qcgd_iteration_rules

pytestmark = pytest.mark.gpu


def _chain_both(hs, n_iter, cap, label):
    """The benchmark iteration n_iter times on the device (states resident,
    ping-pong) and on the oracle, compared after every rule step.  Returns
    the per-step device counters."""
    rules = qcgd_iteration_rules()
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    ref = hs
    stats = []
    for it in range(n_iter):
        for r in rules:
            if r.is_modifier:
                q.simulate(a, r)
                ref = orc.apply_modifier(ref, r)
                continue
            st = q.simulate(a, r, b, None, cap)
            a, b = b, a
            ref, ost = orc.simulate(ref, r, cap)
            what = f"{label} iteration {it + 1} {type(r).__name__}"
            assert st["num_candidates"] == ost["num_candidates"], what
            assert st["num_unique"] == ost["num_unique"], what
            assert st["num_survivors"] == ost["num_survivors"], what
            assert_states_identical(a.download(), ref, what)
            stats.append(st)
    a.close()
    b.close()
    return stats


def test_qcgd12_seeded_against_oracle(ctx):
    """12-node cycle graph, seed 42, cap 2^12, 4 iterations (0.5-0.7 million
    candidates per step from the third iteration on, truncated)."""
    print("qcgd 12 nodes seed 42 cap 2^12 4 iterations", flush=True)
    stats = _chain_both(qcgd_initial(12, 42), 4, 1 << 12, "12 nodes")
    assert max(s["num_candidates"] for s in stats) > 500_000
    assert sum(s["num_unique"] > (1 << 12) for s in stats) >= 4  # the top-K select ran


@pytest.mark.parametrize("chunk_log2", [0, 18])
def test_qcgd16_seeded_against_oracle(ctx, chunk_log2):
    """16-node cycle graph, seed 42, cap 2^12, 3 iterations (1-2.4 million
    candidates per step).  With a candidate budget of 2^18 per pass the
    large steps run in hash-range passes."""
    print(f"qcgd 16 nodes seed 42 cap 2^12 3 iterations chunk_log2={chunk_log2}", flush=True)
    old = ctx.get_tuning()
    ctx.set_tuning(chunk_log2=chunk_log2)
    try:
        stats = _chain_both(qcgd_initial(16, 42), 3, 1 << 12, f"16 nodes chunk {chunk_log2}")
    finally:
        ctx.set_tuning(chunk_log2=old["chunk_log2"])
    ranged = sum(bool(s["reserved"] & 4) for s in stats)
    assert ranged >= (3 if chunk_log2 else 0) and (chunk_log2 or ranged == 0)


def _batch_specs(n_inst, seed):
    """Node counts uniform in 6..16, per-instance particle seeds, (theta, phi)
    uniform in [0, pi/2] x [0, pi] (the same angles drive both rules of an
    instance), as the batched benchmark draws them."""
    rng = np.random.default_rng(seed)
    nodes = rng.integers(6, 17, n_inst)
    seeds = rng.integers(0, 2 ** 31 - 1, n_inst)
    theta = rng.uniform(0, math.pi / 2, n_inst)
    phi = rng.uniform(0, math.pi, n_inst)
    return [(int(nodes[i]), int(seeds[i]), float(theta[i]), float(phi[i])) for i in range(n_inst)]


def _oracle_instance(hs, spec, n_iter, cap):
    """One instance's own oracle run: [(state, counters)] per rule step."""
    orc.set_threads(1)
    _, _, th, ph = spec
    out = []
    for _ in range(n_iter):
        hs = orc.apply_modifier(hs, qcgd.step)
        for r in (qcgd.split_merge(th, ph), qcgd.erase_create(th, ph)):
            hs, st = orc.simulate(hs, r, cap)
            out.append((hs, st))
    return out


def test_batched_256_instances_against_oracle(ctx):
    """256 instances of 6..16 nodes, per-instance angles, cap 2^8 per
    instance, 3 iterations: the batched step (simple truncation active in
    most instance-steps) against every instance's own oracle run."""
    from quids_amd.batch import apply_modifier, batch_t, simulate_batched
    n_inst, n_iter, cap = 256, 3, 1 << 8
    specs = _batch_specs(n_inst, 4242)
    print(f"batched {n_inst} instances cap 2^8 {n_iter} iterations", flush=True)
    states = [HostState.from_objects([qcgd.seeded_graph(n, s)]) for n, s, _, _ in specs]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        refs = list(ex.map(lambda i: _oracle_instance(states[i], specs[i], n_iter, cap), range(n_inst)))
    print("oracle runs done", flush=True)
    a, b = batch_t.from_states(states), batch_t()
    truncated = 0
    k = 0
    for it in range(n_iter):
        apply_modifier(a, qcgd.step)
        for which in (0, 1):
            cls = qcgd.split_merge if which == 0 else qcgd.erase_create
            stats = simulate_batched(a, [cls(th, ph) for _, _, th, ph in specs], b, cap)
            a, b = b, a
            got = a.instances()
            for i in range(n_inst):
                ref, ost = refs[i][k]
                what = f"iteration {it + 1} rule {which} instance {i} ({specs[i][0]} nodes)"
                for key in ("num_candidates", "num_unique", "num_survivors"):
                    assert stats[i][key] == ost[key], (what, key)
                assert stats[i]["total_proba"] == ost["total_proba"], what
                got_i = HostState(got[i].objects, got[i].object_begin, got[i].object_size, got[i].magnitude,
                                  ost["total_proba"])
                ref_i = HostState(ref.objects, ref.object_begin, ref.object_size, ref.magnitude,
                                  ost["total_proba"])
                assert_states_identical(got_i, ref_i, what)
                truncated += ost["num_unique"] > cap
            k += 1
    a.close()
    b.close()
    assert truncated > n_inst  # most instance-steps after the first were cut to the cap


@pytest.mark.parametrize("obs,arg", [(adj.OBS_PARTICLES, 0), (adj.OBS_NODES_EQ, 10)])
def test_adjoint_program_10_nodes_against_oracle(ctx, obs, arg):
    """The adjoint workload: 10-node graph (seed 42), 6 iterations of step ->
    split_merge(t1, p1) -> erase_create(t2, p2), checkpoint every 2
    iterations, every step truncated to 2^12 (the mask held fixed), loss =
    expected particle count or P(10 nodes); gradient in (t1, p1, t2, p2)
    against the oracle's forward mode."""
    print(f"adjoint 10 nodes seed 42 6 iterations cap 2^12 obs={obs}", flush=True)
    hs = qcgd_initial(10, 42)
    it = q.it_t()
    it.upload(hs)
    prog = adj.qcgd_program(6, 1 << 12)
    p = np.array([math.pi / 4, 0.0, math.pi / 3, math.pi / 5])
    v_ref, g_ref, _ = orc.program_jvp(hs, prog, p, obs, arg)
    v, g = adj.gradient(prog, p, it, obs, arg, checkpoint_every=2)
    it.close()
    assert abs(v - v_ref) <= 1e-12 * max(1.0, abs(v_ref)), (v, v_ref)
    scale = np.abs(g_ref).max()
    assert scale > 0
    assert np.abs(np.asarray(g) - g_ref).max() <= 1e-10 * scale, (g, g_ref)


def _small_state(nodes, seed, n_iter):
    """A few dozen objects with varied magnitudes: n_iter oracle iterations
    from a seeded graph."""
    hs = qcgd_initial(nodes, seed)
    for _ in range(n_iter):
        for r in qcgd_iteration_rules(0.7, 0.3, 1.1, 0.4):
            hs = orc.apply_modifier(hs, r) if r.is_modifier else orc.simulate(hs, r)[0]
    return hs


@pytest.mark.parametrize("rule", [qcgd.split_merge(0.7, 0.9, 0.3), qcgd.erase_create(1.1, 0.4, -0.6),
                                  qcgd.coin(0.5, 1.3, 0.2)], ids=["split_merge", "erase_create", "coin"])
def test_step_adjoint_parameter_gradients_match_oracle_jvp(ctx, rule):
    """One step, random complex output cotangent g: the device's dL/dtheta,
    dL/dphi and dL/dxi (L = Re<g, out>) and the input cotangent against
    Re<g, J e> with J e the oracle's forward-mode tangents along every
    parameter and every input magnitude direction.  With a complex g the
    phi and xi derivatives are far from zero (they vanish only for losses
    of |m|^2)."""
    import torch
    hs = _small_state(5, 6, 2)
    n = hs.num_object
    assert 10 <= n <= 200, n
    D = 3 + 2 * n
    tan = np.zeros((n, D), complex)
    for i in range(n):
        tan[i, 3 + 2 * i], tan[i, 3 + 2 * i + 1] = 1, 1j
    from quids_amd._abi import QdRule
    dcoef = list(rule.tangent_descs(3)) + [QdRule(kind=rule.kind)] * (2 * n)
    cols = []  # the oracle carries a few directions per step
    for d0 in range(0, D, orc.MAX_DIRECTIONS):
        d1 = min(D, d0 + orc.MAX_DIRECTIONS)
        ref, t = orc.simulate_jvp(hs, tan[:, d0:d1], rule, dcoef=dcoef[d0:d1])
        cols.append(t)
    t_out = np.concatenate(cols, axis=1)
    it, out = q.it_t(), q.it_t()
    it.upload(hs)
    q.simulate(it, rule, out, None, q.UINT64_MAX)
    assert_states_identical(out.download(), ref, "forward")
    K = out.num_object
    rng = np.random.default_rng(23)
    g = rng.normal(size=K) + 1j * rng.normal(size=K)
    cot = torch.tensor(np.stack([g.real, g.imag], 1), dtype=torch.float64, device="cuda")
    in_cot, grad = adj.step_adjoint(it, rule, out, cot, 3)
    want = np.array([np.real(np.vdot(g, t_out[:, d])) for d in range(D)])
    # parameters: every component to 1e-12 of the largest, phi and xi nonzero
    wp = want[:3]
    assert np.abs(wp[1]) > 1e-3 * np.abs(wp).max() and np.abs(wp[2]) > 1e-3 * np.abs(wp).max(), wp
    assert np.abs(np.array(grad) - wp).max() <= 1e-12 * np.abs(wp).max(), (grad, wp)
    c = in_cot.cpu().numpy()
    got_in = np.concatenate([[c[i, 0], c[i, 1]] for i in range(n)])
    wi = want[3:]
    assert np.abs(got_in - wi).max() <= 1e-12 * np.abs(wi).max()
    it.close()
    out.close()
