"""The device adjoint against the oracle's forward-mode derivative (FP64
reference gradient), against central finite differences of the device
forward pass, and through torch.autograd."""
import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd import adjoint as adj
from quids_amd.rules import qcgd
from quids_amd.rules import quantum_computer as qc
from quids_amd.state import HostState

pytestmark = pytest.mark.gpu

P0 = np.array([0.7, 0.3, 1.1, 0.4])


def _initial(nodes, seed):
    hs = HostState.from_objects([qcgd.seeded_graph(nodes, seed)])
    it = q.it_t()
    it.upload(hs)
    return hs, it


def _gerr(g, ref):
    """Largest component error relative to the largest gradient component
    (the phi components are zero up to rounding, so a per-component relative
    error would be meaningless)."""
    return np.abs(np.asarray(g) - ref).max() / max(np.abs(ref).max(), 1e-300)


@pytest.mark.parametrize("nodes,seed,obs,arg,n_iter,cap", [
    (3, 7, adj.OBS_PARTICLES, 0, 3, None),
    (4, 1, adj.OBS_NODES_EQ, 4, 3, None),
    (5, 3, adj.OBS_NUM_NODES, 0, 3, None),
    (6, 5, adj.OBS_PARTICLES, 0, 4, 2000),
    (8, 2, adj.OBS_PARTICLES, 0, 3, 5000),
])
def test_adjoint_matches_oracle_jvp(nodes, seed, obs, arg, n_iter, cap):
    print(f"adjoint nodes={nodes} seed={seed} obs={obs} iters={n_iter} cap={cap}", flush=True)
    hs, it = _initial(nodes, seed)
    prog = adj.qcgd_program(n_iter, cap if cap else q.UINT64_MAX)
    v_ref, g_ref, _ = orc.program_jvp(hs, prog, P0, obs, arg)
    v, g = adj.gradient(prog, P0, it, obs, arg, checkpoint_every=2)
    assert abs(v - v_ref) <= 1e-12 * max(1.0, abs(v_ref))
    assert _gerr(g, g_ref) <= 1e-10, (g, g_ref)


def test_checkpoint_interval_does_not_change_the_gradient():
    _, it = _initial(5, 4)
    prog = adj.qcgd_program(5, 3000)
    ref = adj.gradient(prog, P0, it, adj.OBS_PARTICLES, 0, checkpoint_every=1)
    for k in (2, 3, 100):
        v, g = adj.gradient(prog, P0, it, adj.OBS_PARTICLES, 0, checkpoint_every=k)
        assert v == ref[0] and np.array_equal(g, ref[1]), k


def test_adjoint_matches_device_finite_differences():
    _, it = _initial(4, 9)
    prog = adj.qcgd_program(3)
    _, g = adj.gradient(prog, P0, it, adj.OBS_PARTICLES)
    h = 1e-5
    fd = []
    for i in range(4):
        e = np.zeros(4)
        e[i] = h
        fd.append((adj.loss(prog, P0 + e, it, adj.OBS_PARTICLES) - adj.loss(prog, P0 - e, it, adj.OBS_PARTICLES)) / (2 * h))
    assert _gerr(g, np.array(fd)) <= 1e-8


def test_torch_autograd():
    import torch
    _, it = _initial(4, 3)
    prog = adj.qcgd_program(3)
    v, g = adj.gradient(prog, P0, it, adj.OBS_PARTICLES)
    p = torch.tensor(P0, dtype=torch.float64, requires_grad=True)
    loss = adj.expectation(p, prog, it, adj.OBS_PARTICLES)
    (3.0 * loss).backward()
    assert loss.item() == v
    # the cotangent is scaled by 3 before the adjoint, so rounding may differ
    assert np.allclose(p.grad.numpy(), 3.0 * g, rtol=1e-13, atol=1e-15)


def test_step_and_modifier_cotangents_match_oracle_jvp():
    """Cotangent of the input magnitudes (no parameters): for every input
    direction e, Re<in_cot, e> = Re<out_cot, J e>, with J e from the oracle's
    forward mode; covers hadamard and the Y / Z / CNOT modifier adjoints."""
    import torch
    objs = [bytes([0, 1, 0, 1]), bytes([1, 1, 0, 0]), bytes([0, 0, 1, 1]), bytes([1, 0, 1, 0])]
    mags = np.array([0.5 + 0.1j, 0.3 - 0.2j, 0.6 + 0.4j, -0.2 + 0.15j])
    hs = HostState.from_objects(objs, mags / np.linalg.norm(mags), align=0)
    n = hs.num_object
    D = 2 * n
    tan = np.zeros((n, D), complex)
    for i in range(n):
        tan[i, 2 * i], tan[i, 2 * i + 1] = 1, 1j
    rng = np.random.default_rng(5)
    old_align = q.align_byte_length
    q.align_byte_length = 0
    try:
        for mod in (qc.Ygate(1), qc.Zgate(0), qc.cnot(0, 2)):
            it = q.it_t()
            it.upload(hs)
            q.simulate(it, mod)
            _, t_out = orc.modifier_jvp(hs, tan, mod)
            g = rng.normal(size=n) + 1j * rng.normal(size=n)
            cot = torch.tensor(np.stack([g.real, g.imag], 1), dtype=torch.float64, device="cuda")
            adj.modifier_adjoint(it, mod, cot)
            c = cot.cpu().numpy()
            got = np.concatenate([[c[i, 0], c[i, 1]] for i in range(n)])
            want = np.array([np.real(np.vdot(g, t_out[:, d])) for d in range(D)])
            assert np.abs(got - want).max() <= 1e-14, mod
        for rule in (qc.hadamard(0), qc.hadamard(3)):
            it, out = q.it_t(), q.it_t()
            it.upload(hs)
            q.simulate(it, rule, out, None, q.UINT64_MAX)
            ref, t_out = orc.simulate_jvp(hs, tan, rule, max_num_object=q.UINT64_MAX, align=0)
            K = out.num_object
            assert K == ref.num_object
            g = rng.normal(size=K) + 1j * rng.normal(size=K)
            cot = torch.tensor(np.stack([g.real, g.imag], 1), dtype=torch.float64, device="cuda")
            in_cot, _ = adj.step_adjoint(it, rule, out, cot, 0)
            c = in_cot.cpu().numpy()
            got = np.concatenate([[c[i, 0], c[i, 1]] for i in range(n)])
            # both outputs are in ascending representative order
            want = np.array([np.real(np.vdot(g, t_out[:, d])) for d in range(D)])
            assert np.abs(got - want).max() <= 1e-13, rule
    finally:
        q.align_byte_length = old_align
