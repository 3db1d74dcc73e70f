"""Brute-force dense state vector of the quantum_computer rules (numpy),
with the reference's gate conventions (quantum_computer.hpp:25-75), against
the oracle on random 8-qubit circuits (CPU).  Independent of the reference
build: it pins Hadamard, CNOT and the phase gates by their definitions."""
import numpy as np
import pytest

from oracle import oracle as orc
from quids_amd.rules import quantum_computer as qc
from quids_amd.state import HostState

from helpers import random_circuit

NQ = 8


def index(reg):
    return sum(int(b) << i for i, b in enumerate(reg))


def dense_apply(psi, op):
    out = np.zeros_like(psi)
    for x in range(psi.size):
        a = psi[x]
        if a == 0:
            continue
        bits = [(x >> i) & 1 for i in range(NQ)]
        if isinstance(op, qc.hadamard):
            b = bits[op.bit]
            flipped = x ^ (1 << op.bit)
            out[flipped] += a / np.sqrt(2)
            out[x] += (-a if b else a) / np.sqrt(2)
        elif isinstance(op, qc.cnot):
            out[x ^ ((bits[op.control_bit]) << op.bit)] += a
        elif isinstance(op, qc.Xgate):
            out[x ^ (1 << op.bit)] += a
        elif isinstance(op, qc.Ygate):
            out[x ^ (1 << op.bit)] += (-1j * a) if bits[op.bit] else (1j * a)
        elif isinstance(op, qc.Zgate):
            out[x ^ (1 << op.bit)] += -a if bits[op.bit] else a
    return out


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_dense_vs_oracle(seed):
    psi = np.zeros(1 << NQ, complex)
    psi[0] = 1
    hs = HostState.from_objects([bytes(NQ)])
    for op in random_circuit(NQ, 24, seed):
        psi = dense_apply(psi, op)
        if op.is_modifier:
            hs = orc.apply_modifier(hs, op)
        else:
            hs, st = orc.simulate(hs, op)
            psi = psi / np.linalg.norm(psi)
            assert abs(hs.total_proba - 1) < 1e-12
    got = np.zeros_like(psi)
    for i in range(hs.num_object):
        o, m = hs.get_object(i)
        got[index(o)] = m
    keep = np.abs(psi) ** 2 > 1e-30
    assert np.allclose(got[keep], psi[keep], atol=1e-13)
    assert np.count_nonzero(got) == np.count_nonzero(keep)
