"""The oracle against the reference's golden vectors (CPU)."""
import pytest

from oracle import oracle as orc

from golden_util import dump_state, fixtures, load, replay_stepwise


def _oracle_rule(hs, rule, max_num_object, tol, align):
    return orc.simulate(hs, rule, max_num_object=max_num_object, tolerance=tol, align=align)


@pytest.mark.parametrize("name", fixtures())
def test_oracle_matches_reference(name):
    n = replay_stepwise(name, _oracle_rule, orc.apply_modifier)
    assert n > 0


@pytest.mark.parametrize("name", fixtures())
def test_product_hasher_matches_reference(name):
    """The library's host-side rule hasher (the functor code the kernels use)
    gives the reference's hash for every object of every dump."""
    from quids_amd.rules import qcgd
    from quids_amd.rules import quantum_computer as qc
    f = load(name)
    rule = qcgd.split_merge(0.0) if int(f["family"]) == 1 else qc.hadamard(0)
    for d in range(int(f["ndumps"])):
        hs, hashes = dump_state(f, d, int(f["align"]))
        for i in range(hs.num_object):
            assert rule.hasher(hs.get_object(i)[0]) == hashes[i]


def test_reversibility_examples():
    """The reference examples undo their gates: quantum_computer_test ends
    on the initial state times the global phase -1 (Zgate squares to -I in
    the reference's definition), qcgd_test ends on the initial state."""
    import numpy as np
    f = load("qc_example")
    last = int(f["ndumps"]) - 1
    hs, _ = dump_state(f, last, 0)
    mags = sorted(hs.magnitude, key=lambda z: (z.real, z.imag))
    init = sorted(-f["in_mag"], key=lambda z: (z.real, z.imag))
    assert np.allclose(mags, init, atol=1e-12)
    f = load("qcgd_example")
    last = int(f["ndumps"]) - 1
    hs, _ = dump_state(f, last, 8)
    assert hs.num_object == 1 and abs(abs(hs.magnitude[0]) - 1) < 1e-12
