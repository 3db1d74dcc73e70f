"""Shared helpers of the tests: workloads and state comparisons."""
import math
import random

import numpy as np

from quids_amd.rules import qcgd
from quids_amd.rules import quantum_computer as qc
from quids_amd.state import HostState

PI = math.pi


def assert_states_identical(a, b, what=""):
    """Bit-exact equality of two HostStates in the same order (HIP path vs
    oracle): objects, offsets, sizes, magnitudes (as values) and total_proba."""
    assert a.num_object == b.num_object, f"{what}: num_object {a.num_object} != {b.num_object}"
    assert np.array_equal(a.object_begin, b.object_begin), f"{what}: object_begin differs"
    assert np.array_equal(a.object_size, b.object_size), f"{what}: object_size differs"
    assert np.array_equal(a.objects, b.objects), f"{what}: object bytes differ"
    if a.num_object:
        da = a.mag_pairs()
        db = b.mag_pairs()
        bad = np.nonzero(np.any(da != db, axis=1))[0]
        assert bad.size == 0, (f"{what}: {bad.size} magnitudes differ, first at {bad[0]}: "
                               f"{a.magnitude[bad[0]]!r} vs {b.magnitude[bad[0]]!r}")
    assert a.total_proba == b.total_proba, f"{what}: total_proba {a.total_proba!r} != {b.total_proba!r}"


def keyed(objs, mags, hashes, fields=None):
    """{hash: (fields-or-bytes, mag)}."""
    out = {}
    for o, m, h in zip(objs, mags, hashes):
        out[int(h)] = (fields(o) if fields else bytes(o), complex(m))
    return out


def compare_keyed(ref, mine, rtol=1e-12, band=1e-12, cut=None, what=""):
    """Set parity keyed by hash (the reference's dedup key): identical keys,
    identical objects, magnitudes within rtol relative to the larger of the
    two, or of the state's largest magnitude times rtol for entries that
    cancelled down.  With a truncation cut (the K-th largest |m|^2), entries
    whose |m|^2 lies within `band` relative of it are compared by count only."""
    scale = max([abs(v[1]) for v in ref.values()] + [1e-300])
    if cut is not None:
        lo, hi = cut * (1 - band), cut * (1 + band)
        in_band = lambda m: lo <= abs(m) ** 2 <= hi
        ref_core = {h: v for h, v in ref.items() if not in_band(v[1])}
        mine_core = {h: v for h, v in mine.items() if not in_band(v[1])}
        n_band_ref = len(ref) - len(ref_core)
        n_band_mine = len(mine) - len(mine_core)
        assert n_band_ref == n_band_mine, f"{what}: {n_band_ref} vs {n_band_mine} objects at the cut"
        ref, mine = ref_core, mine_core
    missing = set(ref) ^ set(mine)
    assert not missing, f"{what}: {len(missing)} keys differ ({len(ref)} vs {len(mine)})"
    worst = 0.0
    for h, (o, m) in ref.items():
        o2, m2 = mine[h]
        assert o == o2, f"{what}: object with hash {h:#x} differs"
        err = abs(m - m2) / max(abs(m), abs(m2), scale * 1e-6)
        worst = max(worst, err)
    assert worst <= rtol, f"{what}: worst magnitude error {worst:.3e} > {rtol}"
    return worst


def state_keyed(hs, rule, fields=None):
    objs = hs.objects_list()
    return keyed(objs, hs.magnitude, [rule.hasher(o) for o in objs], fields)


def kth_norm(hs, k):
    n = np.sort(np.abs(hs.magnitude) ** 2)[::-1]
    return n[k - 1] if 0 < k <= n.size else None


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def qcgd_iteration_rules(theta_sm=PI / 4, phi_sm=0.0, theta_ec=PI / 3, phi_ec=PI / 5):
    """One QCGD iteration of the benchmark workloads: step, split_merge, erase_create."""
    return [qcgd.step, qcgd.split_merge(theta_sm, phi_sm), qcgd.erase_create(theta_ec, phi_ec)]


def qcgd_initial(n_nodes, seed):
    return HostState.from_objects([qcgd.seeded_graph(n_nodes, seed)])


def random_circuit(n_qubits, n_steps, seed):
    """Seeded random circuit: each step is H on a uniform qubit (probability
    1/2) or a modifier: CNOT on distinct random qubits, or Y / Z / X on a
    random qubit (the reference's gate definitions)."""
    rng = random.Random(seed)
    ops = []
    for _ in range(n_steps):
        c = rng.random()
        if c < 0.5:
            ops.append(qc.hadamard(rng.randrange(n_qubits)))
        elif c < 0.7:
            a, b = rng.sample(range(n_qubits), 2)
            ops.append(qc.cnot(a, b))
        elif c < 0.8:
            ops.append(qc.Ygate(rng.randrange(n_qubits)))
        elif c < 0.9:
            ops.append(qc.Zgate(rng.randrange(n_qubits)))
        else:
            ops.append(qc.Xgate(rng.randrange(n_qubits)))
    return ops


def random_graph(n, rng):
    """A graph with random particles and a random history of names, built by
    applying random split/merge children (so pair / .l / .r names occur)."""
    L = [rng.random() < 0.5 for _ in range(n)]
    R = [rng.random() < 0.5 for _ in range(n)]
    return qcgd.with_particles(qcgd.make_graph(n), L, R)


def sparse_graph(n, splits, merges):
    """An n-node graph with particles on both sides of the nodes in splits
    and a merge pair (left on i, right on i + 1) for each i in merges, so
    that split_merge and coin have few decisions even when n is large."""
    left, right = [0] * n, [0] * n
    for i in splits:
        left[i] = right[i] = 1
    for i in merges:  # node i merges with i + 1 (left[i + 1] stays 0)
        left[i] = right[i + 1] = 1
    return qcgd.with_particles(qcgd.make_graph(n), left, right)
