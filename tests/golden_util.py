"""Loading and replaying the golden fixtures of tests/golden/ (written by
tests/golden/gen_golden.py from the reference)."""
import glob
import os

import numpy as np

from quids_amd.rules import qcgd
from quids_amd.rules import quantum_computer as qc
from quids_amd.state import HostState

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OP_RULE, OP_MOD = 1, 2


def fixtures():
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def op_object(row):
    """The rule / modifier object of a program row."""
    typ, kind, a, b, x, y, z = int(row[0]), int(row[1]), int(row[2]), int(row[3]), row[4], row[5], row[6]
    if typ == OP_RULE:
        if kind == 1:
            return qc.hadamard(a)
        return {2: qcgd.erase_create, 3: qcgd.coin, 4: qcgd.split_merge}[kind](x, y, z)
    return {1: lambda: qc.cnot(a, b), 2: lambda: qc.Xgate(b), 3: lambda: qc.Ygate(b), 4: lambda: qc.Zgate(b),
            5: lambda: qcgd.step, 6: lambda: qcgd.reversed_step}[kind]()


def max_num_object(row):
    # stored as float64 in the fixture; UINT64_MAX rounds to 2^64
    v = row[7]
    return (1 << 64) - 1 if v >= 2.0 ** 63 else int(v)


def input_state(f):
    return HostState(f["in_objects"], f["in_begin"], f["in_size"], f["in_mag"])


def dump_state(f, d, align):
    """A dump as a HostState (reference order, objects re-packed with the
    fixture's alignment) and its reference hashes."""
    sizes = f[f"d{d}_size"]
    flat = f[f"d{d}_objects"].tobytes()
    objs, off = [], 0
    for s in sizes:
        objs.append(flat[off:off + int(s)])
        off += int(s)
    hs = HostState.from_objects(objs, f[f"d{d}_mag"], align=align, total_proba=float(f[f"d{d}_total_proba"]))
    return hs, [int(h) for h in f[f"d{d}_hash"]]


def field_key(family):
    if family == 1:
        return qcgd.decode_fields
    return bytes


def compare_to_dump(f, d, got, rule_for_hash, what=""):
    """Compares a computed state with dump d: same object set keyed by hash,
    field-wise equal objects, magnitudes within 1e-12 relative (1e-9 of the
    largest for entries that cancelled), counts and total_proba; when the
    step truncated, entries tied at the cut are compared by count only."""
    family = int(f["family"])
    fk = field_key(family)
    ref_hs, ref_h = dump_state(f, d, int(f["align"]))
    ref = {}
    for i, h in enumerate(ref_h):
        o, m = ref_hs.get_object(i)
        ref[h] = (fk(o), m)
    mine = {}
    for i in range(got.num_object):
        o, m = got.get_object(i)
        mine[rule_for_hash.hasher(o)] = (fk(o), m)
    N = int(f[f"d{d}_N"])
    U = int(f[f"d{d}_U"])
    cut = None
    if N < U:  # truncated: the cut is the smallest kept |m|^2
        cut = min(abs(m) ** 2 for _, m in ref.values())
        lo, hi = cut * (1 - 1e-9), cut * (1 + 1e-9)
        ref = {h: v for h, v in ref.items() if not (lo <= abs(v[1]) ** 2 <= hi)}
        n_tied_ref = N - len(ref)
        mine_core = {h: v for h, v in mine.items() if not (lo <= abs(v[1]) ** 2 <= hi)}
        assert got.num_object - len(mine_core) == n_tied_ref, f"{what}: tied counts differ"
        mine = mine_core
    assert got.num_object == N, f"{what}: {got.num_object} objects, reference {N}"
    assert set(ref) == set(mine), f"{what}: {len(set(ref) ^ set(mine))} hashes differ"
    scale = max([abs(v[1]) for v in ref.values()] + [1e-300])
    worst = 0.0
    for h, (o, m) in ref.items():
        o2, m2 = mine[h]
        assert o == o2, f"{what}: object {h:#x} differs field-wise"
        worst = max(worst, abs(m - m2) / max(abs(m), abs(m2), 1e-3 * scale))
    assert worst <= 1e-12, f"{what}: magnitude error {worst:.3e}"
    assert abs(got.total_proba - float(f[f"d{d}_total_proba"])) <= 1e-12 * max(1.0, got.total_proba), \
        f"{what}: total_proba {got.total_proba!r} vs {float(f[f'd{d}_total_proba'])!r}"
    return worst


def replay_stepwise(name, run_rule, run_mod):
    """Replays every op of a fixture from the reference's own previous state
    (so that tie choices at a truncation cut do not propagate) and compares
    with the dump.  run_rule(hs, rule, max, tol, align) -> (HostState, stats);
    run_mod(hs, mod) -> HostState."""
    f = load(name)
    tol, align = float(f["tolerance"]), int(f["align"])
    ops = f["ops"]
    prev = input_state(f)
    hash_rule = None
    for d, row in enumerate(ops):
        obj = op_object(row)
        what = f"{name} op {d}"
        if int(row[0]) == OP_RULE:
            hash_rule = obj
            got, st = run_rule(prev, obj, max_num_object(row), tol, align)
            assert st["num_candidates"] == int(f[f"d{d}_C"]), f"{what}: C {st['num_candidates']} vs {int(f[f'd{d}_C'])}"
            assert st["num_unique"] == int(f[f"d{d}_U"]), f"{what}: U {st['num_unique']} vs {int(f[f'd{d}_U'])}"
        else:
            got = run_mod(prev, obj)
            got.total_proba = float(f[f"d{d}_total_proba"])
        if hash_rule is None:
            hash_rule = qcgd.split_merge(0.0) if int(f["family"]) == 1 else qc.hadamard(0)
        compare_to_dump(f, d, got, hash_rule, what)
        prev, _ = dump_state(f, d, align)
    return len(ops)
