#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ by running the reference.

The reference headers are compiled in place (oracle/Makefile builds
oracle/_ref/ref_driver against <reference>/src); this script feeds it small
programs and stores inputs and every dumped state as .npz files (no pickles).
Run it in a container where the reference sources exist:

    make -C oracle ref && python tests/golden/gen_golden.py

Each fixture holds: family (0 quantum_computer, 1 qcgd), tolerance, align,
the input state (objects/begin/size/mag), the op list (rows of
type, kind, a, b, x, y, z, max_num_object, dump) and, for every dump d,
arrays d{d}_N/C/U/total_proba and d{d}_hash/mag/size/objects (objects
concatenated in dump order).
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle as orc  # noqa: E402
from quids_amd.rules import qcgd  # noqa: E402
from quids_amd.rules import quantum_computer as qc  # noqa: E402
from quids_amd.state import HostState  # noqa: E402

from helpers import random_circuit  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
U64MAX = (1 << 64) - 1


def save(name, family, hs, ops, tolerance=1e-30, align=8):
    dumps, _ = orc.run_reference(family, hs, ops, tolerance=tolerance, align=align, threads=4)
    arr = {
        "family": np.array(family), "tolerance": np.array(tolerance), "align": np.array(align),
        "in_objects": hs.objects, "in_begin": hs.object_begin, "in_size": hs.object_size, "in_mag": hs.magnitude,
        "ops": np.array([list(op) for op in ops], dtype=np.float64),
        "ndumps": np.array(len(dumps)),
    }
    for d, dp in enumerate(dumps):
        arr[f"d{d}_N"] = np.array(dp["N"], np.uint64)
        arr[f"d{d}_C"] = np.array(dp["C"], np.uint64)
        arr[f"d{d}_U"] = np.array(dp["U"], np.uint64)
        arr[f"d{d}_total_proba"] = np.array(dp["total_proba"])
        arr[f"d{d}_hash"] = dp["hashes"]
        arr[f"d{d}_mag"] = dp["mags"]
        arr[f"d{d}_size"] = np.array([len(o) for o in dp["objects"]], np.uint32)
        arr[f"d{d}_objects"] = np.frombuffer(b"".join(dp["objects"]), np.uint8)
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **arr)
    print(f"{name}: {len(dumps)} dumps, sizes {[int(dp['N']) for dp in dumps]}, {os.path.getsize(path)} bytes")


def ops_of(seq, max_num_object=U64MAX):
    ops = []
    for r in seq:
        if r.is_modifier:
            ops.append(orc.mod_op(r, dump=True))
        else:
            ops.append(orc.rule_op(r, max_num_object, dump=True))
    return ops


def main():
    # examples/quantum_computer_test.cpp:7-50 (align 0, two registers of 4 and 5 qubits)
    hs = HostState.from_objects([bytes([1, 1, 0, 0]), bytes([0, 1, 1, 0, 1])],
                                [1 / math.sqrt(2), 1j / math.sqrt(2)], align=0)
    H1, H2 = qc.hadamard(1), qc.hadamard(2)
    CNOT, X2, Y0, Z3 = qc.cnot(1, 3), qc.Xgate(2), qc.Ygate(0), qc.Zgate(3)
    seq = [H1, H2, CNOT, X2, Y0, Z3, Z3, Y0, X2, CNOT, H2, H1]
    save("qc_example", 0, hs, ops_of(seq), align=0)

    # random 8-qubit circuits, 24 steps
    for seed in (11, 12, 13):
        hs = HostState.from_objects([bytes(8)])
        save(f"qc_circuit_{seed}", 0, hs, ops_of(random_circuit(8, 24, seed)))

    # examples/qcgd_test.cpp:10-50 with an explicit seed (tolerance 1e-15);
    # the example's `coin` is an erase_create(0.25, 0.25)
    hs = HostState.from_objects([qcgd.seeded_graph(6, 5)])
    ec = qcgd.erase_create(0.3333)
    coin_ = qcgd.erase_create(0.25, 0.25)
    sm = qcgd.split_merge(0.25, 0.25, 0.25)
    rsm = qcgd.split_merge(0.25, 0.25, -0.25)
    seq = [qcgd.step, coin_, coin_, ec, sm, qcgd.step, sm, rsm, qcgd.reversed_step, rsm, ec, qcgd.reversed_step]
    save("qcgd_example", 1, hs, ops_of(seq), tolerance=1e-15)

    # QCGD iterations (step, split_merge(pi/4), erase_create(pi/3, pi/5), coin)
    it = [qcgd.step, qcgd.split_merge(math.pi / 4), qcgd.erase_create(math.pi / 3, math.pi / 5),
          qcgd.coin(math.pi / 7, 0.3, 0.2)]
    for n, seed, iters in ((1, 1, 3), (2, 2, 3), (3, 3, 2), (4, 4, 2), (5, 9, 1)):
        hs = HostState.from_objects([qcgd.seeded_graph(n, seed)])
        save(f"qcgd_chain_n{n}_s{seed}", 1, hs, ops_of(it * iters))

    # truncation at a small cap (ties at the cut are compared by count)
    hs = HostState.from_objects([qcgd.seeded_graph(5, 21)])
    save("qcgd_trunc_n5", 1, hs, ops_of(it * 2, max_num_object=60))

    # parents whose first node is a pair with a positive left hmlz and both
    # particles: the first split wraps around (qcgd.hpp:733-748, 837-845)
    cur = HostState.from_objects([qcgd.seeded_graph(4, 4)])
    picked = []
    for _ in range(3):
        for r in it:
            cur = orc.apply_modifier(cur, r) if r.is_modifier else orc.simulate(cur, r)[0]
            for i in range(cur.num_object):
                o, m = cur.get_object(i)
                f = qcgd.decode_fields(o)
                n_, L, R, nb, names = f
                if L[0] and R[0] and names[0][1] >= 0 and names[1][0] > 0:
                    picked.append((o, m))
            if len(picked) >= 40:
                break
        if len(picked) >= 40:
            break
    objs = [o for o, _ in picked[:40]]
    hs = HostState.from_objects(objs, [m for _, m in picked[:40]])
    save("qcgd_first_split_overflow", 1, hs, ops_of([qcgd.split_merge(math.pi / 4), qcgd.split_merge(1.1, 0.7, 0.3)]))


if __name__ == "__main__":
    main()
