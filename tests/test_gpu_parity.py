"""HIP step vs the oracle, bit for bit, on the GPU (through the C ABI)."""
import math

import numpy as np
import pytest

import quids_amd as q
from oracle import oracle as orc
from quids_amd.rules import qcgd
from quids_amd.rules import quantum_computer as qc
from quids_amd.state import HostState

from helpers import assert_states_identical, qcgd_initial, qcgd_iteration_rules, random_circuit, sparse_graph

pytestmark = pytest.mark.gpu

U64MAX = (1 << 64) - 1


def gpu_step(hs, rule, max_num_object=U64MAX, tolerance=1e-30, align=8):
    q.tolerance, q.align_byte_length = tolerance, align
    a, b = q.it_t(), q.it_t()
    a.upload(hs)
    sy = q.sy_it_t()
    st = q.simulate(a, rule, b, sy, max_num_object)
    out = b.download()
    return out, st, sy


def gpu_modifier(hs, mod):
    a = q.it_t()
    a.upload(hs)
    q.simulate(a, mod)
    return a.download()


def test_hadamard_single_object(ctx):
    hs = HostState.from_objects([bytes([1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1])])
    r = qc.hadamard(3)
    out, st, sy = gpu_step(hs, r)
    ref, rst = orc.simulate(hs, r)
    assert st["num_candidates"] == 2 == sy.num_object
    assert_states_identical(out, ref, "hadamard")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_circuit_chain_16_qubits(ctx, seed):
    """A seeded circuit of 24 steps on 16 one-byte qubits, GPU and oracle
    chained independently, compared after every step."""
    hs_g = HostState.from_objects([bytes(16)])
    hs_o = hs_g
    for k, op in enumerate(random_circuit(16, 24, seed)):
        if op.is_modifier:
            hs_g = gpu_modifier(hs_g, op)
            hs_o = orc.apply_modifier(hs_o, op)
        else:
            hs_g, st, _ = gpu_step(hs_g, op)
            hs_o, _ = orc.simulate(hs_o, op)
        assert_states_identical(hs_g, hs_o, f"seed {seed} op {k}")


@pytest.mark.parametrize("nodes,seed", [(3, 1), (4, 2), (6, 7), (8, 11)])
def test_qcgd_chain(ctx, nodes, seed):
    """QCGD iterations (step, split_merge, erase_create), chained on each side."""
    hs_g = hs_o = qcgd_initial(nodes, seed)
    rules = qcgd_iteration_rules()
    for it in range(3):
        for r in rules:
            if r.is_modifier:
                hs_g = gpu_modifier(hs_g, r)
                hs_o = orc.apply_modifier(hs_o, r)
            else:
                hs_g, st, _ = gpu_step(hs_g, r)
                hs_o, ost = orc.simulate(hs_o, r)
                assert st["num_candidates"] == ost["num_candidates"]
                assert st["num_unique"] == ost["num_unique"]
                assert_states_identical(hs_g, hs_o, f"nodes {nodes} it {it} {type(r).__name__}")


@pytest.mark.parametrize("cap", [1, 7, 100, 1000])
def test_qcgd_truncation(ctx, cap):
    """Top-K truncation with ties, and parent pre-truncation."""
    hs = qcgd_initial(6, 3)
    rules = qcgd_iteration_rules()
    hs_g = hs_o = hs
    for it in range(3):
        for r in rules:
            if r.is_modifier:
                hs_g = gpu_modifier(hs_g, r)
                hs_o = orc.apply_modifier(hs_o, r)
            else:
                hs_g, st, _ = gpu_step(hs_g, r, max_num_object=cap)
                hs_o, _ = orc.simulate(hs_o, r, max_num_object=cap)
                assert_states_identical(hs_g, hs_o, f"cap {cap} it {it}")
    # parent pre-truncation: a state larger than the cap
    big = hs_o
    if big.num_object > 2:
        k = big.num_object // 2
        g, _, _ = gpu_step(big, rules[1], max_num_object=k)
        o, _ = orc.simulate(big, rules[1], max_num_object=k)
        assert_states_identical(g, o, "pre-truncation")


@pytest.mark.parametrize("scale", [1.0, 3e-153, 1e-154, 1e-160])
def test_truncation_key_ranges(ctx, scale):
    """Top-K select whose threshold lies among normal |m|^2 (the select
    reads the dedup's 32-bit keys), around the smallest normal and among
    subnormal |m|^2, some of which underflow to 0 (tolerance 0; a threshold
    with a zero exponent field takes the full-record select): identical to
    the oracle.  About 530 uniques from 52 parents."""
    base = qcgd_initial(7, 11)
    rules = qcgd_iteration_rules()
    for r in rules + rules[:2]:
        base = orc.apply_modifier(base, r) if r.is_modifier else orc.simulate(base, r)[0]
    hs = HostState(base.objects, base.object_begin, base.object_size, base.magnitude * scale, base.total_proba)
    rule = qcgd.erase_create(math.pi / 3, math.pi / 5)
    for cap in (5, 37, 200):
        g, st, _ = gpu_step(hs, rule, max_num_object=cap, tolerance=0.0)
        o, ost = orc.simulate(hs, rule, max_num_object=cap, tolerance=0.0)
        assert st["num_unique"] == ost["num_unique"]
        assert_states_identical(g, o, f"scale {scale} cap {cap}")


def test_populate_large_objects(ctx):
    """Objects beyond the populate kernels' LDS buffers and node lists
    (parents of 1.2-1.6 KB, children of up to 74 nodes, graphs of more than
    64 nodes) mixed with small ones in one state, so that survivors written
    by the one-lane path and by the segment path share wavefronts;
    split_merge and coin steps identical to the oracle.  (The sparse graphs
    have few split / merge / coin operations; erase_create would give them
    2^n children.)"""
    objs = [sparse_graph(70, (3, 11, 20, 33, 47, 58, 66), (7, 25, 40)),
            sparse_graph(60, (2, 9, 30, 44, 51), (14, 36)),
            qcgd.seeded_graph(12, 8), qcgd.seeded_graph(9, 2)]
    rng = np.random.default_rng(17)
    mags = rng.normal(size=len(objs)) + 1j * rng.normal(size=len(objs))
    hs_g = hs_o = HostState.from_objects(objs, mags=mags / np.linalg.norm(mags))
    rules = [qcgd.split_merge(0.6, 0.3), qcgd.coin(0.9, 0.4), qcgd.split_merge(1.1, 0.2)]
    for k, r in enumerate(rules):
        hs_g, _, _ = gpu_step(hs_g, r, max_num_object=600)
        hs_o, _ = orc.simulate(hs_o, r, max_num_object=600)
        assert_states_identical(hs_g, hs_o, f"step {k} {type(r).__name__}")
    sizes = hs_o.object_size
    assert int(sizes.max()) > 1024 and int(sizes.min()) < 1024 and hs_o.num_object > 100


def test_coin_and_tolerance(ctx):
    hs = qcgd_initial(7, 5)
    r1 = qcgd.coin(math.pi / 7, 0.3, 0.2)
    g, _, _ = gpu_step(hs, r1, tolerance=1e-15)
    o, _ = orc.simulate(hs, r1, tolerance=1e-15)
    assert_states_identical(g, o, "coin")
    r2 = qcgd.split_merge(0.25, 0.25, -0.25)
    g2, _, _ = gpu_step(g, r2, tolerance=1e-15)
    o2, _ = orc.simulate(o, r2, tolerance=1e-15)
    assert_states_identical(g2, o2, "split_merge xi")


@pytest.mark.parametrize("first_pass_bits", [6, 8])
def test_larger_step_two_level_partition(ctx, first_pass_bits):
    """Small dedup buckets, so that steps of a few hundred thousand
    candidates take two partition passes; with 8 first-pass bits the second
    pass has more first-level buckets than workgroups (several per
    workgroup)."""
    old = ctx.get_tuning()
    ctx.set_tuning(dedup_bucket_target=64, first_pass_bits=first_pass_bits)
    try:
        _two_level_chain()
    finally:
        ctx.set_tuning(dedup_bucket_target=old["dedup_bucket_target"], first_pass_bits=old["first_pass_bits"])


def _two_level_chain():
    hs = qcgd_initial(8, 4)
    rules = qcgd_iteration_rules()
    hs_g = hs_o = hs
    done = False
    for it in range(4):
        for r in rules:
            if r.is_modifier:
                hs_g = gpu_modifier(hs_g, r)
                hs_o = orc.apply_modifier(hs_o, r)
                continue
            hs_g, st, _ = gpu_step(hs_g, r, max_num_object=1 << 16)
            hs_o, _ = orc.simulate(hs_o, r, max_num_object=1 << 16)
            assert_states_identical(hs_g, hs_o, f"it {it}")
            if st["dedup_levels"] == 2:
                done = True
                break
        if done:
            break
    assert done, "workload never reached two partition levels"


def test_empty_state(ctx):
    hs = HostState()
    out, st, _ = gpu_step(hs, qcgd.split_merge(1.0))
    assert out.num_object == 0 and st["num_candidates"] == 0 and out.total_proba == 0


@pytest.mark.parametrize("generic", [0, 1])
def test_qcgd_generate_paths_agree(ctx, generic):
    """The QCGD fast generate (work items sharing walk prefixes) and the
    generic byte-level generate both match the oracle."""
    old = ctx.get_tuning()
    ctx.set_tuning(generic_path=generic)
    try:
        hs_g = hs_o = qcgd_initial(7, 13)
        rules = qcgd_iteration_rules() + [qcgd.coin(0.7, 0.4, 0.1)]
        for it in range(3):
            for r in rules:
                if r.is_modifier:
                    hs_g = gpu_modifier(hs_g, r)
                    hs_o = orc.apply_modifier(hs_o, r)
                else:
                    hs_g, st, _ = gpu_step(hs_g, r, max_num_object=20000)
                    hs_o, _ = orc.simulate(hs_o, r, max_num_object=20000)
                    assert_states_identical(hs_g, hs_o, f"generic={generic} it {it} {type(r).__name__}")
    finally:
        ctx.set_tuning(generic_path=old["generic_path"])
