"""Python access to the CPU checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (quids_amd/).

* ``liboracle.so`` -- the C++ restatement of the reference step
  (oracle/oracle.cpp), loaded with ctypes.
* ``ref_driver`` -- the reference headers compiled in place
  (oracle/_ref/ref_driver, built by oracle/Makefile), driven through
  program files (format in oracle/ref_driver.cpp).
"""
import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_lib = None
MAX_DIRECTIONS = 8  # tangent directions of one forward-mode step (kDir in oracle.cpp)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library {LIB_PATH} missing: run `make -C oracle`")
        from quids_amd._abi import QdModifier, QdParams, QdRule, QdStepStats
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32
        pu64, pu32, pd = ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_double)
        L.or_state_new.argtypes = [u64, vp, pu64, pu32, pd, ctypes.c_double]
        L.or_state_new.restype = vp
        L.or_state_free.argtypes = [vp]
        L.or_state_info.argtypes = [vp, pu64, pu64, pd]
        L.or_state_get.argtypes = [vp, vp, pu64, pu32, pd]
        L.or_simulate.argtypes = [vp, ctypes.POINTER(QdRule), ctypes.POINTER(vp), u64, ctypes.POINTER(QdParams),
                                  ctypes.POINTER(QdStepStats)]
        L.or_modifier.argtypes = [vp, ctypes.POINTER(QdModifier)]
        L.or_num_child.argtypes = [ctypes.POINTER(QdRule), vp, u32, pu32, pu32]
        L.or_populate.argtypes = [ctypes.POINTER(QdRule), vp, u32, u32, vp, pu32, pd]
        L.or_hash.argtypes = [ctypes.POINTER(QdRule), vp, u32]
        L.or_hash.restype = u64
        L.or_std_hash.argtypes = [vp, u64]
        L.or_std_hash.restype = u64
        L.or_hash_combine.argtypes = [u64, u64]
        L.or_hash_combine.restype = u64
        L.or_canonical_sum.argtypes = [pd, u64]
        L.or_canonical_sum.restype = ctypes.c_double
        L.or_average_value.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_average_value.restype = ctypes.c_double
        L.or_state_set_tangent.argtypes = [vp, ctypes.c_int, pd]
        L.or_state_set_tangent.restype = ctypes.c_int
        L.or_state_get_tangent.argtypes = [vp, pd]
        L.or_simulate_jvp.argtypes = [vp, ctypes.POINTER(QdRule), ctypes.POINTER(QdRule), ctypes.POINTER(vp), u64,
                                      ctypes.POINTER(QdParams), ctypes.POINTER(QdStepStats)]
        L.or_average_value_jvp.argtypes = [vp, ctypes.c_int, ctypes.c_int, pd]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_simulate_caps.argtypes = [vp, ctypes.POINTER(QdRule), ctypes.POINTER(vp), u64, u64,
                                       ctypes.POINTER(QdParams), ctypes.POINTER(QdStepStats), ctypes.c_int]
        _lib = L
    return _lib


def _to_or(hs):
    L = _load()
    mag = hs.mag_pairs()
    return L.or_state_new(hs.num_object, hs.objects.ctypes.data_as(ctypes.c_void_p),
                          hs.object_begin.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                          hs.object_size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                          mag.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), hs.total_proba)


def _from_or(h):
    from quids_amd.state import HostState
    L = _load()
    n, b, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
    L.or_state_info(h, ctypes.byref(n), ctypes.byref(b), ctypes.byref(t))
    objects = np.zeros(b.value, np.uint8)
    begin = np.zeros(n.value + 1, np.uint64)
    size = np.zeros(n.value, np.uint32)
    mag = np.zeros((n.value, 2), np.float64)
    L.or_state_get(h, objects.ctypes.data_as(ctypes.c_void_p), begin.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                   size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                   mag.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return HostState(objects, begin, size, mag.view(np.complex128).reshape(-1), t.value)


def simulate(hs, rule, max_num_object=(1 << 64) - 1, tolerance=1e-30, align=8, simple_truncation=True, seed=0):
    """One oracle step on a HostState -> (HostState, stats dict)."""
    from quids_amd._abi import QdParams, QdStepStats
    L = _load()
    h = _to_or(hs)
    out = ctypes.c_void_p()
    st = QdStepStats()
    prm = QdParams(tolerance, align, 1 if simple_truncation else 0, 0, seed)
    d = rule.desc()
    rc = L.or_simulate(h, ctypes.byref(d), ctypes.byref(out), int(max_num_object), ctypes.byref(prm),
                       ctypes.byref(st))
    L.or_state_free(h)
    if rc:
        raise RuntimeError(f"oracle simulate failed: {rc}")
    res = _from_or(out)
    L.or_state_free(out)
    return res, st.as_dict()


def simulate_caps(hs, rule, max_parents, max_survivors, tolerance=1e-30, align=8, normalize=True):
    """One oracle step keeping the top max_parents parents and the top
    max_survivors uniques (the counts a memory-driven search settles on)
    -> (HostState, stats dict).  normalize=False leaves the magnitudes as
    the interference summed them (the values the truncation orders)."""
    from quids_amd._abi import QdParams, QdStepStats
    L = _load()
    h = _to_or(hs)
    out = ctypes.c_void_p()
    st = QdStepStats()
    prm = QdParams(tolerance, align, 1)
    d = rule.desc()
    rc = L.or_simulate_caps(h, ctypes.byref(d), ctypes.byref(out), int(max_parents), int(max_survivors),
                            ctypes.byref(prm), ctypes.byref(st), 1 if normalize else 0)
    L.or_state_free(h)
    if rc:
        raise RuntimeError(f"oracle simulate failed: {rc}")
    res = _from_or(out)
    L.or_state_free(out)
    return res, st.as_dict()


def set_threads(n):
    """OpenMP threads of the calling host thread's later oracle calls."""
    _load().or_set_threads(int(n))


def apply_modifier(hs, mod):
    L = _load()
    h = _to_or(hs)
    d = mod.desc()
    rc = L.or_modifier(h, ctypes.byref(d))
    res = _from_or(h)
    L.or_state_free(h)
    if rc:
        raise RuntimeError(f"oracle modifier failed: {rc}")
    return res


def num_child(rule, obj):
    L = _load()
    nc, ms = ctypes.c_uint32(), ctypes.c_uint32()
    d = rule.desc()
    o = bytes(obj)
    L.or_num_child(ctypes.byref(d), o, len(o), ctypes.byref(nc), ctypes.byref(ms))
    return nc.value, ms.value


def populate(rule, obj, child_id, mag=1.0):
    L = _load()
    o = bytes(obj)
    _, ms = num_child(rule, o)
    buf = ctypes.create_string_buffer(ms + 64)
    sz = ctypes.c_uint32()
    m = (ctypes.c_double * 2)(complex(mag).real, complex(mag).imag)
    d = rule.desc()
    L.or_populate(ctypes.byref(d), o, len(o), int(child_id), buf, ctypes.byref(sz), m)
    return buf.raw[:sz.value], complex(m[0], m[1])


def rule_hash(rule, obj):
    d = rule.desc()
    o = bytes(obj)
    return _load().or_hash(ctypes.byref(d), o, len(o))


def std_hash(data):
    """std::hash<std::string_view> of the host C++ library."""
    b = bytes(data)
    return _load().or_std_hash(b, len(b))


def hash_combine(seed, v):
    return _load().or_hash_combine(seed & (2**64 - 1), v & (2**64 - 1))


def canonical_sum(values):
    a = np.ascontiguousarray(values, np.float64)
    return _load().or_canonical_sum(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0])


def average_value(hs, observable, arg=0):
    """0 = num_nodes, 1 = particle count, 2 = [num_nodes == arg]."""
    L = _load()
    h = _to_or(hs)
    v = L.or_average_value(h, int(observable), int(arg))
    L.or_state_free(h)
    return v


# ---------------------------------------------------------------------------
# reference driver
# ---------------------------------------------------------------------------
OP_RULE, OP_MOD = 1, 2


def rule_op(rule, max_num_object=(1 << 64) - 1, dump=True):
    """Program op for a rule object of quids_amd.rules."""
    from quids_amd._abi import QD_RULE_HADAMARD
    d = rule.desc()
    if d.kind == QD_RULE_HADAMARD:
        return (OP_RULE, d.kind, rule.bit, 0, 0.0, 0.0, 0.0, max_num_object, int(dump))
    return (OP_RULE, d.kind, 0, 0, rule.theta, rule.phi, rule.xi, max_num_object, int(dump))


def mod_op(mod, dump=False):
    d = mod.desc()
    return (OP_MOD, d.kind, d.control_bit, d.bit, 0.0, 0.0, 0.0, 0, int(dump))


def write_program(path, family, hs, ops, tolerance=1e-30, align=8):
    with open(path, "wb") as f:
        f.write(b"QDP1")
        f.write(struct.pack("<IIIdII", family, len(ops), align, tolerance, hs.num_object, 0))
        for i in range(hs.num_object):
            o, m = hs.get_object(i)
            f.write(struct.pack("<IIdd", len(o), 0, m.real, m.imag))
            f.write(o)
        for op in ops:
            f.write(struct.pack("<IIIIdddQII", *op, 0))


def read_dumps(path):
    """List of dicts {N, C, U, total_proba, seconds, objects, mags, hashes}."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        N, C, U, total, secs, nbytes = struct.unpack_from("<QQQddQ", data, pos)
        pos += 48
        objs, mags, hashes = [], [], []
        for _ in range(N):
            sz, _pad, re, im, h = struct.unpack_from("<IIddQ", data, pos)
            pos += 32
            objs.append(data[pos:pos + sz])
            pos += sz
            mags.append(complex(re, im))
            hashes.append(h)
        out.append(dict(N=N, C=C, U=U, total_proba=total, seconds=secs, objects=objs,
                        mags=np.array(mags, np.complex128), hashes=np.array(hashes, np.uint64)))
    return out


def run_reference(family, hs, ops, tolerance=1e-30, align=8, threads=None, timeout=600):
    """Runs the compiled reference on a program; returns (dumps, stdout)."""
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError(f"reference driver {REF_DRIVER} missing: run `make -C oracle ref`")
    with tempfile.TemporaryDirectory() as d:
        prog, out = os.path.join(d, "p.bin"), os.path.join(d, "o.bin")
        write_program(prog, family, hs, ops, tolerance, align)
        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        r = subprocess.run([REF_DRIVER, prog, out], capture_output=True, text=True, env=env, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"ref_driver failed ({r.returncode}): {r.stderr}")
        return read_dumps(out), r.stdout


def program_jvp(hs0, program, params, observable, arg=0, tolerance=1e-30, align=8):
    """Forward-mode derivative of a quids_amd.adjoint.Program's loss
    (average_value of `observable` after the program) with respect to every
    parameter: the value and tangents are carried together through each
    step (oracle simulate with dual magnitudes, the truncation chosen by
    value as in the plain step).  Returns (loss, gradient array, final
    HostState)."""
    from quids_amd._abi import QdParams, QdRule, QdStepStats
    from quids_amd.adjoint import RuleOp
    L = _load()
    params = [float(p) for p in params]
    D = len(params)
    h = _to_or(hs0)
    zeros = (ctypes.c_double * (2 * D * max(hs0.num_object, 1)))()
    L.or_state_set_tangent(h, D, zeros)
    prm = QdParams(tolerance, align, 1)
    try:
        for _ in range(program.n_iter):
            for op in program.ops:
                if isinstance(op, RuleOp):
                    rule = op.make(params)
                    tans = rule.tangent_descs(len(op.index))
                    dcoef = (QdRule * max(D, 1))()
                    for t in range(D):
                        dcoef[t] = tans[op.index.index(t)] if t in op.index else QdRule(kind=rule.kind)
                    out = ctypes.c_void_p()
                    st = QdStepStats()
                    d = rule.desc()
                    rc = L.or_simulate_jvp(h, ctypes.byref(d), dcoef, ctypes.byref(out), int(program.max_num_object),
                                           ctypes.byref(prm), ctypes.byref(st))
                    if rc:
                        raise RuntimeError(f"oracle simulate_jvp failed: {rc}")
                    L.or_state_free(h)
                    h = out.value
                else:
                    d = op.desc()
                    if L.or_modifier(h, ctypes.byref(d)):
                        raise RuntimeError("oracle modifier failed")
        value = L.or_average_value(h, int(observable), int(arg))
        dv = (ctypes.c_double * max(D, 1))()
        L.or_average_value_jvp(h, int(observable), int(arg), dv)
        final = _from_or(h)
    finally:
        L.or_state_free(h)
    return value, np.array([dv[i] for i in range(D)]), final


def program_loss(hs0, program, params, observable, arg=0, tolerance=1e-30, align=8):
    """The loss alone, through the plain oracle steps (finite differences)."""
    hs = hs0
    for _ in range(program.n_iter):
        for op in program.ops:
            if hasattr(op, "make"):
                hs, _ = simulate(hs, op.make(params), program.max_num_object, tolerance, align)
            else:
                hs = apply_modifier(hs, op)
    return average_value(hs, observable, arg)



def _set_tan(h, tan):
    L = _load()
    tan = np.ascontiguousarray(tan, np.complex128)
    D = tan.shape[1] if tan.ndim == 2 else 0
    flat = np.ascontiguousarray(tan.reshape(-1).view(np.float64)) if tan.size else np.zeros(1)
    if L.or_state_set_tangent(h, D, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
        raise ValueError(f"the oracle carries at most {MAX_DIRECTIONS} tangent directions, not {D}")
    return D


def _get_tan(h, n, D):
    out = np.zeros(max(n * D, 1) * 2, np.float64)
    _load().or_state_get_tangent(h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out[:2 * n * D].view(np.complex128).reshape(n, D)


def simulate_jvp(hs, tan, rule, dcoef=None, max_num_object=(1 << 64) - 1, tolerance=1e-30, align=8):
    """One oracle step carrying tangents: tan is (N, D) complex (derivatives
    of the input magnitudes along D directions), dcoef D rule descriptors
    (derivatives of the rule's coefficients; zero rules when None).
    Returns (HostState, (K, D) tangents of the normalised output)."""
    from quids_amd._abi import QdParams, QdRule, QdStepStats
    L = _load()
    h = _to_or(hs)
    D = _set_tan(h, tan)
    arr = (QdRule * max(D, 1))()
    for t in range(D):
        arr[t] = dcoef[t] if dcoef is not None else QdRule(kind=rule.desc().kind)
    out = ctypes.c_void_p()
    st = QdStepStats()
    prm = QdParams(tolerance, align, 1)
    d = rule.desc()
    rc = L.or_simulate_jvp(h, ctypes.byref(d), arr, ctypes.byref(out), int(max_num_object), ctypes.byref(prm),
                           ctypes.byref(st))
    L.or_state_free(h)
    if rc:
        raise RuntimeError(f"oracle simulate_jvp failed: {rc}")
    res = _from_or(out.value)
    t_out = _get_tan(out.value, res.num_object, D)
    L.or_state_free(out.value)
    return res, t_out


def modifier_jvp(hs, tan, mod):
    """A modifier applied to a state carrying tangents: (HostState, tangents)."""
    L = _load()
    h = _to_or(hs)
    D = _set_tan(h, tan)
    d = mod.desc()
    if L.or_modifier(h, ctypes.byref(d)):
        L.or_state_free(h)
        raise RuntimeError("oracle modifier failed")
    res = _from_or(h)
    t_out = _get_tan(h, res.num_object, D)
    L.or_state_free(h)
    return res, t_out
