"""quids_amd -- the QuIDS rule-application step on MI355X (gfx950), Python host side.

Mirrors the reference's API (src/quids.hpp): the globals ``tolerance``,
``align_byte_length`` and ``simple_truncation`` (59-75), the state ``it_t``
(149-335), the symbolic iteration ``sy_it_t`` (338-429) and
``simulate(it, rule, next, sy_it, max_num_object)`` / ``simulate(it, modifier)``
(436-543).  Rules live in ``quids_amd.rules.quantum_computer`` and
``quids_amd.rules.qcgd``.  Every step runs in the HIP library
(libquids_hip.so) through its C ABI; there is no host fallback.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import QdModifier, QdParams, QdRule, QdStepStats, QuidsError, UINT64_MAX, check, lib
from .state import HostState, alignment_offset

__all__ = ["tolerance", "align_byte_length", "simple_truncation", "it_t", "sy_it_t", "simulate", "Context",
           "HostState", "QuidsError", "default_context", "UINT64_MAX"]

# reference globals (src/quids.hpp:59-75)
tolerance = 1e-30
align_byte_length = 8
simple_truncation = True
#: key of the probabilistic truncation's draws (simple_truncation False):
#: Philox2x64-10 keyed by (object hash, truncation_seed)
truncation_seed = 0
# memory-driven truncation (max_num_object == 0): the memory safety margin
# and the equalization factor of the next state's memory (src/quids.hpp:64-66)
safety_margin = 0.2
equalize_factor = 0.25
# the CPU load-balancing granularity (src/quids.hpp:68): read by the QCGD
# command-line flags, no effect on the device step
load_balancing_bucket_per_thread = 32


def _params():
    return QdParams(float(tolerance), int(align_byte_length), 1 if simple_truncation else 0, 0,
                    int(truncation_seed), float(safety_margin), float(equalize_factor))


class Context:
    """A device context: one HIP stream on one GPU and the step's scratch pool."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().qd_ctx_create(ctypes.byref(h), int(device)), "qd_ctx_create")
        self.handle = h
        self.device = device

    def synchronize(self):
        check(lib().qd_ctx_synchronize(self.handle), "qd_ctx_synchronize")

    def scratch_bytes(self):
        v = ctypes.c_uint64()
        check(lib().qd_ctx_scratch_bytes(self.handle, ctypes.byref(v)), "qd_ctx_scratch_bytes")
        return v.value

    def trim(self):
        check(lib().qd_ctx_trim(self.handle), "qd_ctx_trim")

    def get_tuning(self):
        t = _abi.QdTuning()
        check(lib().qd_ctx_get_tuning(self.handle, ctypes.byref(t)), "qd_ctx_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        """Sets launch / layout parameters (fields of qd_tuning); others keep their value."""
        t = _abi.QdTuning()
        check(lib().qd_ctx_get_tuning(self.handle, ctypes.byref(t)), "qd_ctx_get_tuning")
        for k, v in kw.items():
            if not hasattr(t, k) or k == "reserved":
                raise KeyError(k)
            setattr(t, k, int(v))
        check(lib().qd_ctx_set_tuning(self.handle, ctypes.byref(t)), "qd_ctx_set_tuning")

    def copy_probe(self, nbytes, reps=3):
        """(read + write) GB/s of a plain device copy of nbytes."""
        g = ctypes.c_double()
        check(lib().qd_copy_probe(self.handle, int(nbytes), int(reps), ctypes.byref(g)), "qd_copy_probe")
        return g.value

    def set_profiling(self, enable=True):
        """Per-phase GPU timing of simulate() with HIP events."""
        check(lib().qd_ctx_set_profiling(self.handle, 1 if enable else 0), "qd_ctx_set_profiling")

    def reset_phase_times(self):
        check(lib().qd_ctx_reset_phase_times(self.handle), "qd_ctx_reset_phase_times")

    def phase_times(self):
        """{phase: (milliseconds, intervals)} accumulated since the last reset."""
        n = len(_abi.PHASES)
        ms = (ctypes.c_double * n)()
        calls = (ctypes.c_uint64 * n)()
        check(lib().qd_ctx_phase_times(self.handle, ms, calls), "qd_ctx_phase_times")
        return {name: (ms[i], calls[i]) for i, name in enumerate(_abi.PHASES)}

    def close(self):
        if self.handle:
            lib().qd_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class it_t:
    """A wave function resident in device memory (reference: quids::iteration).

    Host-side edits (``append``, ``pop``) go to a host copy that is uploaded
    before the next device call; reads download lazily.
    """

    def __init__(self, obj=None, ctx=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().qd_state_create(self.ctx.handle, ctypes.byref(h)), "qd_state_create")
        self.handle = h
        self._host = HostState()
        self._host_valid = True   # host copy matches the device
        self._dev_valid = True    # device matches the host copy
        if obj is not None:
            self.append(obj)

    # -- host <-> device ----------------------------------------------------
    def upload(self, host_state):
        """Replaces the state with a HostState."""
        hs = host_state
        n = hs.num_object
        mag = hs.mag_pairs()
        check(lib().qd_state_upload(
            self.handle, n, hs.objects.ctypes.data_as(ctypes.c_void_p),
            hs.object_begin.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            hs.object_size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            mag.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(hs.total_proba)), "qd_state_upload")
        self._host = HostState(hs.objects.copy(), hs.object_begin.copy(), hs.object_size.copy(),
                               hs.magnitude.copy(), hs.total_proba)
        self._host_valid = True
        self._dev_valid = True
        return self

    def _sync_to_device(self):
        if not self._dev_valid:
            self.upload(self._host)

    def info(self):
        self._sync_to_device()
        n, b, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        check(lib().qd_state_info(self.handle, ctypes.byref(n), ctypes.byref(b), ctypes.byref(t)), "qd_state_info")
        return n.value, b.value, t.value

    def download(self):
        """The state as a HostState (copied from the device)."""
        if self._host_valid:
            return self._host
        n, nbytes, total = self.info()
        objects = np.zeros(nbytes, np.uint8)
        begin = np.zeros(n + 1, np.uint64)
        size = np.zeros(n, np.uint32)
        mag = np.zeros((n, 2), np.float64)
        check(lib().qd_state_download(
            self.handle, objects.ctypes.data_as(ctypes.c_void_p),
            begin.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            mag.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "qd_state_download")
        self._host = HostState(objects, begin, size, mag.view(np.complex128).reshape(-1), total)
        self._host_valid = True
        return self._host

    def _device_changed(self):
        self._host_valid = False
        self._dev_valid = True

    def copy(self):
        """A device-to-device copy of this state (qd_state_copy)."""
        self._sync_to_device()
        c = it_t(ctx=self.ctx)
        check(lib().qd_state_copy(c.handle, self.handle), "qd_state_copy")
        c._device_changed()
        return c

    # -- reference API ------------------------------------------------------
    @property
    def num_object(self):
        if self._host_valid:
            return self._host.num_object
        return self.info()[0]

    @property
    def total_proba(self):
        if self._host_valid:
            return self._host.total_proba
        return self.info()[2]

    def append(self, obj, mag=1.0):
        """iteration::append, src/quids.hpp:174-188."""
        hs = self.download()
        hs.append(bytes(obj), mag, int(align_byte_length))
        self._host = hs
        self._host_valid = True
        self._dev_valid = False

    def pop(self, n=1, normalize=True):
        """iteration::pop, src/quids.hpp:194-203."""
        if n < 1:
            return
        hs = self.download()
        k = hs.num_object - n
        st = HostState(hs.objects[:int(hs.object_begin[k])].copy(), hs.object_begin[:k + 1].copy(),
                       hs.object_size[:k].copy(), hs.magnitude[:k].copy(), hs.total_proba)
        self.upload(st)
        if normalize:
            self.normalize()

    def normalize(self):
        """iteration::normalize, src/quids.hpp:985-1017."""
        self._sync_to_device()
        t = ctypes.c_double()
        check(lib().qd_normalize(self.ctx.handle, self.handle, ctypes.byref(t)), "qd_normalize")
        self._device_changed()
        return t.value

    def get_object(self, i):
        """iteration::get_object (read-only), src/quids.hpp:254-258: (bytes, mag)."""
        return self.download().get_object(i)

    def average_value(self, observable):
        """iteration::average_value, src/quids.hpp:208-234, for a host callable
        observable(object_bytes) -> float."""
        hs = self.download()
        tot = 0.0
        for i in range(hs.num_object):
            o, m = hs.get_object(i)
            tot += observable(o) * (m.real * m.real + m.imag * m.imag)
        return tot

    def close(self):
        if getattr(self, "handle", None):
            lib().qd_state_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class sy_it_t:
    """Symbolic iteration (reference: quids::symbolic_iteration): the counters
    of the last step; the step's scratch lives in the context."""

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().qd_symbolic_create(self.ctx.handle, ctypes.byref(h)), "qd_symbolic_create")
        self.handle = h
        self.last_stats = None

    def _info(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().qd_symbolic_info(self.handle, ctypes.byref(a), ctypes.byref(b)), "qd_symbolic_info")
        return a.value, b.value

    @property
    def num_object(self):
        return self._info()[0]

    @property
    def num_object_after_interferences(self):
        return self._info()[1]

    def close(self):
        if getattr(self, "handle", None):
            lib().qd_symbolic_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class user_rule:
    """A rule compiled as a device functor by the caller (the reference's
    quids::rule subclass, src/quids.hpp:105-146): ``table`` is the
    qd_user_rule (QdUserRule) that the caller's compiled code filled with
    the launch callbacks of its kernels (quids_amd/user_rule.hpp,
    make_user_rule).  ``keepalive`` holds whatever owns the table's functor
    (its library)."""

    def __init__(self, table, keepalive=None):
        if not isinstance(table, _abi.QdUserRule):
            raise TypeError("user_rule needs a QdUserRule table")
        self.table = table
        self.keepalive = keepalive


def simulate(it, rule, next_iteration=None, symbolic_iteration=None, max_num_object=0, mid_step_function=None):
    """quids::simulate.

    ``simulate(it, modifier)`` applies a modifier in place (src/quids.hpp:436-438).
    ``simulate(it, rule, next, sy_it, max_num_object)`` runs one step
    (src/quids.hpp:448-543): ``max_num_object`` 0 means memory-driven
    truncation, ``UINT64_MAX`` (or -1) means no limit.  Returns the step's
    counters as a dict.
    """
    if getattr(rule, "is_modifier", False):
        it._sync_to_device()
        m = rule.desc()
        check(lib().qd_apply_modifier(it.ctx.handle, it.handle, ctypes.byref(m)), "qd_apply_modifier")
        it._device_changed()
        return None
    if next_iteration is None:
        raise TypeError("simulate(it, rule, next_iteration, ...) needs a next iteration")
    user = isinstance(rule, user_rule)
    if not user and not hasattr(rule, "desc"):
        raise QuidsError(_abi.QD_ERR_UNSUPPORTED, "simulate", "rule has no device implementation")
    if max_num_object is None or max_num_object < 0:
        max_num_object = UINT64_MAX
    it._sync_to_device()
    r = rule.table if user else rule.desc()
    prm = _params()
    st = QdStepStats()
    sy = symbolic_iteration.handle if symbolic_iteration is not None else None
    hook = None
    if mid_step_function is not None:  # called by the library between its phases, with the reference's names
        hook = _abi.PHASE_HOOK(lambda name, _user: mid_step_function(name.decode()))
        check(lib().qd_ctx_set_phase_hook(it.ctx.handle, hook, None), "qd_ctx_set_phase_hook")
    try:
        fn, name = (lib().qd_simulate_user, "qd_simulate_user") if user else (lib().qd_simulate, "qd_simulate")
        check(fn(it.ctx.handle, it.handle, ctypes.byref(r), next_iteration.handle, sy, int(max_num_object),
                 ctypes.byref(prm), ctypes.byref(st)), name)
    finally:
        if hook is not None:
            lib().qd_ctx_set_phase_hook(it.ctx.handle, _abi.PHASE_HOOK(), None)  # a null hook
    next_iteration._device_changed()
    d = st.as_dict()
    if symbolic_iteration is not None:
        symbolic_iteration.last_stats = d
    return d
