"""ctypes binding of the C ABI declared in include/quids_hip.h.

The library is loaded from the package directory (built in-tree by
quids_amd/csrc/Makefile).  There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libquids_hip.so")

QD_OK = 0
QD_ERR_INVALID = 1
QD_ERR_OOM = 2
QD_ERR_HIP = 3
QD_ERR_CAPACITY = 4
QD_ERR_UNSUPPORTED = 5

QD_RULE_HADAMARD = 1
QD_RULE_QCGD_ERASE_CREATE = 2
QD_RULE_QCGD_COIN = 3
QD_RULE_QCGD_SPLIT_MERGE = 4

QD_MOD_CNOT = 1
QD_MOD_XGATE = 2
QD_MOD_YGATE = 3
QD_MOD_ZGATE = 4
QD_MOD_QCGD_STEP = 5
QD_MOD_QCGD_REVERSED_STEP = 6

UINT64_MAX = (1 << 64) - 1

PHASES = ("prep", "scan", "generate", "partition", "dedup", "select", "order", "finalize", "normalize")


class QdRule(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("bit", ctypes.c_uint32),
        ("coef", ctypes.c_double),
        ("do_re", ctypes.c_double),
        ("do_im", ctypes.c_double),
        ("do_not_re", ctypes.c_double),
        ("do_not_im", ctypes.c_double),
        ("do_conj_re", ctypes.c_double),
        ("do_conj_im", ctypes.c_double),
        ("do_not_conj_re", ctypes.c_double),
        ("do_not_conj_im", ctypes.c_double),
    ]


class QdModifier(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("control_bit", ctypes.c_uint32),
        ("bit", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class QdParams(ctypes.Structure):
    _fields_ = [
        ("tolerance", ctypes.c_double),
        ("align_byte_length", ctypes.c_uint32),
        ("simple_truncation", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("safety_margin", ctypes.c_double),
        ("equalize_factor", ctypes.c_double),
    ]

    def __init__(self, tolerance=1e-30, align_byte_length=8, simple_truncation=1, reserved=0, seed=0,
                 safety_margin=0.2, equalize_factor=0.25):
        super().__init__(tolerance, align_byte_length, simple_truncation, reserved, seed, safety_margin,
                         equalize_factor)


class QdStepStats(ctypes.Structure):
    _fields_ = [
        ("num_parent", ctypes.c_uint64),
        ("num_parent_kept", ctypes.c_uint64),
        ("num_candidates", ctypes.c_uint64),
        ("num_unique", ctypes.c_uint64),
        ("num_survivors", ctypes.c_uint64),
        ("output_bytes", ctypes.c_uint64),
        ("total_proba", ctypes.c_double),
        ("dedup_levels", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("mem_target_parents", ctypes.c_uint64),
        ("mem_target_survivors", ctypes.c_uint64),
    ]

    def as_dict(self):
        """Counters by field name; 'reserved' holds flags (bit 0: the second
        partition pass fell back to exact offsets; bit 1: a batched step ran
        in several chunks; bit 2: hash-range passes; bit 3: the fused first
        radix pass was redone unfused); the mem_target fields are the
        memory-driven truncation's targets (0 with an explicit cap)."""
        return {name: getattr(self, name) for name, _ in self._fields_}


class QdTuning(ctypes.Structure):
    _fields_ = [
        ("dedup_bucket_target", ctypes.c_uint32),
        ("first_pass_bits", ctypes.c_uint32),
        ("generate_block", ctypes.c_uint32),
        ("generate_grid", ctypes.c_uint32),
        ("dedup_cap", ctypes.c_uint32),
        ("generic_path", ctypes.c_uint32),
        ("chunk_log2", ctypes.c_uint32),
        ("split_generate", ctypes.c_uint32),
        ("eager_select", ctypes.c_uint32),
        ("free_mem_mib", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


PHASE_HOOK = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_pu32 = ctypes.POINTER(ctypes.c_uint32)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_pd = ctypes.POINTER(ctypes.c_double)
_pvp = ctypes.POINTER(ctypes.c_void_p)

class QdObjectsView(ctypes.Structure):
    _fields_ = [("objects", ctypes.c_void_p), ("begin", ctypes.c_void_p), ("size", ctypes.c_void_p),
                ("mag", ctypes.c_void_p), ("n", ctypes.c_uint64)]


class QdUserRule(ctypes.Structure):
    """qd_user_rule: the launch callbacks of a rule compiled as a device
    functor (filled by the caller's compiled code, quids_amd/user_rule.hpp)."""
    _fields_ = [("self", ctypes.c_void_p), ("num_child", ctypes.c_void_p), ("generate", ctypes.c_void_p),
                ("populate", ctypes.c_void_p), ("hash_objects", ctypes.c_void_p)]


# name -> (argtypes); every entry point returns int except the three noted.
SIGNATURES = {
    "qd_abi_version": [],
    "qd_status_string": [ctypes.c_int],
    "qd_last_error": [],
    "qd_ctx_create": [_pvp, ctypes.c_int],
    "qd_ctx_destroy": [_vp],
    "qd_ctx_synchronize": [_vp],
    "qd_ctx_scratch_bytes": [_vp, _pu64],
    "qd_ctx_trim": [_vp],
    "qd_ctx_set_profiling": [_vp, ctypes.c_int],
    "qd_ctx_get_tuning": [_vp, ctypes.POINTER(QdTuning)],
    "qd_ctx_set_tuning": [_vp, ctypes.POINTER(QdTuning)],
    "qd_ctx_reset_phase_times": [_vp],
    "qd_ctx_phase_times": [_vp, _pd, _pu64],
    "qd_copy_probe": [_vp, _u64, ctypes.c_int, _pd],
    "qd_state_create": [_vp, _pvp],
    "qd_state_destroy": [_vp],
    "qd_state_upload": [_vp, _u64, _vp, _pu64, _pu32, _pd, ctypes.c_double],
    "qd_state_info": [_vp, _pu64, _pu64, _pd],
    "qd_state_download": [_vp, _vp, _pu64, _pu32, _pd],
    "qd_state_device_ptrs": [_vp, _pvp, _pvp, _pvp, _pvp],
    "qd_symbolic_create": [_vp, _pvp],
    "qd_symbolic_destroy": [_vp],
    "qd_symbolic_info": [_vp, _pu64, _pu64],
    "qd_simulate": [_vp, _vp, ctypes.POINTER(QdRule), _vp, _vp, _u64, ctypes.POINTER(QdParams),
                    ctypes.POINTER(QdStepStats)],
    "qd_apply_modifier": [_vp, _vp, ctypes.POINTER(QdModifier)],
    "qd_normalize": [_vp, _vp, _pd],
    "qd_rule_num_child": [ctypes.POINTER(QdRule), _vp, _u32, _pu32, _pu32],
    "qd_rule_populate_child": [ctypes.POINTER(QdRule), _vp, _u32, _vp, _u32, _u32, _pu32, _pd],
    "qd_rule_hash": [ctypes.POINTER(QdRule), _vp, _u32, _pu64],
    "qd_modifier_apply_host": [ctypes.POINTER(QdModifier), _vp, _u32, _pd],
    "qd_state_copy": [_vp, _vp],
    "qd_simulate_batched": [_vp, _vp, _pu64, _u32, ctypes.POINTER(QdRule), _vp, _pu64, _pu64, ctypes.POINTER(QdParams),
                            ctypes.POINTER(QdStepStats)],
    "qd_average_value": [_vp, _vp, ctypes.c_int, ctypes.c_int, _pd, _vp],
    "qd_average_value_noalloc": [_vp, _vp, ctypes.c_int, ctypes.c_int, _pd, _vp],
    "qd_ctx_set_phase_hook": [_vp, PHASE_HOOK, _vp],
    "qd_simulate_adjoint": [_vp, _vp, ctypes.POINTER(QdRule), ctypes.POINTER(QdRule), ctypes.c_int, _vp, _vp, _vp,
                            _pd, _u64],
    "qd_modifier_adjoint": [_vp, _vp, ctypes.POINTER(QdModifier), _vp],
    "qd_ctx_stream": [_vp, _pvp],
    "qd_simulate_user": [_vp, _vp, ctypes.POINTER(QdUserRule), _vp, _vp, _u64, ctypes.POINTER(QdParams),
                         ctypes.POINTER(QdStepStats)],
    "qd_average_value_of": [_vp, _vp, _vp, _pd, _vp],
}
RESTYPES = {"qd_status_string": ctypes.c_char_p, "qd_last_error": ctypes.c_char_p}


class QuidsError(RuntimeError):
    def __init__(self, code, where, message):
        super().__init__(f"{where} failed with status {code}: {message}")
        self.code = code


_lib = None


def lib():
    """The loaded C ABI library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"quids_amd: HIP library {LIB_PATH} is missing; build it with "
                "`make -C quids_amd/csrc` (or __graft_entry__.build())")
        # One HIP runtime per process: torch's HIP library asks for the
        # runtime by its unversioned name, so if this library (linked to
        # libamdhip64.so.7) loaded first, torch would load a second runtime
        # that sees no device.  Loading torch first makes this library reuse
        # torch's runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = RESTYPES.get(name, ctypes.c_int)
        _lib = handle
    return _lib


def check(code, where):
    if code != QD_OK:
        msg = lib().qd_last_error()
        raise QuidsError(code, where, msg.decode() if msg else "")
