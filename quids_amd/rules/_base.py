"""Host-side view of device rules: the reference's rule_t virtual methods
(src/quids.hpp:105-146) evaluated by the library's functors through the C ABI."""
import ctypes

import numpy as np

from .._abi import check, lib


class DeviceRule:
    """Base of rules with a device implementation; subclasses define desc()."""

    is_modifier = False

    def desc(self):  # pragma: no cover - abstract
        raise NotImplementedError

    def get_num_child(self, parent):
        """rule::get_num_child -> (num_child, max_child_size)."""
        p = bytes(parent)
        nc, ms = ctypes.c_uint32(), ctypes.c_uint32()
        d = self.desc()
        check(lib().qd_rule_num_child(ctypes.byref(d), p, len(p), ctypes.byref(nc), ctypes.byref(ms)),
              "qd_rule_num_child")
        return nc.value, ms.value

    def populate_child(self, parent, child_id, mag=1.0):
        """rule::populate_child -> (child bytes, child magnitude)."""
        p = bytes(parent)
        _, ms = self.get_num_child(p)
        buf = ctypes.create_string_buffer(max(ms, 1))
        size = ctypes.c_uint32()
        m = (ctypes.c_double * 2)(complex(mag).real, complex(mag).imag)
        d = self.desc()
        check(lib().qd_rule_populate_child(ctypes.byref(d), p, len(p), buf, max(ms, 1), int(child_id),
                                           ctypes.byref(size), m), "qd_rule_populate_child")
        return buf.raw[:size.value], complex(m[0], m[1])

    def populate_child_simple(self, parent, child_id):
        """rule::populate_child_simple -> child bytes."""
        return self.populate_child(parent, child_id)[0]

    def hasher(self, obj):
        """rule::hasher."""
        o = bytes(obj)
        h = ctypes.c_uint64()
        d = self.desc()
        check(lib().qd_rule_hash(ctypes.byref(d), o, len(o), ctypes.byref(h)), "qd_rule_hash")
        return h.value


class DeviceModifier:
    """Base of modifiers with a device implementation (quids::modifier_t)."""

    is_modifier = True

    def desc(self):  # pragma: no cover - abstract
        raise NotImplementedError

    def __call__(self, obj, mag=1.0):
        """modifier(begin, end, mag) on a host copy -> (new bytes, new mag)."""
        buf = ctypes.create_string_buffer(bytes(obj), len(obj))
        m = (ctypes.c_double * 2)(complex(mag).real, complex(mag).imag)
        d = self.desc()
        check(lib().qd_modifier_apply_host(ctypes.byref(d), buf, len(obj), m), "qd_modifier_apply_host")
        return buf.raw[:len(obj)], complex(m[0], m[1])
