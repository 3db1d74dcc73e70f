"""Quantum causal graph dynamics rules (reference: src/rules/qcgd.hpp).

Graph object layout (qcgd.hpp:28-112): u16 n | bool left[n] | bool right[n]
| u16 name_begin[n+1] | sub_node names[m], sub_node = {i16
hmlz_and_element, i16 right_or_type, 4 padding bytes, u64 hash}.
"""
import ctypes
import ctypes.util
import math
import struct

import numpy as np

from .._abi import (QD_MOD_QCGD_REVERSED_STEP, QD_MOD_QCGD_STEP, QD_RULE_QCGD_COIN, QD_RULE_QCGD_ERASE_CREATE,
                    QD_RULE_QCGD_SPLIT_MERGE, QdModifier, QdRule)
from ._base import DeviceModifier, DeviceRule

DOT_L, DOT_R, ELEMENT = -3, -2, -1


def _polar(rho, theta):
    """std::polar(rho, theta) = (rho cos theta, rho sin theta)."""
    return complex(rho * math.cos(theta), rho * math.sin(theta))


class _CoefRule(DeviceRule):
    """Coefficients do_ = polar(sin theta, phi), do_not = polar(cos theta, xi)
    and their conjugates (qcgd.hpp:466-471, 541-546, 614-619)."""

    kind = 0

    def __init__(self, theta, phi=0.0, xi=0.0):
        self.theta, self.phi, self.xi = float(theta), float(phi), float(xi)
        self.do_ = _polar(math.sin(self.theta), self.phi)
        self.do_not = _polar(math.cos(self.theta), self.xi)
        self.do_conj = self.do_.conjugate()
        self.do_not_conj = self.do_not.conjugate()

    def desc(self):
        return QdRule(kind=self.kind, bit=0, coef=0.0,
                      do_re=self.do_.real, do_im=self.do_.imag,
                      do_not_re=self.do_not.real, do_not_im=self.do_not.imag,
                      do_conj_re=self.do_conj.real, do_conj_im=self.do_conj.imag,
                      do_not_conj_re=self.do_not_conj.real, do_not_conj_im=self.do_not_conj.imag)


    #: names of the real parameters, in the order of tangent_descs
    params = ("theta", "phi", "xi")

    def tangent_descs(self, ndir=2):
        """Rule descriptors whose coefficient fields hold the derivatives of
        this rule's coefficients along theta, phi (and xi when ndir == 3):
        d do_/d theta = polar(cos theta, phi), d do_/d phi = i do_,
        d do_not/d theta = polar(-sin theta, xi), d do_not/d xi = i do_not,
        and the conjugates of these for do_conj / do_not_conj."""
        th, ph, xi = self.theta, self.phi, self.xi
        d_do = [_polar(math.cos(th), ph), 1j * self.do_, 0j]
        d_dn = [_polar(-math.sin(th), xi), 0j, 1j * self.do_not]
        out = []
        for t in range(ndir):
            a, b = complex(d_do[t]), complex(d_dn[t])
            out.append(QdRule(kind=self.kind, bit=0, coef=0.0, do_re=a.real, do_im=a.imag,
                              do_not_re=b.real, do_not_im=b.imag, do_conj_re=a.real, do_conj_im=-a.imag,
                              do_not_conj_re=b.real, do_not_conj_im=-b.imag))
        return out


class erase_create(_CoefRule):
    """qcgd.hpp:459-532: at every node with left == right, either flip both
    particle bits (factor do_, or do_conj when both were set) or not (do_not,
    or -do_not_conj)."""

    kind = QD_RULE_QCGD_ERASE_CREATE


class coin(_CoefRule):
    """qcgd.hpp:534-605: as erase_create at nodes with left != right."""

    kind = QD_RULE_QCGD_COIN


class split_merge(_CoefRule):
    """qcgd.hpp:607-1036: split nodes holding both particles, merge a
    left-moving particle with the next node's right-moving one."""

    kind = QD_RULE_QCGD_SPLIT_MERGE


class _StepModifier(DeviceModifier):
    def __init__(self, kind):
        self.kind = kind

    def desc(self):
        return QdModifier(kind=self.kind)


#: qcgd.hpp:443-449: left particles move to the previous node, right ones to the next.
step = _StepModifier(QD_MOD_QCGD_STEP)
#: qcgd.hpp:451-457: inverse of step.
reversed_step = _StepModifier(QD_MOD_QCGD_REVERSED_STEP)


# ---------------------------------------------------------------------------
# graph utilities (qcgd.hpp:114-120, 214-240)
# ---------------------------------------------------------------------------
def num_nodes(obj):
    return struct.unpack_from("<H", obj, 0)[0]


def left(obj):
    n = num_nodes(obj)
    return np.frombuffer(bytes(obj[2:2 + n]), np.uint8).astype(bool)


def right(obj):
    n = num_nodes(obj)
    return np.frombuffer(bytes(obj[2 + n:2 + 2 * n]), np.uint8).astype(bool)


def make_graph(n):
    """utils::make_graph (qcgd.hpp:214-230): a cycle of n nodes named 0..n-1,
    no particles.  Padding bytes are zero."""
    size = 4 + 20 * n
    b = bytearray(size)
    struct.pack_into("<H", b, 0, n)
    nb = 2 + 2 * n
    for i in range(n + 1):
        struct.pack_into("<H", b, nb + 2 * i, i)
    subs = 4 + 4 * n
    for i in range(n):
        hmlz = -1 if i == 0 else i + 1
        struct.pack_into("<hh4xQ", b, subs + 16 * i, hmlz, ELEMENT, i)
    return bytes(b)


def with_particles(obj, left_bits, right_bits):
    """Copy of a graph with its particle bits replaced."""
    b = bytearray(obj)
    n = num_nodes(obj)
    for i in range(n):
        b[2 + i] = 1 if left_bits[i] else 0
        b[2 + n + i] = 1 if right_bits[i] else 0
    return bytes(b)


_libc = None


def _c_rand():
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL(ctypes.util.find_library("c"))
        _libc.rand.restype = ctypes.c_int
        _libc.srand.argtypes = [ctypes.c_uint]
    return _libc


def seeded_graph(n, seed):
    """flags::read_n_iter("...,seed=s") then read_state("n") (qcgd.hpp:1086-1137):
    srand(seed), then graphs::randomize draws rand() & 1 for left[0],
    right[0], left[1], ... (qcgd.hpp:114-120) with the C library's rand()."""
    libc = _c_rand()
    libc.srand(int(seed))
    L, R = [], []
    for _ in range(n):
        L.append(libc.rand() & 1)
        R.append(libc.rand() & 1)
    return with_particles(make_graph(n), L, R)


def particle_count(obj):
    """Number of particles sum(left + right) of a graph."""
    n = num_nodes(obj)
    return int(np.count_nonzero(np.frombuffer(bytes(obj[2:2 + 2 * n]), np.uint8)))


def decode_fields(obj):
    """(n, left, right, name_begin, [(hmlz, type, hash), ...]) of a graph,
    padding bytes ignored: the field-wise identity used by parity checks."""
    n = num_nodes(obj)
    L = tuple(bool(x) for x in obj[2:2 + n])
    R = tuple(bool(x) for x in obj[2 + n:2 + 2 * n])
    nb = struct.unpack_from("<%dH" % (n + 1), obj, 2 + 2 * n)
    subs = 4 + 4 * n
    m = nb[-1]
    names = tuple(struct.unpack_from("<hh4xQ", obj, subs + 16 * k) for k in range(m))
    return (n, L, R, tuple(nb), names)
