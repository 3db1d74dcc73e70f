"""Quantum causal graph dynamics rules (reference: src/rules/qcgd.hpp).

Graph object layout (qcgd.hpp:28-112): u16 n | bool left[n] | bool right[n]
| u16 name_begin[n+1] | sub_node names[m], sub_node = {i16
hmlz_and_element, i16 right_or_type, 4 padding bytes, u64 hash}.
"""
import ctypes
import ctypes.util
import math
import struct
import threading

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


_rand_lock = threading.Lock()


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
    L, R = [], []
    with _rand_lock:  # srand / rand share one process-wide state
        libc.srand(int(seed))
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


# ---------------------------------------------------------------------------
# statistics and printing (qcgd.hpp:212, 232-353) and the command-line flags
# of the Quantum-graph-simulation runs (qcgd.hpp:1038-1176)
# ---------------------------------------------------------------------------
def _host_view(state):
    """The state's objects and magnitudes on the host (downloads an it_t)."""
    from ..state import HostState
    return state if isinstance(state, HostState) else state.download()


def _name(obj, k):
    """Name tree of the sub_node k (qcgd.hpp:243-266)."""
    n = num_nodes(obj)
    hmlz, typ = struct.unpack_from("<hh", obj, 4 + 4 * n + 16 * k)
    if typ == ELEMENT:
        return str(abs(hmlz) - 1)
    if typ in (DOT_L, DOT_R):
        return "(" + _name(obj, k + 1) + (").l" if typ == DOT_L else ").r")
    return "(" + _name(obj, k + 1) + ")∨(" + _name(obj, k + typ) + ")"


def _density(obj):
    n = num_nodes(obj)
    d = 0.0
    for i in range(n):
        d += obj[2 + i] + obj[2 + n + i]
    return d / (2 * n)


class utils:
    """qcgd::utils: graph construction, particle drawing, printing and
    statistics of a state (an it_t or a HostState)."""
    #: graphs printed by print() (qcgd.hpp:212)
    max_print_num_graphs = (1 << 64) - 1
    make_graph = staticmethod(make_graph)

    @staticmethod
    def randomize(state):
        """qcgd.hpp:232-240: draws the particles of every graph of the state,
        rand() & 1 for left[0], right[0], left[1], ... graph after graph."""
        from ..state import HostState
        libc = _c_rand()
        hs = _host_view(state)
        buf = bytearray(hs.objects.tobytes())
        for i in range(hs.num_object):
            b = int(hs.object_begin[i])
            n = struct.unpack_from("<H", buf, b)[0]
            for k in range(n):
                buf[b + 2 + k] = libc.rand() & 1
                buf[b + 2 + n + k] = libc.rand() & 1
        out = HostState(np.frombuffer(bytes(buf), np.uint8).copy(), hs.object_begin, hs.object_size,
                        hs.magnitude, hs.total_proba)
        if isinstance(state, HostState):
            state.objects = out.objects
        else:
            state.upload(out)
        return state

    @staticmethod
    def print(state, file=None):
        """qcgd.hpp:242-303: the max_print_num_graphs most probable graphs
        (stable order among equal |m|^2), one per line: the magnitude (parts
        below the tolerance shown as 0, 5 decimals) and every node as
        -|<|name|>|-.  Writes to `file` (stdout by default) and returns the
        text."""
        import sys
        import quids_amd
        hs = _host_view(state)
        mags = np.asarray(hs.magnitude)
        order = np.argsort(-(mags.real * mags.real + mags.imag * mags.imag), kind="stable")
        num = min(hs.num_object, utils.max_print_num_graphs)
        tol = quids_amd.tolerance
        lines = []
        for g in order[:num]:
            obj = hs.get_object(int(g))[0]
            m = complex(mags[g])
            re = 0.0 if abs(m.real) < tol else m.real
            im = 0.0 if abs(m.imag) < tol else m.imag
            s = "\t%.5f%s%.5fi  " % (re, " - " if im <= -1e-5 else " + ", abs(im))
            n = num_nodes(obj)
            for k in range(n):
                nb = struct.unpack_from("<H", obj, 2 + 2 * n + 2 * k)[0]
                s += "-|" + ("<" if obj[2 + k] else " ") + "|" + _name(obj, nb) + "|" + (">" if obj[2 + n + k] else " ") + "|-"
            lines.append(s + "\n")
        if num < hs.num_object:
            lines.append("\t...and %d other graphs\n" % (hs.num_object - num))
        text = "".join(lines)
        (file or sys.stdout).write(text)
        return text

    @staticmethod
    def serialize(state, sy_it, indentation=0, file=None):
        """qcgd.hpp:309-353: a JSON object of the state's statistics
        (total_proba, num_graphs, mean and spread of the node count and of
        the particle density, interference and deletion ratios of the last
        step; numbers as C++ streams print them).  As in the reference,
        std_dev_size is sqrt(avg_squared_size) once the variance passes the
        tolerance.  sy_it: a sy_it_t, or (num_object,
        num_object_after_interferences).  Returns the text."""
        import sys
        import quids_amd
        hs = _host_view(state)
        c, u = (sy_it.num_object, sy_it.num_object_after_interferences) if hasattr(sy_it, "num_object") else sy_it
        interference_ratio = deletion_ratio = 1.0
        if c > 0:
            interference_ratio = u / c
            deletion_ratio = hs.num_object / u
        nrm = [complex(m).real * complex(m).real + complex(m).imag * complex(m).imag for m in hs.magnitude]
        avg_size = avg_sq_size = avg_density = avg_sq_density = 0.0
        for i in range(hs.num_object):
            obj = hs.get_object(i)[0]
            n = float(num_nodes(obj))
            d = _density(obj)
            avg_size += n * nrm[i]
            avg_sq_size += n * n * nrm[i]
            avg_density += d * nrm[i]
            avg_sq_density += d * d * nrm[i]
        tol = quids_amd.tolerance
        std_dev_size = avg_sq_size - avg_size * avg_size
        std_dev_size = 0.0 if std_dev_size < tol else math.sqrt(avg_sq_size)
        std_dev_density = avg_sq_density - avg_density * avg_density
        std_dev_density = 0.0 if std_dev_density < tol else math.sqrt(std_dev_density)
        ind = "\t" * indentation
        rows = [("total_proba", "%g" % hs.total_proba), ("num_graphs", "%d" % hs.num_object),
                ("avg_size", "%g" % avg_size), ("std_dev_size", "%g" % std_dev_size),
                ("avg_density", "%g" % avg_density), ("std_dev_density", "%g" % std_dev_density),
                ("interference_ratio", "%g" % interference_ratio), ("deletion_ratio", "%g" % deletion_ratio)]
        text = "{\n" + ",\n".join(ind + '\t"%s" : %s' % kv for kv in rows) + "\n" + ind + "}"
        (file or sys.stdout).write(text)
        return text


def _strip(text, sep):
    """(text before the first sep, text after it); all of text when absent."""
    at = text.find(sep)
    return (text, "") if at < 0 else (text[:at], text[at + len(sep):])


def _value_of(text, key, sep):
    """The text after key, with the reference's substring length (first
    separator of the whole text minus the key's position; past the end when
    the key comes after that separator)."""
    at = text.find(key)
    if at < 0:
        return ""
    end = text.find(sep)
    end = len(text) if end < 0 else end
    start = at + len(key)
    return text[start:] if end < at else text[start:start + (end - at)]


def _atoi(s):
    import re
    m = re.match(r"\s*([+-]?\d+)", s)
    return int(m.group(1)) if m else 0


def _atof(s):
    import re
    m = re.match(r"\s*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?))", s)
    return float(m.group(1)) if m else 0.0


def _int_or(text, key, otherwise):
    v = _value_of(text, key, ",")
    return otherwise if v == "" else _atoi(v)


def _float_or(text, key, otherwise):
    """parse_float_with_default: a C float (the default converted too)."""
    v = _value_of(text, key, ",")
    return float(np.float32(otherwise if v == "" else _atof(v)))


class flags:
    """qcgd::flags: the command-line flags of the Quantum-graph-simulation
    runs, "<n_iter flags>|<state>|<rules>"."""

    @staticmethod
    def read_n_iter(arg):
        """qcgd.hpp:1086-1117: "n_iter,seed=,reversed_n_iter=,
        max_print_num_graphs=,tolerance=,safety_margin=,align=,
        simple_truncate=,load_balancing_bucket_per_thread=,max_num_object="
        -> (n_iter, reversed_n_iter, max_num_object).  Seeds the C library's
        rand() (with the time when no seed is given) and sets the globals;
        the tolerance passes through a C float, as in the reference."""
        import time
        import quids_amd
        head, rest = _strip(arg, ",")
        n_iters = _atoi(head)
        seed = _value_of(rest, "seed=", ",")
        libc = _c_rand()
        libc.srand(_atoi(seed) & 0xFFFFFFFF if seed != "" else int(time.time()) & 0xFFFFFFFF)
        reversed_n_iters = _int_or(rest, "reversed_n_iter=", 0)
        mp = _int_or(rest, "max_print_num_graphs=", -1 if utils.max_print_num_graphs >= (1 << 63) else
                     int(utils.max_print_num_graphs))
        utils.max_print_num_graphs = mp & ((1 << 64) - 1)
        quids_amd.tolerance = _float_or(rest, "tolerance=", quids_amd.tolerance)
        quids_amd.safety_margin = _float_or(rest, "safety_margin=", quids_amd.safety_margin)
        quids_amd.align_byte_length = _int_or(rest, "align=", int(quids_amd.align_byte_length))
        quids_amd.simple_truncation = bool(_int_or(rest, "simple_truncate=", int(quids_amd.simple_truncation)))
        quids_amd.load_balancing_bucket_per_thread = _int_or(rest, "load_balancing_bucket_per_thread=",
                                                             quids_amd.load_balancing_bucket_per_thread)
        max_num_object = _int_or(rest, "max_num_object=", 0) & ((1 << 64) - 1)
        return n_iters, reversed_n_iters, max_num_object

    @staticmethod
    def read_state(arg, state):
        """qcgd.hpp:1119-1137: "n_nodes,n_graphs=,real=,imag=;..." appends
        n_graphs cycle graphs of n_nodes with magnitude (real + i imag) /
        sqrt(n_graphs) (in C floats) to the state (an it_t or a HostState),
        then draws the particles of every graph of the state."""
        import quids_amd
        rest = arg
        while True:
            spec, rest = _strip(rest, ";")
            if spec == "":
                break
            head, opts = _strip(spec, ",")
            n_node = _atoi(head)
            n_graphs = _int_or(opts, "n_graphs=", 1)
            root = np.sqrt(np.float32(n_graphs))
            re = float(np.float32(_float_or(opts, "real=", 1.0)) / root)
            im = float(np.float32(_float_or(opts, "imag=", 0.0)) / root)
            g = make_graph(n_node)
            for _ in range(n_graphs):
                if hasattr(state, "objects"):
                    state.append(g, complex(re, im), align=quids_amd.align_byte_length)
                else:
                    state.append(g, complex(re, im))
        utils.randomize(state)
        return state

    @staticmethod
    def read_rule(arg):
        """qcgd.hpp:1139-1164: "name,theta=,phi=,xi=,n_iter=;..." with the
        angles in units of pi (computed in C floats) -> a list of (n_iter,
        is_rule, modifier, rule, reversed modifier, reversed rule); a rule's
        reverse has -xi, step and reversed_step reverse each other."""
        sim = []
        rest = arg
        while True:
            spec, rest = _strip(rest, ";")
            if spec == "":
                break
            name, opts = _strip(spec, ",")
            theta = float(np.float32(math.pi * _float_or(opts, "theta=", 0.25)))
            phi = float(np.float32(math.pi * _float_or(opts, "phi=", 0.0)))
            xi = float(np.float32(math.pi * _float_or(opts, "xi=", 0.0)))
            n = _int_or(opts, "n_iter=", 1)
            cls = {"split_merge": split_merge, "erase_create": erase_create, "coin": coin}.get(name)
            if cls is not None:
                sim.append((n, True, None, cls(theta, phi, xi), None, cls(theta, phi, -xi)))
            elif name == "step":
                sim.append((n, False, step, None, reversed_step, None))
            elif name == "reversed_step":
                sim.append((n, False, reversed_step, None, step, None))
        return sim

    @staticmethod
    def parse_simulation(arg, state):
        """qcgd.hpp:1166-1175: "<n_iter flags>|<state>|<rules>" -> (n_iter,
        reversed_n_iter, simulator, max_num_object); fills the state."""
        a, rest = _strip(arg, "|")
        n_iter, reversed_n_iters, max_num_object = flags.read_n_iter(a)
        b, rest = _strip(rest, "|")
        flags.read_state(b, state)
        return n_iter, reversed_n_iters, flags.read_rule(rest), max_num_object


def run_simulation(arg, step_rule=None, apply_modifier=None, state=None, file=None):
    """A Quantum-graph-simulation style run of "<n_iter flags>|<state>|<rules>":
    n_iter iterations of the rule list (each entry n_iter times), then
    reversed_n_iter iterations of the reversed list; prints utils::serialize
    after the parse and after every iteration as the elements of a JSON
    list, then utils::print of the final state.  step_rule(state, rule,
    max_num_object) -> (next state, (num_object,
    num_object_after_interferences)) and apply_modifier(state, modifier) ->
    state default to the device (simulate on it_t).  Returns the text."""
    import io
    import sys
    from .. import it_t, simulate, sy_it_t
    out = io.StringIO()
    if step_rule is None:
        sy = sy_it_t()

        def step_rule(st, rule, max_num_object):
            nxt = it_t()
            simulate(st, rule, nxt, sy, max_num_object)
            return nxt, (sy.num_object, sy.num_object_after_interferences)

        def apply_modifier(st, mod):
            simulate(st, mod)
            return st
    if state is None:
        state = it_t()
    n_iter, n_rev, sim, max_num_object = flags.parse_simulation(arg, state)
    counts = (0, 0)
    out.write("[\n\t")
    utils.serialize(state, counts, 1, file=out)

    def apply(entries, reverse):
        nonlocal state, counts
        for n, is_rule, mod, rule, rmod, rrule in entries:
            for _ in range(n):
                if is_rule:
                    state, counts = step_rule(state, rrule if reverse else rule, max_num_object)
                else:
                    state = apply_modifier(state, rmod if reverse else mod)

    for _ in range(n_iter):
        apply(sim, False)
        out.write(",\n\t")
        utils.serialize(state, counts, 1, file=out)
    for _ in range(n_rev):
        apply(list(reversed(sim)), True)
        out.write(",\n\t")
        utils.serialize(state, counts, 1, file=out)
    out.write("\n]\n")
    utils.print(state, file=out)
    text = out.getvalue()
    (file or sys.stdout).write(text)
    return text


if __name__ == "__main__":
    import sys as _sys
    if len(_sys.argv) != 2:
        _sys.exit('usage: python -m quids_amd.rules.qcgd "<n_iter flags>|<state>|<rules>"')
    run_simulation(_sys.argv[1])
