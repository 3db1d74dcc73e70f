"""quantum_computer rules (reference: src/rules/quantum_computer.hpp).

``hadamard(bit)`` is a rule (two children); ``cnot``, ``Xgate``, ``Ygate`` and
``Zgate`` are modifiers.  Objects are byte registers, one byte per qubit.
"""
import math

from .._abi import (QD_MOD_CNOT, QD_MOD_XGATE, QD_MOD_YGATE, QD_MOD_ZGATE, QD_RULE_HADAMARD, QdModifier,
                    QdRule)
from ._base import DeviceModifier, DeviceRule


class hadamard(DeviceRule):
    """quantum_computer.hpp:31-50: child 0 flips ``bit`` with factor 1/sqrt(2),
    child 1 keeps it with factor -1/sqrt(2) if the bit is set, else 1/sqrt(2)."""

    def __init__(self, bit):
        self.bit = int(bit)
        self.sqrt2 = 1 / math.sqrt(2)  # quantum_computer.hpp:41

    def desc(self):
        return QdRule(kind=QD_RULE_HADAMARD, bit=self.bit, coef=self.sqrt2)


class cnot(DeviceModifier):
    """quantum_computer.hpp:25-29: begin[bit] ^= begin[control_bit]."""

    def __init__(self, control_bit, bit):
        self.control_bit, self.bit = int(control_bit), int(bit)

    def desc(self):
        return QdModifier(kind=QD_MOD_CNOT, control_bit=self.control_bit, bit=self.bit)


class Xgate(DeviceModifier):
    """quantum_computer.hpp:52-56."""

    def __init__(self, bit):
        self.bit = int(bit)

    def desc(self):
        return QdModifier(kind=QD_MOD_XGATE, bit=self.bit)


class Ygate(DeviceModifier):
    """quantum_computer.hpp:58-66: mag *= i, times -1 if the bit is set, then flip."""

    def __init__(self, bit):
        self.bit = int(bit)

    def desc(self):
        return QdModifier(kind=QD_MOD_YGATE, bit=self.bit)


class Zgate(DeviceModifier):
    """quantum_computer.hpp:68-75: times -1 if the bit is set, and the bit is
    flipped as well (the reference's definition)."""

    def __init__(self, bit):
        self.bit = int(bit)

    def desc(self):
        return QdModifier(kind=QD_MOD_ZGATE, bit=self.bit)
