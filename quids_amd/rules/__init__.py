"""Rules and modifiers with a device implementation (reference: src/rules/)."""
from . import qcgd, quantum_computer  # noqa: F401
