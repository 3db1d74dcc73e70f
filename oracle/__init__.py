"""CPU checker for the HIP library (test infrastructure only)."""
