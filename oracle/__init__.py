"""CPU checker for the radiative solver (test infrastructure only)."""
