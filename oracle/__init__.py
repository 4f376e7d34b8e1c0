"""CPU oracle of the GENFIRE hot path (test infrastructure only; see oracle.py)."""
