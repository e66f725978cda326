"""CPU oracle for the CAMP hot path: test infrastructure only (see oracle.py)."""
