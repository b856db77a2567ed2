"""CPU checkers for delfem_amd (test infrastructure only; see oracle.py)."""
