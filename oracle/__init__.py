"""CPU oracle for the MDS likelihood -- test infrastructure only (see oracle.py)."""
