"""CPU oracle for the NMF multiplicative-update path (test infrastructure only)."""
