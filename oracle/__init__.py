"""CPU oracle for the Helmholtz SL-MG path (test infrastructure only; see the module headers)."""
