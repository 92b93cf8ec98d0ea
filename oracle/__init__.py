"""CPU oracle for the Limber C_ell path (test infrastructure only)."""
