"""CPU oracle (test infrastructure only; see mco.c)."""
