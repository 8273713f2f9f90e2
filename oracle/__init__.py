"""CPU oracle (test infrastructure only; see sr_oracle.c)."""
