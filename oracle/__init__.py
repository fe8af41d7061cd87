"""CPU oracle of the correlator path (test infrastructure only)."""
