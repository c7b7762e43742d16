"""CPU oracle of the SPH step (test infrastructure only)."""
