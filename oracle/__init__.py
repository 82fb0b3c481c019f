"""CPU oracle of the 2-D Euler FV step (test infrastructure only)."""
