"""CPU oracle of the CT reconstruction path (test infrastructure only)."""
