"""CPU checkers for the HIP path (test infrastructure only)."""
