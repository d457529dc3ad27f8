"""CPU FP64 oracle of the wave-equation hot path (test infrastructure only)."""
