"""FP64 oracle of the projector arithmetic (test infrastructure only)."""
