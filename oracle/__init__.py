"""FP64 CPU oracle of the reference path (test infrastructure only)."""
