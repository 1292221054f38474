"""FP64 CPU oracle of the quasi-static xi-slice sweep (test infrastructure only)."""
