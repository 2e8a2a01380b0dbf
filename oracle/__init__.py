"""FP64 CPU oracle of the complex-Langevin update (test infrastructure only)."""
