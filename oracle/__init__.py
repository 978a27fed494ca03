"""FP64 CPU oracle of the felbm Lee-Liu step (test infrastructure only)."""
