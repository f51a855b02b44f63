"""FP64 CPU oracle (test infrastructure only)."""
