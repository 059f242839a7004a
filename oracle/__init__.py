"""FP64 oracle package (test infrastructure only)."""
