"""FP64 CPU oracle of the SIMUS / PFIELD path (test infrastructure only)."""
