"""CPU oracle (test infrastructure): FP64 restatement of ReaDDy's CPU kernel per-step path."""
