"""CPU oracle (test infrastructure only): a restatement of the reference's CPU engine."""
