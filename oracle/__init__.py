"""CPU restatement of the reference semantics (test infrastructure only)."""
