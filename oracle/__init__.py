"""CPU oracle for the CorePotts hot path (test infrastructure only)."""
