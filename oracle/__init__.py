"""CPU oracle for the nalgebra dense hot path (test infrastructure only)."""
