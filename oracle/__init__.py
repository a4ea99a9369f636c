"""CPU oracle of the hot path (test infrastructure only: imported by tests/, the smoke
check and the benchmark's CPU-baseline leg, never by the product)."""
