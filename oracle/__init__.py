"""CPU oracle of the rasterization path (test infrastructure only)."""
