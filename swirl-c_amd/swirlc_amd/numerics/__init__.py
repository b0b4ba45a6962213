"""numerics."""
