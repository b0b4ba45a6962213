"""boundary."""
