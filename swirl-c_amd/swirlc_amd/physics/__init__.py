"""physics."""
