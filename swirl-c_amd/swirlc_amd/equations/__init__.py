"""equations."""
