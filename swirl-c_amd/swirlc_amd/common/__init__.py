"""common."""
