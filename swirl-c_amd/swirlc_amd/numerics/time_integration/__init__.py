"""numerics/time_integration."""
