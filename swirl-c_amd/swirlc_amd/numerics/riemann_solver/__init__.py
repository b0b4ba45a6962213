"""numerics/riemann_solver."""
