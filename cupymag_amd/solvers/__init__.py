from .linear_solvers import solve_cg  # noqa: F401
