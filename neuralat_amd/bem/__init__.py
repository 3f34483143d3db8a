"""BEM Helmholtz solver on MI355X (drop-in for the reference's src/bem)."""
from .solver import (BEM_Solver, Mesh, bicgstab, check_tensor, combine, gmres, map_triangle2vertex, matvec,
                     preprocess, solve_linear_equation)

__all__ = ["BEM_Solver", "Mesh", "bicgstab", "check_tensor", "combine", "gmres", "map_triangle2vertex", "matvec",
           "preprocess", "solve_linear_equation"]
