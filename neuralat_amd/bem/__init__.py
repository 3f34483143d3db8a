"""BEM Helmholtz solver on MI355X (drop-in for the reference's src/bem)."""
from .solver import (BEM_Solver, GeometrySet, Mesh, bicgstab, boundary2potential_samples, burton_miller_samples,
                     check_tensor, combine, gmres, map_triangle2vertex, matvec, matvec_strided, preprocess,
                     solve_linear_equation)

__all__ = ["BEM_Solver", "GeometrySet", "Mesh", "bicgstab", "boundary2potential_samples", "burton_miller_samples",
           "check_tensor", "combine", "gmres", "map_triangle2vertex", "matvec", "matvec_strided", "preprocess",
           "solve_linear_equation"]
