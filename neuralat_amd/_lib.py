"""ctypes binding of libbem_hip.so (C ABI in include/bem_hip.h).

The library is built in-tree by ``neuralat_amd/csrc/Makefile`` (see
``__graft_entry__.build``).  There is no fallback: if the library is missing
or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbem_hip.so")

BEM_OK, BEM_ERR_BAD_ARG, BEM_ERR_BAD_MESH, BEM_ERR_HIP, BEM_ERR_NOT_CONVERGED, BEM_ERR_NO_DEVICE, BEM_ERR_OOM = range(7)

OPS = {
    "single": 0,
    "double": 1,
    "adjointdouble": 2,
    "hypersingular": 3,
    "bm_lhs": 4,
    "bm_rhs": 5,
}
BASES = {"P1": 0, "DP0": 1}
QUADS = {"full": 0, "approx": 1, "approx_1": 2}


class BemError(RuntimeError):
    """A HIP or library failure."""


class NotConvergedError(BemError):
    """A solver stopped before reaching its tolerance."""

    def __init__(self, msg, info=None):
        super().__init__(msg)
        self.info = info


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("rel_residual", ctypes.c_double),
                ("implicit_rel_residual", ctypes.c_double)]

    def as_dict(self):
        return {"iterations": int(self.iterations), "converged": bool(self.converged),
                "rel_residual": float(self.rel_residual), "implicit_rel_residual": float(self.implicit_rel_residual)}


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double

# name -> (restype, argtypes); exactly the functions include/bem_hip.h declares
SIGNATURES = {
    "bem_version": (ctypes.c_char_p, []),
    "bem_last_error": (ctypes.c_char_p, []),
    "bem_mesh_create": (_I32, [_P, _I64, _P, _I64, _I32, _P, ctypes.POINTER(_P)]),
    "bem_mesh_destroy": (_I32, [_P]),
    "bem_mesh_info": (_I32, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                             ctypes.POINTER(ctypes.c_int32)]),
    "bem_mesh_normals_curls": (_I32, [_P, _P, _P, _P]),
    "bem_assemble": (_I32, [_P, _F, _I32, _I32, _I32, _I32, _P, _I64, _P]),
    "bem_assemble_ex": (_I32, [_P, _F, _I32, _I32, _I32, _I32, _I32, _P, _I64, _P]),
    "bem_assemble_dk": (_I32, [_P, _F, _I32, _I32, _I32, _P, _I64, _P]),
    "bem_burton_miller_system": (_I32, [_P, _F, _I32, _I32, _P, _P, _I64, _P, _P]),
    "bem_burton_miller_systems": (_I32, [_P, _P, _I32, _I32, _I32, _P, _P, _I64, _I64, _P, _P]),
    "bem_burton_miller_samples": (_I32, [_P, _P, _I32, _I32, _P, _P, _I64, _I64, _P, _P]),
    "bem_assemble_batch_k": (_I32, [_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _I64, _I64, _P]),
    "bem_cbie_work_size": (_I64, [_P]),
    "bem_cbie_dp0": (_I32, [_P, _F, _I32, _I32, _P, _P, _I64, _P, _P, _P]),
    "bem_cbie_dp0_ex": (_I32, [_P, _F, _I32, _I32, _I32, _P, _P, _I64, _P, _P, _P]),
    "bem_add_identity": (_I32, [_P, _I32, _F, _F, _P, _I64, _P]),
    "bem_assemble_rows": (_I32, [_P, _F, _I32, _I32, _I32, _I64, _I64, _P, _I64, _P]),
    "bem_geomset_create": (_I32, [_P, _P, _I32, _P, ctypes.POINTER(_P)]),
    "bem_geomset_destroy": (_I32, [_P]),
    "bem_geomset_count": (_I32, [_P]),
    "bem_assemble_samples": (_I32, [_P, _P, _I32, _I32, _I32, _I32, _P, _I64, _I64, _P]),
    "bem_add_identity_samples": (_I32, [_P, _I32, _P, _P, _I64, _I64, _P]),
    "bem_potential_samples": (_I32, [_P, _P, _I64, _P, _I32, _I32, _I32, _P, _I64, _I64, _P]),
    "bem_matvec_strided": (_I32, [_P, _I64, _I64, _I64, _I64, _P, _I64, _P, _I64, _I32, _F, _F, _F, _F, _P]),
    "bem_boundary_matrix": (_I32, [_P, _I64, _P, _I64, _F, _I32, _I32, _P, _P]),
    "bem_identity": (_I32, [_P, _I32, _P, _I64, _P]),
    "bem_triangle2vertex": (_I32, [_P, _P, _P, _P]),
    "bem_triangle2vertex_work_size": (_I64, [_I64, _I64]),
    "bem_triangle2vertex_raw": (_I32, [_P, _I64, _I64, _P, _P, _P, _P]),
    "bem_potential": (_I32, [_P, _P, _I64, _F, _I32, _I32, _I32, _P, _I64, _P]),
    "bem_potential_dk": (_I32, [_P, _P, _I64, _F, _I32, _I32, _I32, _P, _I64, _P]),
    "bem_mesh_set_vertices": (_I32, [_P, _P, _I32, _P]),
    "bem_potential_batch_k": (_I32, [_P, _P, _I64, _P, _I32, _I32, _I32, _I32, _P, _I64, _I64, _P]),
    "bem_matvec_work_size": (_I64, [_I64, _I64, _I32]),
    "bem_matvec": (_I32, [_P, _I64, _I64, _I64, _I32, _P, _P, _F, _F, _F, _F, _P, _P]),
    "bem_combine": (_I32, [_P, _I64, _P, _I64, _P, _I64, _I64, _I64, _F, _F, _F, _F, _P]),
    "bem_gmres_work_size": (_I64, [_I64, _I32]),
    "bem_gmres": (_I32, [_P, _I64, _I64, _P, _P, _F, _I32, _I32, _P, ctypes.POINTER(SolveInfo), _P]),
    "bem_bicgstab_work_size": (_I64, [_I64]),
    "bem_bicgstab": (_I32, [_P, _I64, _I64, _P, _P, _I32, _D, _D, _P, ctypes.POINTER(SolveInfo), _P]),
    "bem_lowp_ld": (_I64, [_I64]),
    "bem_lowp_plane_elems": (_I64, [_I64, _I64]),
    "bem_convert_lowp": (_I32, [_P, _I64, _I64, _I64, _I32, _F, _P, _P, _I64, ctypes.POINTER(_F), _P]),
    "bem_matmul_work_size": (_I64, [_I64]),
    "bem_matmul": (_I32, [_P, _I64, _I64, _I64, _P, _I64, _I32, _P, _I64, _F, _F, _F, _F, _P, _P]),
    "bem_matmul_lowp_work_size": (_I64, [_I64]),
    "bem_matmul_lowp": (_I32, [_P, _P, _I32, _I64, _I64, _I64, _F, _P, _I64, _I32, _P, _I64, _F, _F, _F, _F, _P,
                               _P]),
    "bem_gmres_multi_work_size": (_I64, [_I64, _I32, _I32, _I64]),
    "bem_gmres_multi": (_I32, [_P, _I64, _P, _P, _I32, _I64, _F, _I64, _I32, _P, _I64, _P, _I64, _F, _I32, _I32, _P,
                               ctypes.POINTER(SolveInfo), _P]),
    "bem_solve_ir_work_size": (_I64, [_I64, _I32, _I32, _I64]),
    "bem_solve_ir": (_I32, [_P, _I64, _P, _P, _I32, _I64, _F, _I64, _I32, _P, _I64, _P, _I64, _F, _F, _I32, _I32,
                            _I32, _P, ctypes.POINTER(SolveInfo), ctypes.POINTER(ctypes.c_int32), _P]),
    "bem_batch_cbie_work_size": (_I64, [_P, _I32]),
    "bem_batch_cbie": (_I32, [_P, _I32, _I32, _I32, _P, _P]),
    "bem_batch_autotune": (_I32, [_P, _I32, _I32, _P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_F), _P]),
    "bem_batch_gmres_work_size": (_I64, [_P, _I32, _I32]),
    "bem_batch_gmres": (_I32, [_P, _I32, _F, _I32, _I32, _I32, _P, ctypes.POINTER(SolveInfo), _P]),
    "bem_batch_fgmres_work_size": (_I64, [_P, _I32, _I32, _I32]),
    "bem_batch_fgmres": (_I32, [_P, _I32, _I32, _F, _F, _I32, _I32, _I32, _I32, _I32, _P,
                                ctypes.POINTER(SolveInfo), _P, _P]),
    "bem_batch_last_error": (ctypes.c_char_p, []),
    "bem_fem_create": (_I32, [_P, _I64, _P, _I64, ctypes.POINTER(ctypes.c_void_p)]),
    "bem_fem_free": (None, [_P]),
    "bem_fem_num_blocks": (_I64, [_P]),
    "bem_fem_pattern": (_I32, [_P, _P, _P, _P, _P]),
    "bem_fem_assemble": (_I32, [_P, _I32, _F, _F, _F, _I32, _P, _P]),
    "bem_fem_spmm": (_I32, [_P, _I32, _P, _P, _I64, _I32, _P, _I64, _P]),
    "bem_fem_last_error": (ctypes.c_char_p, []),
    "bem_gmres_ex": (_I32, [_P, _I64, _I64, _I32, _P, _P, _F, _I32, _I32, _P, ctypes.POINTER(SolveInfo), _P]),
    "bem_gmres_noalloc_capacity": (_I64, []),
    "bem_gmres_noalloc": (_I32, [_P, _I64, _I64, _I32, _P, _P, _F, _I32, _I32, ctypes.POINTER(SolveInfo), _P]),
    "bem_gmres_timed": (_I32, [_P, _I64, _I64, _I32, _P, _P, _F, _I32, _I32, _P, ctypes.POINTER(SolveInfo),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64), _P]),
    "bem_adjoint_work_size": (_I64, [_P]),
    "bem_adjoint_cbie_dp0": (_I32, [_P, _F, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "bem_adjoint_potential_dp0": (_I32, [_P, _P, _I64, _F, _I32, _P, _P, _P, _P, _P, _P]),
    "bem_adjoint_bm_p1": (_I32, [_P, _F, _I32, _P, _P, _P, _P, _P]),
    "bem_adjoint_potential_p1": (_I32, [_P, _P, _I64, _F, _I32, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load libbem_hip.so once; raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C neuralat_amd/csrc` "
                          "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().bem_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "bem"):
    if rc == BEM_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc in (BEM_ERR_BAD_ARG, BEM_ERR_BAD_MESH):
        raise ValueError(msg)
    if rc == BEM_ERR_NOT_CONVERGED:
        raise NotConvergedError(msg)
    if rc == BEM_ERR_OOM:
        raise MemoryError(msg)
    raise BemError(f"{msg} (status {rc})")
