"""ctypes binding of libfeaxhip.so (the C ABI declared in include/feaxhip.h).

The library is the only compute backend.  If it is missing, importing the
compute entry points raises immediately: there is no CPU fallback.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libfeaxhip.so")

FX_MAT_LINEAR_ELASTIC = 0
FX_MAT_SIMP = 1
FX_MAT_NEO_HOOKEAN = 2
FX_MAT_CONDUCTION = 3
FX_MAT_SIMP_CONDUCTION = 4
FX_PARAM_NODE = 0
FX_PARAM_CELL = 1
FX_PCG_BLOCK3 = 1
FX_PCG_DIFF3 = 2
FX_PCG_TIME = 4
FX_PCG_AUTOTUNE = 8
FX_PCG_NO_GRAPH = 16
FX_PCG_SEG_FORCE = 64
FX_GMG_F32 = 32
FX_BAND_LDLT = 1
FX_ASM_DIFF3 = 1
FX_ASM_COMPUTE32 = 2
FX_ASM_UPPER = 4


class FxMaterial(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("param_loc", ctypes.c_int32), ("nu", ctypes.c_double),
                ("E0", ctypes.c_double), ("Emin", ctypes.c_double), ("penal", ctypes.c_double),
                ("kappa", ctypes.c_double), ("temperature", ctypes.c_void_p), ("alpha", ctypes.c_double),
                ("t_ref", ctypes.c_double)]


class FxGmgLevel(ctypes.Structure):
    """fx_gmg_level: one grid of the multigrid hierarchy."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("indptr", ctypes.c_void_p), ("bcol", ctypes.c_void_p), ("data", ctypes.c_void_p),
                ("dinv", ctypes.c_void_p), ("bc_mask", ctypes.c_void_p), ("dense_inv", ctypes.c_void_p),
                ("lmax", ctypes.c_double), ("lmin", ctypes.c_double), ("x", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("r", ctypes.c_void_p), ("d", ctypes.c_void_p)]


class FxB3Map(ctypes.Structure):
    """fx_b3map: block-column map of a block-3 CSR (device pointers)."""
    _fields_ = [("lidx", ctypes.c_void_p), ("uptr", ctypes.c_void_p), ("ucol", ctypes.c_void_p),
                ("nodes_per_block", ctypes.c_int32), ("max_unique", ctypes.c_int32)]


class FxB3Sym(ctypes.Structure):
    """fx_b3sym: upper-block storage map of a symmetric block-3 pattern (device pointers)."""
    _fields_ = [("code", ctypes.c_void_p), ("uptr", ctypes.c_void_p), ("ucol", ctypes.c_void_p),
                ("usup", ctypes.c_void_p), ("up_ptr", ctypes.c_void_p), ("nodes_per_block", ctypes.c_int32),
                ("max_unique", ctypes.c_int32), ("max_degree", ctypes.c_int32), ("num_upper", ctypes.c_int32)]


class B3Sym:
    """An FxB3Sym and the device tensors (code, uptr, ucol, usup, up_ptr) it points into."""

    def __init__(self, tensors, nodes_per_block, max_unique, max_degree, num_upper):
        self.tensors = tensors
        code, uptr, ucol, usup, up_ptr = tensors
        self.code, self.uptr, self.ucol, self.usup, self.up_ptr = tensors
        self.num_upper = int(num_upper)
        self.struct = FxB3Sym(code.data_ptr(), uptr.data_ptr(), ucol.data_ptr(), usup.data_ptr(), up_ptr.data_ptr(),
                              int(nodes_per_block), int(max_unique), int(max_degree), int(num_upper))


class B3Map:
    """An FxB3Map and the device tensors (lidx, uptr, ucol) it points into."""

    def __init__(self, tensors, nodes_per_block, max_unique):
        self.tensors = tensors
        lidx, uptr, ucol = tensors
        self.struct = FxB3Map(lidx.data_ptr(), uptr.data_ptr(), ucol.data_ptr(), int(nodes_per_block),
                              int(max_unique))


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F64 = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); every symbol include/feaxhip.h declares.
SIGNATURES = {
    "fx_last_error": (_I32, [ctypes.c_char_p, _SZ]),
    "fx_version": (_I32, [ctypes.c_char_p, _SZ]),
    "fx_hex8_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex8_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_tet4_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_tet4_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_tet4_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_tet4_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_tet4_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_tet4_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_tet4_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                    _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_hex8ut_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex8ut_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_hex8ut_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_hex8ut_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_hex8ut_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_hex8ut_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                      _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_hex20_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex20_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_hex20_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_hex20_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_hex20_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_hex20_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_hex20_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                    _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_hex20_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_hex20_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                     _P, _SZ]),
    "fx_hex20_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex20_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_csr_is_block3": (_I32, [_I32, _P, _P, _P]),
    "fx_csr_spmv_f64": (_I32, [_P, _I32, _I64, _P, _P, _P, _P, _P]),
    "fx_csr_spmv_f32": (_I32, [_P, _I32, _I64, _P, _P, _P, _P, _P]),
    "fx_csr_block3_index": (_I32, [_P, _I32, _P, _P, _P]),
    "fx_csr_spmv_b3_f64": (_I32, [_P, _I32, _I64, _P, _P, _P, _P, _P]),
    "fx_csr_spmv_b3_f32": (_I32, [_P, _I32, _I64, _P, _P, _P, _P, _P]),
    "fx_csr_spmv_diff3_f32f64": (_I32, [_P, _I32, _I64, _P, _P, _P, _P, _P]),
    "fx_b3sym_compact_f32": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P]),
    "fx_csr_spmv_b3sym_f32f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P]),
    "fx_pcg_b3sym_f32f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P, _P,
                                   _P, _P, _I32, _P, _SZ]),
    "fx_b3sym_compact_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P]),
    "fx_csr_spmv_b3sym_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P]),
    "fx_pcg_b3sym_f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P, _P,
                                _P, _P, _I32, _P, _SZ]),
    "fx_csr_jacobi_f64": (_I32, [_P, _I32, _P, _P, _P, _F64, _P]),
    "fx_csr_jacobi_f32": (_I32, [_P, _I32, _P, _P, _P, _F64, _P]),
    "fx_csr_diagonal_f64": (_I32, [_P, _I32, _P, _P, _P, _P]),
    "fx_csr_diagonal_f32": (_I32, [_P, _I32, _P, _P, _P, _P]),
    "fx_pcg_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_pcg_f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P,
                          _P, _P, _P, _I32, _P, _SZ]),
    "fx_pcg_f64_noalloc": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P,
                                  _P, _P]),
    "fx_pcg_f32": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P,
                          _P, _P, _P, _I32, _P, _SZ]),
    "fx_pcg_f32f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P,
                             _P, _P, _P, _P, _I32, _P, _SZ]),
    "fx_csr_block3_lanes": (_I32, [_I32, _I64, _I32]),
    "fx_pcg_b3map_f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P, _P,
                                _P, _P, _I32, _P, _SZ]),
    "fx_pcg_b3map_f32f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P,
                                   _P, _P, _P, _I32, _P, _SZ]),
    "fx_krylov_project_workspace_bytes": (_SZ, [_I32]),
    "fx_krylov_project_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _SZ]),
    "fx_dot_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _SZ]),
    "fx_axpy_f32_into_f64": (_I32, [_P, _I32, _F64, _P, _P]),
    "fx_convert_f64_to_f32": (_I32, [_P, _I32, _F64, _P, _P]),
    "fx_axpby_f64": (_I32, [_P, _I32, _F64, _P, _F64, _P, _P]),
    "fx_bc_lift_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P]),
    "fx_hex8_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _I32, _I32, _F64, _P, _P, _P, _SZ, _P, _P]),
    "fx_hex8_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _I32, _I32, _F64, _P, _P, _P, _SZ, _P, _P]),
    "fx_hex8_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_hex8_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_hex8_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                    _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_hex8_jvp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_pcg_matfree_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P, _F64,
                                  _F64, _I32, _I32, _P, _P, _P, _P, _SZ]),
    "fx_pcg_batched_workspace_bytes": (_SZ, [_I32, _P, _I64, _I32]),
    "fx_pcg_batched_f64": (_I32, [_P, _I32, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P,
                                  _P, _P, _P, _P, _SZ]),
    "fx_pcg_batched_f32f64": (_I32, [_P, _I32, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32,
                                     _P, _P, _P, _P, _P, _SZ]),
    "fx_pcg_batched_record_f32f64": (_I32, [_P, _I32, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32,
                                            _I32, _P, _P, _P, _P, _P, _P, _I32, _P, _SZ]),
    "fx_seg_project_workspace_bytes": (_SZ, [_I32, _P, _I32]),
    "fx_seg_project_f32": (_I32, [_P, _I32, _P, _I32, _P, _P, _P, _P, _P, _P, _SZ]),
    "fx_seg_dot_f64": (_I32, [_P, _I32, _P, _P, _P, _P]),
    "fx_hex8_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_hex8_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                     _P, _SZ]),
    "fx_hex8_thermal_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P, _P,
                                       _SZ]),
    "fx_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex8_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_tet4_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_tet4_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                     _P, _SZ]),
    "fx_tet4_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_tet4_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_quad4s_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_quad4s_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_quad4s_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_quad4s_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_quad4s_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_quad4s_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_quad4s_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                     _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_quad4s_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_quad4s_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                      _P, _SZ]),
    "fx_quad4s_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_quad4s_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_tet4s_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_tet4s_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_tet4s_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_tet4s_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_tet4s_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_tet4s_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_tet4s_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                     _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_tet4s_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_tet4s_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                      _P, _SZ]),
    "fx_tet4s_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_tet4s_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_hex8s_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex8s_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_hex8s_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_hex8s_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_hex8s_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_hex8s_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_hex8s_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                     _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_hex8s_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_hex8s_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                      _P, _SZ]),
    "fx_hex8s_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_hex8s_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_quad4_pattern": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_quad4_face_incidence": (_I32, [_I32, _I32, _P, _I32, _P, _P, _P]),
    "fx_quad4_assemble_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_quad4_assemble_f32": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I32, _I32, _F64, _P, _P, _P, _SZ]),
    "fx_quad4_assemble_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_quad4_residual_workspace_bytes": (_SZ, [_I32]),
    "fx_quad4_residual_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                     _P, _P, _I32, _P, _P, _I32, _P, _P, _SZ]),
    "fx_quad4_param_vjp_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_quad4_param_vjp_f64": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P,
                                      _P, _SZ]),
    "fx_quad4_simp_compliance_sens_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fx_quad4_traction_vjp_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _P]),
    "fx_tet4_energy_workspace_bytes": (_SZ, [_I32]),
    "fx_tet4_energy_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ]),
    "fx_hex8_energy_workspace_bytes": (_SZ, [_I32]),
    "fx_hex8_energy_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ]),
    "fx_bicgstab_workspace_bytes": (_SZ, [_I32]),
    "fx_bicgstab_f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P, _P,
                               _P, _SZ]),
    "fx_band_ldab": (_I64, [_I32]),
    "fx_band_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_band_from_csr_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _I32, _I64, _P, _P, _SZ]),
    "fx_band_cholesky_f64": (_I32, [_P, _I32, _I32, _I64, _P, _P, _P, _SZ]),
    "fx_band_ldlt_f64": (_I32, [_P, _I32, _I32, _I64, _P, _P, _P, _SZ]),
    "fx_band_inv_block": (_I32, [_I32]),
    "fx_band_inv_bytes": (_SZ, [_I32, _I32]),
    "fx_band_solve_inv_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_band_invert_diag_f64": (_I32, [_P, _I32, _I32, _I64, _I32, _P, _P]),
    "fx_band_solve_inv_f64": (_I32, [_P, _I32, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _SZ]),
    "fx_band_solve_f64": (_I32, [_P, _I32, _I32, _I64, _I32, _P, _P, _P, _P, _P, _SZ]),
    "fx_gmres_workspace_bytes": (_SZ, [_I32, _I32, _I32]),
    "fx_gmres_f64": (_I32, [_P, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _F64, _F64, _I32, _I32, _I32, _P, _P,
                            _P, _P, _SZ]),
    "fx_gmg_eig_bounds_workspace_bytes": (_SZ, [_I32, _I32]),
    "fx_gmg_eig_bounds_f64": (_I32, [_P, _I32, _P, _P, _P, _P, _I32, _P, _P, _P, _SZ]),
    "fx_gmg_vcycle": (_I32, [_P, _I32, _I32, _P, _I32, _I32, _F64, _P, _P]),
    "fx_pcg_gmg_workspace_bytes": (_SZ, [_I32]),
    "fx_pcg_gmg_f64": (_I32, [_P, _I32, _I32, _P, _I32, _I32, _F64, _P, _P, _P, _F64, _F64, _I32, _I32, _P, _P,
                              _P, _P, _SZ]),
    "fx_gmg_coarsen_f64": (_I32, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P]),
    "fx_csr_to_dense_f64": (_I32, [_P, _I32, _P, _P, _P, _P]),
    "fx_gmg_galerkin_workspace_bytes": (_SZ, [_I32, _I32, _I32, _I32, _I32, _I32]),
    "fx_gmg_galerkin_dense_f64": (_I32, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _SZ]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load libfeaxhip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is missing: build it with `python -m feax_amd._build` "
                "(there is no CPU fallback)")
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7) as the one the library links against.  Loading
        # torch first makes the process use one runtime, torch's, for the
        # streams and device memory it hands to the library; loading the
        # library first would put a second runtime beside torch's.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib().fx_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise NativeError(f"{what} failed with code {rc}: {last_error()}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def ptr(t):
    """Raw pointer of a torch tensor / numpy array / None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def stream():
    """hipStream_t of torch's current stream on the current device."""
    import torch
    return torch.cuda.current_stream().cuda_stream


def material_struct(kind, param_loc, nu=0.3, E0=1.0, Emin=1e-9, penal=3.0, kappa=0.0, temperature=None,
                    alpha=0.0, t_ref=0.0):
    return FxMaterial(int(kind), int(param_loc), float(nu), float(E0), float(Emin), float(penal), float(kappa),
                      temperature, float(alpha), float(t_ref))
