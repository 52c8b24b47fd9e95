"""ctypes binding of libfem_mi355x.so (declarations in include/fem_mi355x.h).

The library is the only compute path: if it is missing or no GPU is visible,
``lib()`` raises instead of falling back to anything else.
"""
import ctypes as C
import os

import torch

LIB_NAME = 'libfem_mi355x.so'
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

FEM_OK, FEM_ERR_BAD_ARG, FEM_ERR_NOT_CONVERGED, FEM_ERR_NAN, FEM_ERR_HIP, FEM_ERR_CAPACITY = 0, -1, -2, -3, -4, -5
FEM_F32, FEM_F64, FEM_BF16X3 = 0, 1, 2
FEM_FORM_STANDARD, FEM_FORM_DIFFERENCE, FEM_FORM_SCALED_DIFF16 = 0, 1, 2
FEM_MAT_POISSON, FEM_MAT_LINEAR_ELASTIC, FEM_MAT_SIMP, FEM_MAT_NEO_HOOKEAN, FEM_MAT_J2_PLASTIC = 1, 2, 3, 4, 5
FEM_MAT_TABULATED = 6
FEM_MAT_NONE = 0


def FEM_TAB_STRIDE(vec):
    return 6 * vec + 9 * vec * vec

_vp = C.c_void_p
_i32, _i64, _f64 = C.c_int32, C.c_int64, C.c_double


class FemMesh(C.Structure):
    _fields_ = [('nodes_per_cell', _i32), ('num_quads', _i32), ('vec', _i32), ('num_face_quads', _i32),
                ('num_cells', _i64), ('num_nodes', _i64),
                ('points', _vp), ('cells', _vp), ('shape_vals', _vp), ('shape_grads_ref', _vp),
                ('quad_weights', _vp), ('face_shape_vals', _vp), ('face_shape_grads_ref', _vp),
                ('face_weights', _vp), ('face_normals', _vp)]


class FemMaterial(C.Structure):
    _fields_ = [('kind', _i32), ('has_param', _i32), ('c', _f64 * 6), ('param', _vp), ('state', _vp)]


class FemPattern(C.Structure):
    _fields_ = [('num_rows', _i64), ('nnzb', _i64), ('num_slices', _i64), ('sell_blocks', _i64),
                ('vec', _i32), ('num_cells', _i32), ('nodes_per_cell', _i32), ('pad', _i32),
                ('row_ptr', _vp), ('col_idx', _vp), ('diag_slot', _vp), ('contrib_ptr', _vp),
                ('contrib_idx', _vp), ('slice_ptr', _vp), ('sell_col', _vp), ('sell_row', _vp),
                ('sell_pos', _vp)]


class FemOperator(C.Structure):
    _fields_ = [('mesh', C.POINTER(FemMesh)), ('mat', C.POINTER(FemMaterial)), ('u', _vp),
                ('inc_ptr', _vp), ('inc_idx', _vp), ('n_bc', _i64), ('bc_dofs', _vp),
                ('cell_scratch', _vp), ('mass_M', _vp), ('cell_K', _vp)]


class FemConstraint(C.Structure):
    _fields_ = [('rows', _i64), ('cols', _i64), ('p_ptr', _vp), ('p_idx', _vp), ('p_val', _vp),
                ('pt_ptr', _vp), ('pt_idx', _vp), ('pt_val', _vp)]


class FemAssemblyPlan(C.Structure):
    _fields_ = [('slice_cell_ptr', _vp), ('slice_cells', _vp), ('slice_pair_ptr', _vp), ('slice_pairs', _vp),
                ('contrib_loc', _vp), ('max_slots', _i32), ('max_pairs', _i32)]


class FemPcgOptions(C.Structure):
    _fields_ = [('rtol', _f64), ('atol', _f64), ('maxiter', _i32), ('check_every', _i32),
                ('precision', _i32), ('precond', _i32), ('reliable_delta', _f64),
                ('xcd_remap', _i32), ('timing', _i32)]


class FemPcgResult(C.Structure):
    _fields_ = [('iterations', _i32), ('replacements', _i32), ('status', _i32), ('pad', _i32),
                ('b_norm', _f64), ('res_norm', _f64), ('spmv_ms', _f64), ('update_ms', _f64),
                ('timed_launches', _i32), ('pad2', _i32)]


class FemBcsr(C.Structure):
    _fields_ = [('rows', _i64), ('cols', _i64), ('nnzb', _i64), ('br', _i32), ('bc', _i32), ('f32', _i32),
                ('pad', _i32), ('row_ptr', _vp), ('col_idx', _vp), ('vals', _vp)]


class FemAmgLevel(C.Structure):
    _fields_ = [('pat', C.POINTER(FemPattern)), ('sell_vals', _vp), ('A', FemBcsr), ('P', FemBcsr),
                ('R', FemBcsr), ('dinv', _vp), ('lmax', _f64)]


class FemAmg(C.Structure):
    _fields_ = [('num_levels', _i32), ('precision', _i32), ('degree', _i32), ('level0_form', _i32),
                ('levels', C.POINTER(FemAmgLevel)), ('coarse_n', _i64), ('coarse_inv', _vp)]


class FemAmgOptions(C.Structure):
    _fields_ = [('rtol', _f64), ('atol', _f64), ('maxiter', _i32), ('check_every', _i32), ('timing', _i32),
                ('pad', _i32), ('reliable_delta', _f64)]


_P = C.POINTER
_SIGNATURES = {
    'fem_last_error': (C.c_char_p, []),
    'fem_version': (C.c_int, []),
    'fem_incidence': (C.c_int, [_i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp]),
    'fem_scan_ws_bytes': (_i64, [_i64]),
    'fem_exclusive_scan': (C.c_int, [_i64, _vp, _vp, _vp, _vp]),
    'fem_pattern_rows': (C.c_int, [_P(FemMesh), _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_pattern_fill': (C.c_int, [_P(FemMesh), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_sell_order': (C.c_int, [_i64, _i32, _vp, _vp, _vp, _vp]),
    'fem_sell_slice_sizes': (C.c_int, [_i64, _vp, _vp, _vp, _vp]),
    'fem_sell_fill': (C.c_int, [_P(FemPattern), _vp, _vp]),
    'fem_geometry': (C.c_int, [_P(FemMesh), _vp, _vp, _vp, _vp]),
    'fem_cell_tangent': (C.c_int, [_P(FemMesh), _P(FemMaterial), _i32, _vp, _vp, _vp]),
    'fem_quad_grad': (C.c_int, [_P(FemMesh), _vp, _vp, _vp, _vp]),
    'fem_cell_mass_tangent': (C.c_int, [_P(FemMesh), _vp, _i32, _vp, _vp]),
    'fem_cell_mass_jvp': (C.c_int, [_P(FemMesh), _vp, _vp, _vp, _vp]),
    'fem_cell_matvec': (C.c_int, [_P(FemMesh), _vp, _vp, _vp, _vp]),
    'fem_cell_residual': (C.c_int, [_P(FemMesh), _P(FemMaterial), _vp, _vp, _vp]),
    'fem_cell_jvp': (C.c_int, [_P(FemMesh), _P(FemMaterial), _vp, _vp, _vp, _vp]),
    'fem_cell_load': (C.c_int, [_P(FemMesh), _vp, _vp, _vp]),
    'fem_cell_load_adjoint': (C.c_int, [_P(FemMesh), _vp, _vp, _vp]),
    'fem_cell_adjoint': (C.c_int, [_P(FemMesh), _P(FemMaterial), _vp, _vp, _vp, _vp]),
    'fem_face_geometry': (C.c_int, [_P(FemMesh), _i64, _vp, _vp, _vp, _vp, _vp]),
    'fem_face_load': (C.c_int, [_P(FemMesh), _i64, _vp, _vp, _vp, _vp]),
    'fem_face_load_adjoint': (C.c_int, [_P(FemMesh), _i64, _vp, _vp, _vp, _vp]),
    'fem_gather_nodal': (C.c_int, [_i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_gather_matrix': (C.c_int, [_P(FemPattern), _i32, _vp, _i32, _i32, _vp, _vp, _vp]),
    'fem_matrix_bc_rows': (C.c_int, [_P(FemPattern), _i32, _vp, _i64, _vp, _vp, _vp]),
    'fem_jacobi': (C.c_int, [_P(FemPattern), _i32, _vp, _vp, _vp]),
    'fem_residual_bc': (C.c_int, [_i64, _vp, _vp, _vp, _vp, _f64, _vp]),
    'fem_spmv': (C.c_int, [_P(FemPattern), _i32, _i32, _vp, _vp, _vp, _vp]),
    'fem_export_csr': (C.c_int, [_P(FemPattern), _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    'fem_operator_apply': (C.c_int, [_P(FemOperator), _vp, _vp, _vp]),
    'fem_assembly_plan_counts': (C.c_int, [_P(FemPattern), _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_assembly_plan_fill': (C.c_int, [_P(FemPattern), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_assemble': (C.c_int, [_P(FemMesh), _P(FemMaterial), _P(FemPattern), _P(FemAssemblyPlan), _vp, _i32,
                               _i32, _vp, _vp, _vp, _vp, _vp]),
    'fem_csr_spmv': (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_bicgstab_ws_bytes': (_i64, [_P(FemPattern), _P(FemConstraint)]),
    'fem_bicgstab_solve': (C.c_int, [_P(FemPattern), _vp, _vp, _P(FemConstraint), _vp, _vp, _P(FemPcgOptions),
                                     _vp, _vp, _P(FemPcgResult), _vp]),
    'fem_pcg_ws_bytes': (_i64, [_P(FemPattern), _i32]),
    'fem_pcg_solve': (C.c_int, [_P(FemPattern), _vp, _vp, _P(FemOperator), _vp, _vp, _P(FemPcgOptions),
                                _vp, _vp, _P(FemPcgResult), _vp]),
    'fem_pcg_solve_noalloc': (C.c_int, [_P(FemPattern), _vp, _vp, _P(FemOperator), _vp, _vp, _P(FemPcgOptions),
                                        _vp, _P(FemPcgResult), _vp]),
    'fem_dot': (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    'fem_amg_aggregate_ws_bytes': (_i64, [_i64]),
    'fem_amg_aggregate': (C.c_int, [_i64, _vp, _vp, _vp, _P(_i64), _vp, _vp, _vp]),
    'fem_amg_members': (C.c_int, [_i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    'fem_amg_tentative': (C.c_int, [_i64, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fem_amg_bc_mask': (C.c_int, [_i64, _i64, _vp, _vp, _vp]),
    'fem_amg_sell_to_bcsr': (C.c_int, [_P(FemPattern), _vp, _vp, _i32, _vp]),
    'fem_amg_sell_to_f32_diff': (C.c_int, [_P(FemPattern), _vp, _vp, _vp]),
    'fem_amg_sell_to_f32_copies': (C.c_int, [_P(FemPattern), _vp, _vp, _vp, _vp, _vp]),
    'fem_amg_dinv': (C.c_int, [_P(FemBcsr), _i32, _vp, _vp]),
    'fem_amg_lmax_ws_bytes': (_i64, [_i64]),
    'fem_amg_lmax': (C.c_int, [_P(FemBcsr), _P(FemPattern), _vp, _i32, _vp, _i32, _P(_f64), _vp, _vp]),
    'fem_amg_lmax_async': (C.c_int, [_P(FemBcsr), _P(FemPattern), _vp, _i32, _vp, _i32, _vp, _vp, _vp]),
    'fem_bcsr_spgemm_symbolic': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _vp, _vp, _P(_i64), _vp, _vp, _vp]),
    'fem_bcsr_spgemm_plan': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _i64, _P(FemBcsr), _vp, _vp, _vp, _vp, _vp,
                                       _vp]),
    'fem_bcsr_spgemm_numeric': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _P(FemBcsr), _vp, _vp, _vp]),
    'fem_amg_smooth_prolongator': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _P(FemBcsr), _vp, _vp, _vp, _vp, _f64,
                                             _vp]),
    'fem_amg_smooth_prolongator_dev': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _P(FemBcsr), _vp, _vp, _vp, _vp, _vp,
                                                 _vp]),
    'fem_bcsr_transpose_structure': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _vp, _vp, _vp]),
    'fem_bcsr_transpose_values': (C.c_int, [_P(FemBcsr), _P(FemBcsr), _vp, _vp]),
    'fem_amg_coarse_inverse': (C.c_int, [_P(FemBcsr), _vp, _vp, _vp]),
    'fem_amg_to_f32': (C.c_int, [_i64, _vp, _vp, _vp]),
    'fem_amg_sell_to_f16_scaled': (C.c_int, [_P(FemPattern), _vp, _vp, _vp, _vp]),
    'fem_amg_ws_bytes': (_i64, [_P(FemAmg)]),
    'fem_amg_vcycle': (C.c_int, [_P(FemAmg), _vp, _vp, _vp, _vp]),
    'fem_amg_solve_ws_bytes': (_i64, [_P(FemPattern), _P(FemAmg)]),
    'fem_amg_pcg_solve': (C.c_int, [_P(FemPattern), _vp, _vp, _P(FemAmg), _vp, _vp, _P(FemAmgOptions), _vp, _vp,
                                    _P(FemPcgResult), _vp]),
}

EXPORTED = tuple(_SIGNATURES)

# include/fem_mi355x.h: device scratch of fem_pcg_solve_noalloc, per GPU
FEM_PCG_NOALLOC_SCRATCH_BYTES = 128 << 20

_LIB = None


class FemError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the shared library and attach signatures (no GPU needed)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise FemError(f"{LIB_NAME} is not built ({path}); run __graft_entry__.build() "
                           "or make -C jax-fem_amd/csrc")
        torch.cuda  # the HIP runtime of torch is the one the library binds to
        lib = C.CDLL(path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def lib():
    """The library, for compute calls: requires a visible GPU."""
    if not torch.cuda.is_available():
        raise FemError("no GPU visible: the device path needs an MI355X (there is no CPU fallback)")
    return load()


def check(rc, what=''):
    if rc != FEM_OK:
        msg = load().fem_last_error().decode(errors='replace')
        raise FemError(f"{what} failed with status {rc}: {msg}")
    return rc


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def stream_ptr():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)
