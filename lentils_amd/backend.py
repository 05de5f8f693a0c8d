"""ctypes binding of liblentils_hip.so -- the HIP backend of this package.

Mirrors lentils/backend/load_backend.py: the same by-value struct layouts
(common.h / deflect.h) and the same entry-point names for the drop-in
symbols, plus the lt_* device-resident API.  The library is loaded eagerly,
like the reference's four backends (load_backend.py:59,68,93,122); there is
no CPU fallback: if the library is missing or no GPU is visible, calls fail
with an error.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, 'liblentils_hip.so')

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)
c_bool_p = C.POINTER(C.c_bool)
c_void_p = C.c_void_p


class image_space_ctype(C.Structure):
    _fields_ = [('nx', C.c_int), ('ny', C.c_int), ('num_channels', C.c_int), ('num_stokes', C.c_int),
                ('xmin', C.c_double), ('xmax', C.c_double), ('ymin', C.c_double), ('ymax', C.c_double),
                ('dx', C.c_double), ('dy', C.c_double), ('_c_channels', c_double_p), ('_c_mask', c_bool_p)]


class delaunay_space_ctype(C.Structure):
    """lentils/backend/common.h:25-36 (load_backend.py:33-35)."""
    _fields_ = [('num_points', C.c_int), ('num_triangles', C.c_int), ('num_channels', C.c_int),
                ('num_stokes', C.c_int), ('_c_channels', c_double_p), ('_c_points', c_double_p),
                ('_c_triangles', c_int_p)]


class fourier_space_ctype(C.Structure):
    _fields_ = [('nu', C.c_int), ('nv', C.c_int), ('half_nv', C.c_int), ('num_channels', C.c_int),
                ('num_stokes', C.c_int), ('du', C.c_double), ('dv', C.c_double), ('gcx', C.c_double),
                ('gcy', C.c_double), ('_c_channels', c_double_p)]


class visibility_space_ctype(C.Structure):
    _fields_ = [('num_rows', C.c_int), ('num_channels', C.c_int), ('num_stokes', C.c_int),
                ('_c_uvw', c_double_p), ('_c_channels', c_double_p)]


class global_lens_model_ctype(C.Structure):
    _fields_ = [('num_lenses', C.c_int), ('num_zplanes', C.c_int), ('_c_lenses', c_void_p)]


generic_mass_model_ctype = np.dtype([('type', 'u4'), ('_pad', 'u4'), ('z_l', 'f8'), ('z_s', 'f8'), ('d_l', 'f8'),
                                     ('d_s', 'f8'), ('d_ls', 'f8'), ('sigma_c', 'f8'), ('beta', 'f8'),
                                     ('fpars', 'f8', 24), ('ipars', 'i4', 8), ('flags', 'b', 32)])

LT_OK = 0
LT_ERR_ARG, LT_ERR_HIP, LT_ERR_NYQUIST, LT_ERR_UNSUPPORTED, LT_ERR_CAPACITY, LT_ERR_FFT, LT_ERR_NOT_CONVERGED = 1, 2, 3, 4, 5, 6, 7
LT_F64, LT_F32 = 0, 1
LT_OP_NUFFT, LT_OP_BCB, LT_OP_NUFFT_EXPLICIT = 0, 1, 2
LT_PC_NONE, LT_PC_JACOBI, LT_PC_LEARNED, LT_PC_LENS = 0, 1, 2, 3
RASTER_STACK_OVERFLOW, RASTER_FRAG_OVERFLOW, RASTER_POP_LIMIT, RASTER_BAD_LENS = 1, 2, 4, 8


class BackendError(RuntimeError):
    pass


_lib = None


def lib():
    """Load liblentils_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise BackendError(f'{LIBPATH} is missing: build it with python -m lentils_amd.build')
    L = C.CDLL(LIBPATH)
    vp, i, d, l = C.c_void_p, C.c_int, C.c_double, C.c_long
    ip = C.POINTER(C.c_int)
    lp = C.POINTER(C.c_long)
    dp = C.POINTER(C.c_double)
    sig = {
        'lt_init': ([i], i),
        'lt_last_error_message': ([], C.c_char_p),
        'lt_version': ([], C.c_char_p),
        'deflect_points': ([global_lens_model_ctype, vp, i, vp, i, vp], None),
        'manifold_lens_matrix_csr': ([image_space_ctype, image_space_ctype, global_lens_model_ctype, vp, vp, vp],
                                     None),
        'grid_cpu': ([visibility_space_ctype, vp, fourier_space_ctype, vp, i, d, i], None),
        'dft_matrix_csr': ([visibility_space_ctype, image_space_ctype, vp, vp, vp], None),
        'curv_matrix_findiff_csr': ([image_space_ctype, vp, vp, vp], None),
        'grad_matrix_findiff_csr': ([image_space_ctype, vp, vp, vp], None),
        'triangle_gradient_csr': ([delaunay_space_ctype, vp, vp, vp], None),
        'delaunay_lens_matrix_csr': ([image_space_ctype, delaunay_space_ctype, i, vp, vp, vp, vp, vp], None),
        'triangle_geometry_from_inds': ([vp, vp, vp, i, vp], None),
        'lt_triangle_gradient_csr': ([delaunay_space_ctype, vp, vp, vp], i),
        'lt_delaunay_lens_matrix_csr': ([image_space_ctype, delaunay_space_ctype, i, vp, vp, vp, vp, vp, l, lp], i),
        'lt_deflect_points': ([C.POINTER(global_lens_model_ctype), vp, l, vp], i),
        'lt_deflect_points_deriv': ([C.POINTER(global_lens_model_ctype), vp, l, vp, vp], i),
        'lt_lens_matrix_nnz': ([image_space_ctype, image_space_ctype, global_lens_model_ctype, lp, ip], i),
        'lt_lens_matrix_csr': ([image_space_ctype, image_space_ctype, global_lens_model_ctype, vp, vp, vp, l, lp,
                                ip], i),
        'lt_deflect_lattice': ([image_space_ctype, global_lens_model_ctype, vp], i),
        'lt_lens_matrix_from_lattice': ([image_space_ctype, image_space_ctype, vp, vp, vp, vp, l, lp, ip], i),
        'lt_lensbatch_create': ([image_space_ctype, vp, i, vp, vp, C.POINTER(vp)], i),
        'lt_lensbatch_info': ([vp, i, lp, ip], i),
        'lt_lensbatch_download': ([vp, i, vp, vp, vp], i),
        'lt_lensbatch_destroy': ([vp], i),
        'lt_nufft_create': ([visibility_space_ctype, image_space_ctype, i, i, d, i, C.POINTER(vp)], i),
        'lt_nufft_forward_host': ([vp, vp, vp], i),
        'lt_nufft_adjoint_host': ([vp, vp, vp], i),
        'lt_nufft_degrid_host': ([vp, vp, vp], i),
        'lt_nufft_grid_host': ([vp, vp, vp], i),
        'lt_nufft_info': ([vp, ip, ip, ip, dp, lp], i),
        'lt_nufft_time': ([vp, i, i, dp], i),
        'lt_kb_taps': ([d, i, i, vp, dp], i),
        'lt_kb_taps_window': ([d, i, i, vp, dp], i),
        'lt_dirty_beam_host': ([visibility_space_ctype, image_space_ctype, vp, i, d, vp, dp], i),
        'lt_bcb_apply_host': ([image_space_ctype, vp, i, vp, vp], i),
        'lt_nufft_destroy': ([vp], i),
        'lt_nufft_set_weights': ([vp, vp], i),
        'lt_nufft_set_stencil_shape': ([vp, i], i),
        'lt_nufft_set_degrid_form': ([vp, i], i),
        'lt_nufft_set_row_form': ([vp, i], i),
        'lt_nufft_set_col_form': ([vp, i], i),
        'lt_nufft_stencil_time': ([vp, i, i, dp, dp, dp, lp, lp], i),
        'lt_nufft_stencil_form': ([vp, ip], i),
        'lt_device_alloc': ([l, C.POINTER(vp)], i),
        'lt_device_upload': ([vp, vp, l], i),
        'lt_device_download': ([vp, vp, l], i),
        'lt_device_free': ([vp], i),
        'lt_solver_create_resident': ([vp, vp, vp, i, d, d, i, C.POINTER(vp)], i),
        'lt_solver_create': ([vp, vp, vp, vp, d, d, i, C.POINTER(vp)], i),
        'lt_solver_create_dirty': ([vp, vp, vp, vp, i, d, d, i, C.POINTER(vp)], i),
        'lt_pcnet_param_count': ([], l),
        'lt_pcnet_create': ([i, i, i, vp, C.c_uint, C.POINTER(vp)], i),
        'lt_pcnet_get_params': ([vp, vp], i),
        'lt_pcnet_set_params': ([vp, vp], i),
        'lt_pcnet_forward_host': ([vp, vp, i, vp], i),
        'lt_pcnet_grad_host': ([vp, vp, vp, i, vp, vp], i),
        'lt_pcnet_train': ([vp, vp, vp, l, i, i, C.c_float, C.c_float, C.c_uint, vp], i),
        'lt_pcnet_eval': ([vp, vp, vp, l, dp], i),
        'lt_pcnet_destroy': ([vp], i),
        'lt_solver_set_pcnet': ([vp, vp], i),
        'lt_solver_diag': ([vp, vp], i),
        'lt_solver_train_pcnet': ([vp, vp, i, i, C.c_float, C.c_float, C.c_uint, vp], i),
        'lt_solver_pc_fallbacks': ([vp, vp, vp], i),
        'lt_solver_solve': ([vp, d, i, i, vp, vp, vp], i),
        'lt_solver_rhs': ([vp, vp], i),
        'lt_solver_set_rhs': ([vp, vp], i),
        'lt_solver_set_precision': ([vp, i, i], i),
        'lt_solver_set_lens_pc': ([vp, i, i, d], i),
        'lt_solver_set_graphs': ([vp, i], i),
        'lt_solver_set_flexible_depth': ([vp, i], i),
        'lt_solver_time_products': ([vp, i, dp, dp, lp, lp], i),
        'lt_solver_set_learned_iterations': ([vp, i], i),
        'lt_solver_lens_pc_matrix': ([vp, vp, vp, vp, vp, vp], i),
        'lt_solver_apply': ([vp, vp, vp], i),
        'lt_solver_destroy': ([vp], i),
        'lt_fft_r2c_host': ([i, i, i, vp, vp], i),
        'lt_fft_c2r_host': ([i, i, i, vp, vp], i),
        'lt_csr_spmv_host': ([l, l, vp, vp, vp, vp, vp], i),
        'lt_csr_spmv_host_kernel': ([l, l, vp, vp, vp, vp, vp, i], i),
        'lt_csr_spmv_device_noalloc': ([l, l, l, vp, vp, vp, vp, vp], i),
        'lt_stream_copy': ([l, i, dp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names every build of the library must export (include/lentils_hip.h)."""
    return ['lt_init', 'lt_last_error_message', 'lt_version', 'deflect_points', 'manifold_lens_matrix_csr',
            'grid_cpu', 'dft_matrix_csr', 'curv_matrix_findiff_csr', 'grad_matrix_findiff_csr', 'triangle_gradient_csr', 'delaunay_lens_matrix_csr',
            'triangle_geometry_from_inds', 'lt_triangle_gradient_csr', 'lt_delaunay_lens_matrix_csr',
            'lt_deflect_points', 'lt_deflect_points_deriv',
            'lt_lens_matrix_nnz', 'lt_lens_matrix_csr', 'lt_deflect_lattice', 'lt_lens_matrix_from_lattice',
            'lt_lensbatch_create', 'lt_lensbatch_info', 'lt_lensbatch_download', 'lt_lensbatch_destroy',
            'lt_nufft_create', 'lt_nufft_forward_host', 'lt_nufft_adjoint_host', 'lt_nufft_degrid_host',
            'lt_nufft_grid_host', 'lt_nufft_info', 'lt_nufft_time', 'lt_nufft_destroy', 'lt_nufft_set_weights',
            'lt_nufft_set_stencil_shape', 'lt_nufft_set_degrid_form', 'lt_nufft_set_row_form', 'lt_nufft_set_col_form',
            'lt_nufft_stencil_time', 'lt_nufft_stencil_form',
            'lt_device_alloc', 'lt_device_upload', 'lt_device_download', 'lt_device_free', 'lt_solver_create_resident',
            'lt_kb_taps', 'lt_kb_taps_window', 'lt_dirty_beam_host', 'lt_bcb_apply_host',
            'lt_solver_create', 'lt_solver_create_dirty', 'lt_pcnet_param_count', 'lt_pcnet_create',
            'lt_pcnet_get_params', 'lt_pcnet_set_params', 'lt_pcnet_forward_host', 'lt_pcnet_grad_host',
            'lt_pcnet_train', 'lt_pcnet_eval', 'lt_pcnet_destroy', 'lt_solver_set_pcnet', 'lt_solver_diag', 'lt_solver_train_pcnet',
            'lt_solver_pc_fallbacks', 'lt_solver_solve', 'lt_solver_rhs', 'lt_solver_set_rhs', 'lt_solver_set_precision', 'lt_solver_set_lens_pc', 'lt_solver_set_graphs', 'lt_solver_set_flexible_depth', 'lt_solver_set_learned_iterations', 'lt_solver_time_products', 'lt_solver_lens_pc_matrix', 'lt_solver_apply', 'lt_solver_destroy', 'lt_fft_r2c_host',
            'lt_fft_c2r_host', 'lt_csr_spmv_host', 'lt_csr_spmv_host_kernel', 'lt_csr_spmv_device_noalloc', 'lt_stream_copy']


def check(status):
    """Raise BackendError with the library's message on a nonzero status."""
    if status != LT_OK:
        msg = lib().lt_last_error_message()
        raise BackendError(f'lentils_hip error {status}: {msg.decode() if msg else ""}')


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)
