"""ctypes binding of liboaph.so (include/oaph.h).

The library is the product: there is no fallback.  If liboaph.so is missing
or cannot be loaded, every entry point raises OapLibraryError.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liboaph.so"


class OapLibraryError(RuntimeError):
    pass


class OapError(RuntimeError):
    """Non-zero oaph_status returned by the library."""

    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: status {status}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


class oaph_matrix(C.Structure):
    """math::Matrix layout (oapMatrix/Matrix.h:31-44); planes are device pointers."""

    _fields_ = [
        ("columns", C.c_uint32),
        ("rows", C.c_uint32),
        ("realColumns", C.c_uint32),
        ("realRows", C.c_uint32),
        ("reValues", C.c_void_p),
        ("imValues", C.c_void_p),
    ]


class oaph_matrix_info(C.Structure):
    _fields_ = [("isRe", C.c_int32), ("isIm", C.c_int32), ("columns", C.c_uint32),
                ("rows", C.c_uint32)]


class oaph_arnoldi_stats(C.Structure):
    _fields_ = [("restarts", C.c_uint32), ("matvecs", C.c_uint32), ("matvecs_low", C.c_uint32),
                ("nconv", C.c_uint32), ("seconds", C.c_double), ("max_resid", C.c_double),
                ("fallbacks", C.c_uint32), ("mixed_declined", C.c_uint32),
                ("restarts_perfect_shift", C.c_uint32),
                ("restarts_inverse_iteration", C.c_uint32), ("fp64_polish", C.c_uint32),
                ("matvecs_f16", C.c_uint32), ("check_resid", C.c_double)]


class oaph_matrix_ex(C.Structure):
    """MatrixEx (oapMatrix/MatrixEx.h:26-32): a region of a matrix."""

    _fields_ = [("column", C.c_uint32), ("columns", C.c_uint32), ("row", C.c_uint32),
                ("rows", C.c_uint32)]


MP = C.POINTER(oaph_matrix)
DP = C.POINTER(C.c_double)
U32 = C.c_uint32
DIMS = (C.c_uint32 * 2) * 3
CDIMS = ((C.c_uint32 * 2) * 2) * 2
MEXP = C.POINTER(oaph_matrix_ex)
MULTIPLY_FN = C.CFUNCTYPE(None, MP, MP, C.c_void_p, C.c_void_p, C.c_int)
SORT_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p)
CHECK_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_double, MP, C.c_uint32, C.c_uint32, C.c_void_p)

# name -> (restype, argtypes)
_SIGS = {
    "oaph_status_string": (C.c_char_p, [C.c_int]),
    "oaph_last_error": (C.c_char_p, []),
    "oaph_init": (C.c_int, [C.c_int]),
    "oaph_device_synchronize": (C.c_int, []),
    "oaph_device_count": (C.c_int, []),
    "oaph_matrix_new": (C.c_int, [C.POINTER(MP), C.c_int, C.c_int, U32, U32]),
    "oaph_matrix_free": (C.c_int, [MP]),
    "oaph_matrix_info_get": (C.c_int, [MP, C.POINTER(oaph_matrix_info)]),
    "oaph_matrix_copy_h2d": (C.c_int, [MP, DP, DP]),
    "oaph_matrix_copy_d2h": (C.c_int, [DP, DP, MP]),
    "oaph_matrix_copy_d2d": (C.c_int, [MP, MP]),
    "oaph_matrix_get_value": (C.c_int, [MP, U32, DP, DP]),
    "oaph_matrix_set_value": (C.c_int, [MP, U32, C.c_double, C.c_double]),
    "oaph_matrix_set_zero_row": (C.c_int, [MP, U32, C.c_int, C.c_int]),
    "oaph_matrix_copy_d2d_dims": (C.c_int, [MP, MP, CDIMS]),
    "oaph_matrix_copy_h2d_dims": (C.c_int, [MP, MP, CDIMS]),
    "oaph_matrix_copy_d2h_dims": (C.c_int, [MP, MP, CDIMS]),
    "oaph_matrix_set_submatrix": (C.c_int, [MP, MP, U32, U32, C.c_int]),
    "oaph_api_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "oaph_api_destroy": (C.c_int, [C.c_void_p]),
    "oaph_api_synchronize": (C.c_int, [C.c_void_p]),
    "oaph_api_stream": (C.c_void_p, [C.c_void_p]),
    "oaph_dot_product": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_dot_product_dims": (C.c_int, [C.c_void_p, MP, MP, MP, DIMS]),
    "oaph_dot_product_shared": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_transpose": (C.c_int, [C.c_void_p, MP, MP]),
    "oaph_conjugate_transpose": (C.c_int, [C.c_void_p, MP, MP]),
    "oaph_tensor_product": (C.c_int, [C.c_void_p, MP, MP, MP, DIMS]),
    "oaph_hadamard_product": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_hadamard_product_vec": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_add": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_substract": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_multiply_re_constant": (C.c_int, [C.c_void_p, MP, MP, C.c_double]),
    "oaph_multiply_constant": (C.c_int, [C.c_void_p, MP, MP, C.c_double, C.c_double]),
    "oaph_set_diagonal": (C.c_int, [C.c_void_p, MP, C.c_double, C.c_double]),
    "oaph_set_identity": (C.c_int, [C.c_void_p, MP]),
    "oaph_set_zero_matrix": (C.c_int, [C.c_void_p, MP]),
    "oaph_set_vector": (C.c_int, [C.c_void_p, MP, U32, MP, U32]),
    "oaph_get_vector": (C.c_int, [C.c_void_p, MP, U32, MP, U32]),
    "oaph_magnitude": (C.c_int, [C.c_void_p, DP, MP]),
    "oaph_magnitude2": (C.c_int, [C.c_void_p, DP, MP]),
    "oaph_calc_triangular_h_step": (C.c_int, [C.c_void_p, MP, MP]),
    "oaph_hadamard_product_cr": (C.c_int, [C.c_void_p, MP, MP, MP, U32, U32]),
    "oaph_hadamard_product_vec_cr": (C.c_int, [C.c_void_p, MP, MP, MP, U32, U32]),
    "oaph_tensor_product_cr": (C.c_int, [C.c_void_p, MP, MP, MP, U32, U32]),
    "oaph_add_cr": (C.c_int, [C.c_void_p, MP, MP, MP, U32, U32]),
    "oaph_substract_cr": (C.c_int, [C.c_void_p, MP, MP, MP, U32, U32]),
    "oaph_add_substract": (C.c_int, [C.c_void_p, MP, MP, MP, U32, U32]),
    "oaph_add_dot_product": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_dot_product_periodic": (C.c_int, [C.c_void_p, MP, MP, MP]),
    "oaph_dot_product_dim_periodic": (C.c_int, [C.c_void_p, MP, MP, MP, DIMS, U32]),
    "oaph_dot_product_ex": (C.c_int, [C.c_void_p, MP, MP, MP, MEXP]),
    "oaph_transpose_ex": (C.c_int, [C.c_void_p, MP, MP, MEXP]),
    "oaph_magnitude_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, MP]),
    "oaph_sum": (C.c_int, [C.c_void_p, DP, DP, MP]),
    "oaph_qrgr": (C.c_int, [C.c_void_p, MP, MP, MP, MP, MP, MP, MP]),
    "oaph_is_upper_triangular": (C.c_int, [C.c_void_p, MP, C.POINTER(C.c_int)]),
    "oaph_calc_triangular_h": (C.c_int, [C.c_void_p, MP, MP, MP, MP, MP, MP, MP, MP, MP]),
    "oaph_set_threads_blocks": (C.c_int, [C.POINTER(U32), C.POINTER(U32), U32, U32, U32]),
    "oaph_arnoldi_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "oaph_arnoldi_destroy": (C.c_int, [C.c_void_p]),
    "oaph_arnoldi_api": (C.c_void_p, [C.c_void_p]),
    "oaph_arnoldi_set_rho": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_arnoldi_set_blimit": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_arnoldi_set_qr_type": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_vec_init_type": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_sort_type": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_check_type": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_check_counts": (C.c_int, [C.c_void_p, U32]),
    "oaph_arnoldi_set_check_tolerance": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_arnoldi_set_outputs_eigenvalues": (C.c_int, [C.c_void_p, DP, DP]),
    "oaph_arnoldi_set_outputs_eigenvectors": (C.c_int, [C.c_void_p, C.POINTER(MP)]),
    "oaph_arnoldi_set_output_type": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_calc_triangular_h_type": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_callback": (C.c_int, [C.c_void_p, MULTIPLY_FN, C.c_void_p]),
    "oaph_arnoldi_set_sort_fn": (C.c_int, [C.c_void_p, SORT_FN, C.c_void_p]),
    "oaph_arnoldi_set_check_callback": (C.c_int, [C.c_void_p, CHECK_FN, C.c_void_p]),
    "oaph_arnoldi_set_matrix": (C.c_int, [C.c_void_p, MP]),
    "oaph_arnoldi_execute": (C.c_int, [C.c_void_p, U32, U32, C.POINTER(oaph_matrix_info), C.c_int]),
    "oaph_arnoldi_test_outcome": (C.c_int, [C.c_void_p, U32, DP]),
    "oaph_arnoldi_set_seed": (C.c_int, [C.c_void_p, C.c_uint64]),
    "oaph_arnoldi_set_norm_a": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_arnoldi_set_max_restarts": (C.c_int, [C.c_void_p, U32]),
    "oaph_arnoldi_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_projection": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_mixed_tolerance": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_arnoldi_set_mixed_check": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "oaph_arnoldi_set_f16_phase": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "oaph_arnoldi_apply_operator": (C.c_int, [C.c_void_p, C.c_int, MP, MP]),
    "oaph_arnoldi_set_restart_guard_weight": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_arnoldi_set_restart_keep": (C.c_int, [C.c_void_p, C.c_uint32]),
    "oaph_arnoldi_set_schur_method": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_arnoldi_set_start_vector": (C.c_int, [C.c_void_p, DP, DP, U32]),
    "oaph_arnoldi_get_stats": (C.c_int, [C.c_void_p, C.POINTER(oaph_arnoldi_stats)]),
    "oaph_arnoldi_time_matvec": (C.c_int, [C.c_void_p, C.c_int, U32, DP]),
    "oaph_arnoldi_restart_profile": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "oaph_synth_matrix": (C.c_int, [MP, C.c_int, C.c_uint64, C.c_double, C.c_double]),
    "oaph_set_matvec_shape": (C.c_int, [C.c_int, C.c_int]),
    "oaph_set_matvec_fallback_shape": (C.c_int, [C.c_int, C.c_int]),
    "oaph_set_gemm_path": (C.c_int, [C.c_int]),
    "oaph_matvec_shape_info": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]),
    "oaph_synth_cdr": (C.c_int, [MP, U32, C.c_double, C.c_double, C.c_double]),
    "oaph_time_dot_product": (C.c_int, [C.c_void_p, MP, MP, MP, U32, DP]),
    "oaph_time_reduction": (C.c_int, [C.c_void_p, C.c_int, MP, U32, DP]),
    "oaph_set_reduce_blocks": (C.c_int, [C.c_int]),
    "oaph_activation": (C.c_int, [C.c_void_p, C.c_int, MP, MP]),
    "oaph_activation_dims": (C.c_int, [C.c_void_p, C.c_int, MP, MP, C.POINTER(U32)]),
    "oaph_activation_dims_periodic": (C.c_int, [C.c_void_p, C.c_int, MP, MP, C.POINTER(U32)]),
    "oaph_network_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p]),
    "oaph_network_destroy": (C.c_int, [C.c_void_p]),
    "oaph_network_add_layer": (C.c_int, [C.c_void_p, U32, C.c_int, C.c_int]),
    "oaph_network_set_learning_rate": (C.c_int, [C.c_void_p, C.c_double]),
    "oaph_network_set_host_weights": (C.c_int, [C.c_void_p, U32, DP]),
    "oaph_network_get_host_weights": (C.c_int, [C.c_void_p, U32, DP]),
    "oaph_network_set_inputs": (C.c_int, [C.c_void_p, DP]),
    "oaph_network_set_expected": (C.c_int, [C.c_void_p, DP]),
    "oaph_network_forward": (C.c_int, [C.c_void_p]),
    "oaph_network_get_outputs": (C.c_int, [C.c_void_p, DP]),
    "oaph_network_accumulate_errors": (C.c_int, [C.c_void_p]),
    "oaph_network_backward": (C.c_int, [C.c_void_p]),
    "oaph_network_update_weights": (C.c_int, [C.c_void_p]),
    "oaph_network_post_step": (C.c_int, [C.c_void_p]),
    "oaph_network_calculate_error": (C.c_int, [C.c_void_p, DP]),
    "oaph_mlp_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(U32), U32, U32]),
    "oaph_mlp_destroy": (C.c_int, [C.c_void_p]),
    "oaph_mlp_num_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "oaph_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "oaph_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "oaph_mlp_set_output_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "oaph_mlp_predict": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), U32, C.POINTER(C.c_float)]),
    "oaph_mlp_gradients": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), U32,
                                     C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "oaph_mlp_set_dataset": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       U32]),
    "oaph_mlp_train": (C.c_int, [C.c_void_p, U32, U32, C.c_float, C.c_uint64,
                                 C.POINTER(C.c_float)]),
    "oaph_matrix_to_string": (C.c_int, [MP, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "oaph_host_matrix_to_string": (C.c_int, [U32, U32, DP, DP, C.c_char_p, C.c_size_t,
                                             C.POINTER(C.c_size_t)]),
    "oaph_matrix_parse": (C.c_int, [C.c_char_p, C.POINTER(MP)]),
    "oaph_host_parse_matrix": (C.c_int, [C.c_char_p, C.POINTER(oaph_matrix_info), DP, DP,
                                         C.c_size_t]),
    "oaph_host_parse_array": (C.c_int, [C.c_char_p, U32, DP, C.c_size_t, C.POINTER(C.c_size_t)]),
    "oaph_matrix_save": (C.c_int, [MP, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "oaph_host_matrix_save": (C.c_int, [C.c_int, U32, U32, DP, DP, C.c_void_p, C.c_size_t,
                                        C.POINTER(C.c_size_t)]),
    "oaph_matrix_load": (C.c_int, [C.POINTER(MP), C.c_void_p, C.c_size_t,
                                   C.POINTER(C.c_size_t)]),
    "oaph_host_matrix_load": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t),
                                        C.POINTER(C.c_int), C.POINTER(oaph_matrix_info), DP, DP,
                                        C.c_size_t]),
    "oaph_buffer_write_file": (C.c_int, [C.c_char_p, C.c_void_p, C.c_size_t]),
    "oaph_buffer_read_file": (C.c_int, [C.c_char_p, C.c_void_p, C.c_size_t,
                                        C.POINTER(C.c_size_t)]),
    "oaph_png_read": (C.c_int, [C.c_char_p, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32),
                                C.c_void_p, C.c_size_t]),
    "oaph_png_write": (C.c_int, [C.c_char_p, U32, U32, U32, C.c_void_p]),
    "oaph_image_opt_size": (C.c_int, [C.c_void_p, U32, U32, U32, C.POINTER(U32)]),
    "oaph_dataloader_plan": (C.c_int, [C.POINTER(U32), U32, C.POINTER(U32), C.POINTER(U32)]),
    "oaph_dataloader_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_char_p), U32,
                                         C.c_int]),
    "oaph_dataloader_destroy": (C.c_int, [C.c_void_p]),
    "oaph_dataloader_info": (C.c_int, [C.c_void_p, C.POINTER(oaph_matrix_info)]),
    "oaph_dataloader_window": (C.c_int, [C.c_void_p, U32, C.POINTER(U32), C.POINTER(U32)]),
    "oaph_dataloader_column": (C.c_int, [C.c_void_p, U32, DP]),
    "oaph_dataloader_matrix": (C.c_int, [C.c_void_p, MP]),
    "oaph_arnoldi_set_restart_stages": (C.c_int, [C.c_void_p, DP, U32]),
    "oaph_arnoldi_set_restart_guard": (C.c_int, [C.c_void_p, U32]),
    "oaph_rts_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, MP]),
    "oaph_rts_destroy": (C.c_int, [C.c_void_p]),
    "oaph_rts_info": (C.c_int, [C.c_void_p, C.POINTER(oaph_matrix_info)]),
    "oaph_rts_square": (C.c_int, [C.c_void_p, MP]),
    "oaph_rts_sub_matrix": (C.c_int, [C.c_void_p, U32, U32, MP]),
    "oaph_rts_row_vector": (C.c_int, [C.c_void_p, U32, MP]),
    "oaph_rts_multiply": (C.c_int, [C.c_void_p, C.c_void_p, MP, MP]),
    "oaph_batch_c3_sizes": (C.c_int, [U32, C.c_uint64, C.POINTER(U32)]),
    "oaph_batch_create": (C.c_int, [C.POINTER(C.c_void_p), U32, C.POINTER(U32), U32]),
    "oaph_batch_destroy": (C.c_int, [C.c_void_p]),
    "oaph_batch_upload_x": (C.c_int, [C.c_void_p, U32, C.POINTER(C.c_float)]),
    "oaph_batch_download_x": (C.c_int, [C.c_void_p, U32, C.POINTER(C.c_float)]),
    "oaph_batch_synth_c3": (C.c_int, [C.c_void_p, C.c_uint64]),
    "oaph_batch_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_batch_variant_info": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "oaph_batch_variant_shape": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]),
    "oaph_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "oaph_batch_solve": (C.c_int, [C.c_void_p, U32, U32, C.c_int, C.c_double, U32, C.c_uint64]),
    "oaph_batch_autotune": (C.c_int, [C.c_void_p, U32, U32, C.c_int, C.c_double, U32, C.c_uint64,
                                      U32, C.c_char_p, C.POINTER(C.c_int), DP]),
    "oaph_batch_results": (C.c_int, [C.c_void_p, DP, DP, C.POINTER(U32), C.POINTER(U32),
                                     C.POINTER(U32), DP]),
    "oaph_batch_phase_cycles": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "oaph_batch_eigenvectors": (C.c_int, [C.c_void_p, U32, DP]),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load liboaph.so (once) and attach signatures; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OapLibraryError(
            f"{p} is missing: build it with `python -m oap_amd.build` (there is no CPU fallback)")
    try:
        lib = C.CDLL(str(p))
    except OSError as e:  # pragma: no cover - depends on the environment
        raise OapLibraryError(f"cannot load {p}: {e}") from e
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    return list(_SIGS)


def check(status: int, where: str) -> None:
    if status != 0:
        lib = load()
        detail = lib.oaph_status_string(status).decode()
        last = lib.oaph_last_error().decode()
        if last:
            detail += "; " + last
        raise OapError(status, where, detail)
