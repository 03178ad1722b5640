"""ctypes binding of libmchip.so (the C ABI declared in include/mchip.h).

The library is built in-tree by chem_amd.build / __graft_entry__.build().
Loading fails loudly when it is missing: there is no Python fallback for any
entry point.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmchip.so")

_lib = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_int64 = ctypes.c_int64
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P_int = ctypes.POINTER(ctypes.c_int)
P_double = ctypes.POINTER(ctypes.c_double)
P_u8 = ctypes.POINTER(ctypes.c_uint8)
P_float = ctypes.POINTER(ctypes.c_float)
P_int64 = ctypes.POINTER(ctypes.c_int64)


class MchipError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MchipError(
                "libmchip.so not found at %s; run __graft_entry__.build() first" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


def check(status):
    if status != 0:
        msg = lib().mchip_last_error().decode("utf-8", "replace")
        raise MchipError(msg)


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_double)


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(P_float)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(P_int)


def u8ptr(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(P_u8)


# (name, restype, argtypes); every entry point returns an int status unless noted
_SIGS = [
    ("mchip_last_error", c_char_p, []),
    ("mchip_basis_create", c_int, [c_int, P_int, P_int, P_double, P_double, P_double, P_int,
                                   ctypes.POINTER(c_void_p)]),
    ("mchip_basis_destroy", c_int, [c_void_p]),
    ("mchip_basis_nbf", c_int, [c_void_p, P_int]),
    ("mchip_basis_norm_coefs", c_int, [c_void_p, P_double]),
    ("mchip_boys", c_int, [c_int, c_double, P_double]),
    ("mchip_int1e", c_int, [c_void_p, c_int, P_double, P_double, P_double, P_double, P_double]),
    ("mchip_int2c2e", c_int, [c_void_p, P_double]),
    ("mchip_int1e_gpu", c_int, [c_void_p, c_void_p, c_int, P_double, P_double, P_double, P_double, P_double]),
    ("mchip_int2c2e_gpu", c_int, [c_void_p, c_void_p, P_double]),
    ("mchip_schwarz_ao_gpu", c_int, [c_void_p, c_void_p, P_double]),
    ("mchip_schwarz", c_int, [c_void_p, c_void_p, P_double, P_double]),
    ("mchip_int3c2e_shell", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, P_double]),
    ("mchip_btensor_build", c_int, [c_void_p, c_void_p, c_int, P_int, c_int, P_int, P_u8,
                                    c_double, ctypes.POINTER(c_void_p)]),
    ("mchip_btensor_from_dense", c_int, [c_int, P_int, c_int, P_int, P_u8, P_double,
                                         ctypes.POINTER(c_void_p)]),
    ("mchip_btensor_destroy", c_int, [c_void_p]),
    ("mchip_btensor_info", c_int, [c_void_p, P_int, P_int64]),
    ("mchip_btensor_mask", c_int, [c_void_p, P_u8]),
    ("mchip_btensor_to_dense", c_int, [c_void_p, P_double]),
    ("mchip_btensor_dense_rows", c_int, [c_void_p, c_int, c_int, P_double]),
    ("mchip_laplace_minimax", c_int, [c_int, c_double, c_double, P_double, P_double, c_int,
                                      P_double, P_double, P_double, P_double]),
    ("mchip_device_count", c_int, [P_int]),
    ("mchip_ctx_create", c_int, [c_int, ctypes.POINTER(c_void_p)]),
    ("mchip_ctx_destroy", c_int, [c_void_p]),
    ("mchip_ctx_sync", c_int, [c_void_p]),
    ("mchip_ctx_stream", c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    ("mchip_dbt_upload", c_int, [c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
    ("mchip_dbt_destroy", c_int, [c_void_p]),
    ("mchip_dbt_build_gpu", c_int, [c_void_p, c_void_p, c_void_p, c_int, P_int, c_int, P_int, P_u8,
                                    c_double, ctypes.POINTER(c_void_p)]),
    ("mchip_dbt_info", c_int, [c_void_p, P_int, P_int64]),
    ("mchip_dbt_mask", c_int, [c_void_p, P_u8]),
    ("mchip_dbt_dense_rows", c_int, [c_void_p, c_int, c_int, P_double]),
    ("mchip_dbt_download", c_int, [c_void_p, P_double]),
    ("mchip_metric_create", c_int, [c_void_p, c_int, P_double, ctypes.POINTER(c_void_p)]),
    ("mchip_metric_destroy", c_int, [c_void_p]),
    ("mchip_metric_get", c_int, [c_void_p, c_void_p, P_double, P_double]),
    ("mchip_dbt_transform", c_int, [c_void_p, c_void_p]),
    ("mchip_dfj_dev", c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    ("mchip_dfj", c_int, [c_void_p, c_void_p, P_double, c_double, P_double]),
    ("mchip_dfj_dev_noalloc", c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    ("mchip_dfk_mo_dev", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    ("mchip_dfk_mo", c_int, [c_void_p, c_void_p, P_double, c_int, c_int, P_double]),
    ("mchip_scf_run", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, P_int, P_double, P_double,
                              c_double, c_double, c_int, c_int, c_int, c_int, c_int, c_double,
                              c_int, c_int, P_double, P_int, P_int, P_double, P_double, c_int,
                              P_double, P_double, P_double, P_double, P_int]),
    ("mchip_mp2_sos", c_int, [c_void_p, c_void_p, c_int, c_int, P_double, P_double, P_double, c_int,
                              P_double, P_double, c_int, P_int, c_int, c_double, P_double, P_double,
                              P_double]),
    ("mchip_mp2_sos_batched", c_int, [c_void_p, c_void_p, c_int, c_int, P_double, P_double, P_double, c_int,
                              P_double, P_double, c_int, P_int, c_int, c_double, P_double, P_double,
                              P_double, P_double]),
    ("mchip_mp2_sos_z", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, P_double, P_double, P_double, c_int,
                                P_double, P_double, c_int, P_int, c_int, c_double, P_double, P_double, P_double]),
    ("mchip_z_compute", c_int, [c_void_p, c_int, c_int, P_double, c_int, P_double, P_double, c_int, P_double]),
    ("mchip_mp2_z", c_int, [c_void_p, c_void_p, c_int, c_int, P_double, P_double, P_double, c_int,
                            P_double, P_double, c_int, P_int, c_int, c_double, c_int, P_double, P_double]),
    ("mchip_ctx_set_timing", c_int, [c_void_p, c_int]),
    ("mchip_ctx_set_escalation", c_int, [c_void_p, c_double, c_int]),
    ("mchip_ctx_trim", c_int, [c_void_p]),
    ("mchip_ctx_set_guess", c_int, [c_void_p, c_int, P_double]),
    ("mchip_server_set_batching", c_int, [c_void_p, c_int]),
    ("mchip_dfk_energy_tilde64_dev", c_int, [c_void_p, c_void_p, c_int, c_int, P_double]),
    ("mchip_dfk_energy_dev", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, P_double]),
    ("mchip_server_batch_stats", c_int, [c_void_p, P_double]),
    ("mchip_dfk_mo_bf16_group_dev", c_int, [c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), P_int,
                                            P_int, ctypes.POINTER(c_void_p)]),
    ("mchip_dev_alloc", c_int, [c_void_p, ctypes.c_size_t, ctypes.POINTER(c_void_p)]),
    ("mchip_dev_free", c_int, [c_void_p, c_void_p]),
    ("mchip_dev_copy_in", c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t]),
    ("mchip_dev_copy_out", c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t]),
    ("mchip_dbt_make_f32", c_int, [c_void_p]),
    ("mchip_dfk_mo_f32_dev", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    ("mchip_dfk_mo_f32", c_int, [c_void_p, P_double, c_int, c_int, P_double]),
    ("mchip_dbt_make_bf16", c_int, [c_void_p]),
    ("mchip_dbt_make_lowp", c_int, [c_void_p, c_void_p, c_int, c_int, c_int]),
    ("mchip_dfk_mo_tilde64", c_int, [c_void_p, P_double, c_int, c_int, P_double]),
    ("mchip_dfk_mo_bf16_dev", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    ("mchip_dfk_mo_tilde64_dev", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    ("mchip_dfk_mo_bf16", c_int, [c_void_p, P_double, c_int, c_int, P_double]),
    ("mchip_dfk_bf16_tile", c_int, [c_void_p, c_int, P_int]),
    ("mchip_dfk_bf16_plan", c_int, [c_void_p, c_int, P_int, P_float, P_int]),
    ("mchip_dfk_f32_plan", c_int, [c_void_p, c_int, P_int, P_float, P_int]),
    ("mchip_dfk_f32_force", c_int, [c_void_p, c_int, c_int, c_int]),
    ("mchip_dfk_f32_shapes", c_int, [P_int, P_int]),
    ("mchip_dfk_f64_plan", c_int, [c_void_p, c_int, P_int, P_float, P_int]),
    ("mchip_dfk_f64_force", c_int, [c_void_p, c_int, c_int, c_int]),
    ("mchip_dfk_f64_shapes", c_int, [P_int, P_int]),
    ("mchip_dfk_bf16_force", c_int, [c_void_p, c_int, c_int, c_int]),
    ("mchip_dfk_bf16_shapes", c_int, [P_int, P_int, P_int]),
    ("mchip_dfk_ao_dev", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    ("mchip_dfk_ao", c_int, [c_void_p, c_void_p, P_double, c_int, c_int, P_double]),
    ("mchip_int4c2e", c_int, [c_void_p, P_double]),
    ("mchip_sad_guess", c_int, [c_void_p, c_int, P_double, P_double, P_double]),
    ("mchip_sad_atom", c_int, [c_void_p, c_double, P_double, c_double, P_double, P_double, P_int, P_int,
                               P_double]),
    ("mchip_schwarz_mask", c_int, [c_void_p, c_void_p, c_int, P_int, c_int, P_int, c_double, P_u8, P_double,
                                   P_double]),
    ("mchip_dfk_lmo_uv", c_int, [c_void_p, c_void_p, P_double, P_double, c_int, c_int, c_int, P_double]),
    ("mchip_server_create", c_int, [c_int, c_double, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("mchip_server_destroy", c_int, [c_void_p]),
    ("mchip_server_submit", c_int, [c_void_p, c_void_p, c_void_p, c_int, P_double, P_double, c_int, P_int,
                                    c_int, P_int, c_int, c_double, c_double, c_int, c_double, P_int]),
    ("mchip_server_result", c_int, [c_void_p, c_int, c_int, P_int, P_double, P_int, P_int, P_double]),
    ("mchip_server_stats", c_int, [c_void_p, P_double]),
    ("mchip_scf_last_extra", c_int, [P_double, P_double]),
    ("mchip_ctx_phase_times", c_int, [c_void_p, P_double, c_int]),
    ("mchip_ctx_set_graphs", c_int, [c_void_p, c_int]),
    ("mchip_dbt_graph_stats", c_int, [c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
]


def _declare(L):
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def exported_symbols():
    return [s[0] for s in _SIGS]
