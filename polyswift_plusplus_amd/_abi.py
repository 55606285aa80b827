"""ctypes binding of include/psx_scft.h (libpsx_scft.so, built in-tree).

Every exported symbol of the header is listed in ``EXPORTS``; loading fails
loudly if the library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsx_scft.so")

PSX_FP64 = 0
PSX_FP32 = 1

STATUS_NAMES = {
    0: "PSX_OK", 1: "PSX_ERR_INVALID", 2: "PSX_ERR_SHAPE", 3: "PSX_ERR_STEPS",
    4: "PSX_ERR_NAN", 5: "PSX_ERR_HIP", 6: "PSX_ERR_NOMEM", 7: "PSX_ERR_STATE",
}

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class Diblock(ctypes.Structure):
    _fields_ = [("fA", ctypes.c_double), ("ds", ctypes.c_double), ("bA", ctypes.c_double),
                ("bB", ctypes.c_double), ("chiN", ctypes.c_double)]


PSX_MAX_BLOCKS = 16
PSX_MAX_JOIN = 8


class BlockDesc(ctypes.Structure):
    _fields_ = [("species", ctypes.c_int), ("lengthfrac", ctypes.c_double),
                ("bSegRatio", ctypes.c_double), ("nhead", ctypes.c_int), ("ntail", ctypes.c_int),
                ("head", ctypes.c_int * PSX_MAX_JOIN), ("tail", ctypes.c_int * PSX_MAX_JOIN)]


PSX_MAX_POLYMERS = 8


class PolymerDesc(ctypes.Structure):
    _fields_ = [("volfrac", ctypes.c_double), ("length", ctypes.c_double), ("nblocks", ctypes.c_int)]


class Mixing(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("lambda0", ctypes.c_double),
                ("lambda1", ctypes.c_double), ("depth", ctypes.c_int),
                ("beta", ctypes.c_double), ("anderson_start", ctypes.c_double),
                ("f_order", ctypes.c_int)]


# name -> (restype, argtypes)
EXPORTS = {
    "psx_last_error": (None, [ctypes.c_char_p, ctypes.c_size_t]),
    "psx_version": (ctypes.c_char_p, []),
    "psx_fft_plan_create": (ctypes.c_int, [ctypes.c_int] * 3 + [ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "psx_fft_plan_destroy": (ctypes.c_int, [_vp]),
    "psx_fft_size": (ctypes.c_size_t, [_vp]),
    "psx_fft_scaled_pair": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "psx_fft_scaled_pair_host": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "psx_fft_scaled_pair_noalloc": (ctypes.c_int, [ctypes.c_int] * 4 + [_vp, _vp, _vp, _vp]),
    "psx_fft_forward_abs": (ctypes.c_int, [_vp, _vp, _vp]),
    "psx_fft_convolve_re": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "psx_fft_scaled_pair_im": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "psx_fft_forward_re": (ctypes.c_int, [_vp, _vp, _vp]),
    "psx_fft_backward_re": (ctypes.c_int, [_vp, _vp, _vp]),
    "psx_fft_spec_filter": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, _ip,
                                           ctypes.c_double]),
    "psx_fft_poisson_update": (ctypes.c_int, [_vp, _vp, _vp] + [ctypes.c_double] * 4),
    "psx_fft_synchronize": (ctypes.c_int, [_vp]),
    "psx_buffer_create": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "psx_buffer_destroy": (ctypes.c_int, [_vp]),
    "psx_buffer_upload": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "psx_buffer_download": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "psx_probe_stream": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, _dp]),
    "psx_scft_create": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Diblock), ctypes.c_int,
                                       _vp, ctypes.POINTER(_vp)]),
    "psx_scft_create_blocks": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.c_int, ctypes.POINTER(BlockDesc),
                                              ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp,
                                              ctypes.POINTER(_vp)]),
    "psx_scft_create_blend": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.c_int, ctypes.POINTER(PolymerDesc),
                                             ctypes.POINTER(BlockDesc), ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "psx_scft_destroy": (ctypes.c_int, [_vp]),
    "psx_scft_graph_info": (ctypes.c_int, [_vp, _ip, _ip, _ip]),
    "psx_scft_get_block_slice_host": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "psx_scft_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "psx_scft_block_steps": (ctypes.c_int, [_vp, _ip, _ip]),
    "psx_scft_set_fields": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "psx_scft_get_fields": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "psx_scft_set_fields_host": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "psx_scft_get_fields_host": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "psx_scft_spec_filter": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip]),
    "psx_rng_create": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.POINTER(_vp)]),
    "psx_rng_destroy": (ctypes.c_int, [_vp]),
    "psx_rng_uniform": (ctypes.c_int, [_vp, ctypes.c_longlong, _dp]),
    "psx_rng_gaussian": (ctypes.c_int, [_vp, ctypes.c_longlong, _dp]),
    "psx_scft_set_noise": (ctypes.c_int, [_vp, ctypes.c_double, _vp]),
    "psx_scft_propagate": (ctypes.c_int, [_vp]),
    "psx_scft_Q": (ctypes.c_int, [_vp, _dp]),
    "psx_scft_polymer_Q": (ctypes.c_int, [_vp, _dp, _ip]),
    "psx_scft_densities": (ctypes.c_int, [_vp, _vp, _vp]),
    "psx_scft_get_densities_host": (ctypes.c_int, [_vp, _dp, _dp]),
    "psx_scft_get_slice_host": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "psx_scft_update_steepest": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _dp]),
    "psx_scft_iterate": (ctypes.c_int, [_vp, ctypes.POINTER(Mixing), ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, _dp, _dp, _dp, _ip]),
    "psx_scft_reset_history": (ctypes.c_int, [_vp]),
    "psx_scft_free_energy": (ctypes.c_int, [_vp, _dp]),
    "psx_scft_box_stress": (ctypes.c_int, [_vp, _dp]),
    "psx_scft_backward": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _vp, _vp, _dp]),
    "psx_scft_backward_host": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]),
    "psx_scft_set_quadrature": (ctypes.c_int, [_vp, ctypes.c_int]),
    "psx_scft_set_quad_kernel": (ctypes.c_int, [_vp, ctypes.c_int]),
    "psx_scft_set_checkpointing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "psx_scft_set_cell": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "psx_scft_param_grad": (ctypes.c_int, [_vp, _dp, _dp]),
    "psx_scft_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "psx_scft_set_variant": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "psx_scft_set_sweep_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "psx_scft_sweep_mode": (ctypes.c_int, [_vp, _ip, _ip]),
    "psx_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Grid), ctypes.POINTER(Diblock),
                                        ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "psx_batch_destroy": (ctypes.c_int, [_vp]),
    "psx_batch_size": (ctypes.c_int, [_vp, _ip, _ip]),
    "psx_batch_instance": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "psx_batch_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "psx_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "psx_batch_set_sweep_mode": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "psx_batch_sweep_mode": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "psx_batch_iterate": (ctypes.c_int, [_vp, ctypes.POINTER(Mixing), ctypes.c_int, ctypes.c_int,
                                         _dp, _dp, _ip]),
    "psx_scft_time_step": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
}

_lib = None


class PsxError(RuntimeError):
    """A non-zero psx_status, with the library's thread-local message."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def load(path: str = LIB_PATH):
    """Load libpsx_scft.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m polyswift_plusplus_amd.build_ext` "
            "(the engine has no CPU fallback)")
    L = ctypes.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().psx_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != 0:
        raise PsxError(status, last_error())
