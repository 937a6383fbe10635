"""Loader for libkymatio_hip.so (the C ABI in include/kymatio_hip.h).

The library is built in-tree by ``kymatio_amd/csrc/build.py``.  There is no
CPU fallback: if the library is missing or no HIP device is present, every
call raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libkymatio_hip.so')

_lib = None

STATUS_EXC = {1: TypeError, 2: RuntimeError, 3: ValueError, 4: RuntimeError, 5: NotImplementedError}

_c_ll = ctypes.c_longlong
_c_vp = ctypes.c_void_p
_c_i = ctypes.c_int

# name -> argtypes; every function returns int status (except noted)
_SIGNATURES = {
    'kym_version': [],
    'kym_scat2d_create': [_c_i, _c_i, _c_i, _c_i, _c_i, _c_i, _c_i, _c_ll, ctypes.POINTER(_c_vp)],
    'kym_scat2d_destroy': [_c_vp],
    'kym_scat2d_info': [_c_vp, ctypes.POINTER(_c_ll)],
    'kym_scat2d_filter_offsets': [_c_vp, ctypes.POINTER(_c_ll)],
    'kym_scat2d_leaf_taps': [_c_vp, ctypes.POINTER(_c_ll), _c_i],
    'kym_scat2d_create_tuned': [_c_i] * 7 + [_c_ll, ctypes.POINTER(_c_i), _c_i, ctypes.POINTER(_c_vp)],
    'kym_scat2d_graph_create': [ctypes.POINTER(_c_vp), _c_i] + [ctypes.POINTER(_c_vp)] * 3
                               + [ctypes.POINTER(_c_ll), ctypes.POINTER(_c_vp), ctypes.POINTER(_c_ll),
                                  ctypes.POINTER(_c_vp)],
    'kym_scat2d_graph_create_counted': [ctypes.POINTER(_c_vp), _c_i] + [ctypes.POINTER(_c_vp)] * 3
                                       + [ctypes.POINTER(_c_ll), ctypes.POINTER(_c_vp), ctypes.POINTER(_c_ll), _c_vp,
                                          ctypes.POINTER(_c_vp)],
    'kym_scat2d_graph_launch': [_c_vp, _c_vp],
    'kym_scat2d_graph_destroy': [_c_vp],
    'kym_scat2d_forward': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp, _c_ll, _c_vp],
    'kym_scat2d_forward_noalloc': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp],
    'kym_scat2d_forward_train': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp, _c_ll, _c_vp],
    'kym_scat2d_backward': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp, _c_ll, _c_vp],
    'kym_scat2d_launch_info': [_c_vp, _c_i, ctypes.POINTER(_c_ll)],
    'kym_scat2d_profile': [_c_vp, _c_i],
    'kym_scat2d_profile_read': [_c_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_c_ll), _c_i],
    'kym_scat1d_create': [_c_i] * 8 + [ctypes.POINTER(_c_i)] * 2 + [_c_i] + [ctypes.POINTER(_c_i)] * 2
                         + [_c_i] + [ctypes.POINTER(_c_i)] * 2 + [_c_i, _c_i, _c_ll, ctypes.POINTER(_c_vp)],
    'kym_scat1d_destroy': [_c_vp],
    'kym_scat1d_info': [_c_vp, ctypes.POINTER(_c_ll)],
    'kym_scat1d_layout': [_c_vp, ctypes.POINTER(_c_ll), ctypes.POINTER(_c_ll)],
    'kym_scat1d_forward': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp, _c_ll, _c_vp],
    'kym_scat1d_set_streams': [_c_vp, _c_i],
    'kym_scat1d_forward_train': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp, _c_ll, _c_vp],
    'kym_scat1d_backward': [_c_vp, _c_vp, _c_vp, _c_vp, _c_ll, _c_vp, _c_ll, _c_vp],
    'kym_op_pad2d': [_c_vp, _c_vp, _c_ll] + [_c_i] * 7 + [_c_vp],
    'kym_op_pad1d': [_c_vp, _c_vp, _c_ll, _c_i, _c_i, _c_i, _c_i, _c_vp],
    'kym_op_cdgmm': [_c_vp, _c_vp, _c_vp, _c_ll, _c_ll, _c_i, _c_i, _c_vp],
    'kym_op_modulus': [_c_vp, _c_vp, _c_ll, _c_i, _c_vp],
    'kym_op_subsample': [_c_vp, _c_vp, _c_ll, _c_i, _c_i, _c_i, _c_i, _c_i, _c_vp],
    'kym_op_unpad2d': [_c_vp, _c_vp, _c_ll, _c_i, _c_i, _c_i, _c_vp],
    'kym_op_fft': [_c_vp, _c_vp, _c_vp, _c_ll, _c_i, _c_i, _c_i, _c_i, _c_ll, _c_vp],
    'kym_op_modulus_rotation': [_c_vp, _c_vp, _c_vp, _c_ll, _c_i, _c_vp],
    'kym_op_filter_ifft': [_c_vp, _c_vp, _c_vp, _c_ll, _c_ll, _c_i, _c_i, _c_i, _c_vp],
    'kym_op_filter_ifft_t': [_c_vp, _c_vp, _c_vp, _c_ll, _c_ll, _c_i, _c_i, _c_ll, _c_i, _c_vp],
    'kym_op_filter_ifft_g': [_c_vp, _c_vp, _c_vp, _c_ll, _c_ll, _c_i, _c_i, _c_i, _c_i, _c_i, _c_i, _c_i, _c_vp],
    'kym_op_freq_scatter': [_c_vp, ctypes.POINTER(_c_vp), ctypes.POINTER(_c_vp), ctypes.POINTER(_c_i), _c_i, _c_ll,
                            _c_i, _c_i, _c_i, _c_i, _c_i, _c_vp],
    'kym_op_lowpass': [_c_vp, _c_vp, _c_vp, _c_ll, _c_ll, _c_ll, _c_i, _c_i, _c_vp],
    'kym_op_integrals': [_c_vp, _c_vp, _c_ll, _c_ll, _c_vp, _c_i, _c_i, _c_vp],
    'kym_op_harmonic_products': [_c_vp, _c_vp, _c_vp, _c_ll, _c_i, _c_ll, _c_i, _c_vp],
    'kym_op_modulus_rotation_sum': [_c_vp, _c_vp, _c_ll, _c_i, _c_ll, _c_ll, _c_i, _c_vp],
    'kym_dec_conv_fwd': [_c_vp] * 5 + [_c_i] * 8 + [_c_vp],
    'kym_dec_conv_bwd_weight': [_c_vp] * 4 + [_c_i] * 8 + [_c_vp],
    'kym_dec_up_relu_adjoint': [_c_vp] * 3 + [_c_i] * 5 + [_c_vp],
    'kym_dec_bias_grad': [_c_vp, _c_vp, _c_i, _c_i, _c_i, _c_vp],
    'kym_dec_mse': [_c_vp] * 4 + [_c_ll, _c_vp],
    'kym_dec_adam': [_c_vp] * 4 + [_c_ll] + [ctypes.c_float] * 4 + [_c_i, _c_vp],
    'kym_dec_pack_weights': [_c_vp] * 3 + [_c_i] * 4 + [_c_vp],
    'kym_chirps': [_c_vp, _c_i, _c_i, ctypes.c_ulonglong, _c_ll, ctypes.c_float, _c_vp],
    'kym_mfma_layout_check': [_c_vp, _c_vp, _c_vp],
}


def signatures():
    """name -> ctypes argtypes of every bound entry point."""
    return dict(_SIGNATURES)


def exported_symbols():
    """Names the Python side binds (all declared in include/kymatio_hip.h)."""
    return ['kym_last_error'] + list(_SIGNATURES)


def lib():
    """Load the HIP library (once).  Raises ImportError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('libkymatio_hip.so is not built; run python kymatio_amd/csrc/build.py '
                              '(or __graft_entry__.build())')
        L = ctypes.CDLL(LIB_PATH)
        L.kym_last_error.restype = ctypes.c_char_p
        L.kym_last_error.argtypes = []
        for name, args in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().kym_last_error().decode()
        raise STATUS_EXC.get(status, RuntimeError)(msg)


def require_device(t):
    """The HIP path runs only on GPU tensors."""
    if not t.is_cuda:
        raise TypeError('The input should be a tensor on a HIP device (MI355X); this backend has no CPU '
                        'path. Move the module and the input with .cuda() / .to(device).')


def stream_of(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


DTYPE_CODE = {torch.float32: 0, torch.float64: 1}
