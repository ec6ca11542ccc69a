"""Kubelka-Munk spectral mix on MFMA (include/canvaz_km.h) from Python:
device buffers are torch tensors on the HIP device, the kernel runs on
torch's current stream."""
import ctypes as C

import numpy as np

import canvaz_film as czf

_sig_done = False


def _lib():
    global _sig_done
    L = czf.lib()
    if not _sig_done:
        L.cz_km_prepare.restype = C.c_int
        L.cz_km_prepare.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32,
                                    C.POINTER(C.c_uint16), C.POINTER(C.c_uint32)]
        L.cz_km_mix.restype = C.c_int
        L.cz_km_mix.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p]
        L.cz_km_last_error.restype = C.c_char_p
        _sig_done = True
    return L


def prepare(k_tab, s_tab):
    """BF16 hi/lo operand tiles of the FP32 tables [12][bands] (numpy u16)."""
    k = np.ascontiguousarray(k_tab, np.float32)
    s = np.ascontiguousarray(s_tab, np.float32)
    bands = k.shape[1]
    nt = (bands + 15) // 16
    ops = np.zeros(2 * nt * 64 * 8, np.uint16)
    ntiles = C.c_uint32()
    rc = _lib().cz_km_prepare(k.ctypes.data_as(C.POINTER(C.c_float)), s.ctypes.data_as(C.POINTER(C.c_float)),
                              bands, ops.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(ntiles))
    if rc:
        raise czf.FilmError(_lib().cz_km_last_error().decode())
    return ops


def mix(conc, ops, bands, out=None):
    """R_inf per (cell, band) for conc (torch bfloat16 [cells, 12] on the device)."""
    import torch
    assert conc.dtype == torch.bfloat16 and conc.is_cuda and conc.shape[1] == 12 and conc.is_contiguous()
    cells = conc.shape[0]
    if out is None:
        out = torch.empty((cells, bands), dtype=torch.float32, device=conc.device)
    stream = torch.cuda.current_stream().cuda_stream
    rc = _lib().cz_km_mix(conc.data_ptr(), ops.data_ptr(), cells, bands, out.data_ptr(), stream)
    if rc:
        raise czf.FilmError(_lib().cz_km_last_error().decode())
    return out
