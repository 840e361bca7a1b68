"""ctypes binding of ``libfastsimus.so`` (declarations in ``include/fastsimus.h``).

The library is loaded from this package directory only; there is no
fallback: if it is missing, every compute call raises ``RuntimeError``.
Status codes map to the exception types the reference raises for the same
conditions (``ValueError`` for bad inputs, ``NotImplementedError`` for
unsupported options, ``RuntimeError`` for resource / runtime failures).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfastsimus.so"

FS_OK, FS_ERR_INVALID, FS_ERR_UNSUPPORTED, FS_ERR_WORKSPACE, FS_ERR_LAUNCH, FS_ERR_RESOURCE, FS_ERR_SCRATCH = range(7)
FS_STATIC_SCRATCH_BYTES = 64 << 20
FS_METHOD_FP32, FS_METHOD_MFMA_BF16X3 = 0, 1
FS_ADJOINT_AUTO, FS_ADJOINT_MATRIX, FS_ADJOINT_VECTOR = 0, 1, 2

# every symbol declared in include/fastsimus.h
EXPORTS = (
    "fs_version",
    "fs_status_string",
    "fs_simus_workspace_bytes",
    "fs_simus_spectrum",
    "fs_simus_spectrum_static",
    "fs_simus_batch_workspace_bytes",
    "fs_simus_spectrum_batch",
    "fs_tuning_set",
    "fs_tuning_get",
    "fs_rf_workspace_bytes",
    "fs_spectrum_to_rf",
    "fs_rf_adjoint_workspace_bytes",
    "fs_rf_adjoint",
    "fs_simus_adjoint_workspace_bytes",
    "fs_simus_adjoint",
    "fs_pfield_workspace_bytes",
    "fs_pfield",
    "fs_max_distance",
    "fs_rf2iq_workspace_bytes",
    "fs_rf2iq",
    "fs_das",
    "fs_set_kernel_timing",
    "fs_kernel_time_ms",
)


class SimusDesc(C.Structure):
    """Mirror of ``fs_simus_desc``."""

    _fields_ = [
        ("n_scat", C.c_int32), ("n_elem", C.c_int32), ("n_sub", C.c_int32), ("n_freq", C.c_int32),
        ("n_tx", C.c_int32), ("full_frequency_directivity", C.c_int32), ("baffle_kind", C.c_int32),
        ("pad0", C.c_int32),
        ("scat_x", C.c_void_p), ("scat_z", C.c_void_p), ("rc", C.c_void_p),
        ("sub_x", C.c_void_p), ("sub_z", C.c_void_p), ("elem_cos", C.c_void_p), ("elem_sin", C.c_void_p),
        ("tx_delay", C.c_void_p), ("tx_apod", C.c_void_p),
        ("pulse_re", C.c_void_p), ("pulse_im", C.c_void_p), ("probe", C.c_void_p),
        ("f0", C.c_double), ("df", C.c_double), ("speed_of_sound", C.c_double), ("attenuation", C.c_double),
        ("freq_center", C.c_double), ("seg_length", C.c_double), ("radius", C.c_double),
        ("apex_offset", C.c_double), ("baffle_impedance", C.c_double), ("correction", C.c_double),
    ]


class Iir(C.Structure):
    """Mirror of ``fs_iir``: cascaded second-order sections, odd-extension length, section steady states."""

    _fields_ = [("n_sections", C.c_int32), ("padlen", C.c_int32), ("sos", (C.c_double * 6) * 8),
                ("zi", (C.c_double * 2) * 8)]


class Tuning(C.Structure):
    """Mirror of ``fs_tuning`` (zero fields mean default)."""

    _fields_ = [("msm_fbk", C.c_int32), ("msm_units", C.c_int32), ("msm_item_div", C.c_int32),
                ("msm_table", C.c_int32), ("reserved", C.c_int32 * 4)]


_lib = None


def load() -> C.CDLL:
    """Load and type the shared library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH.name} is not built; run `python -m fastsimus_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    D = C.POINTER(SimusDesc)
    lib.fs_version.restype = C.c_int
    lib.fs_status_string.restype = C.c_char_p
    lib.fs_status_string.argtypes = [C.c_int]
    lib.fs_simus_workspace_bytes.restype = C.c_size_t
    lib.fs_simus_workspace_bytes.argtypes = [D, C.c_int]
    lib.fs_simus_spectrum.restype = C.c_int
    lib.fs_simus_spectrum.argtypes = [D, C.c_int, P, P, C.c_size_t, P]
    lib.fs_simus_spectrum_static.restype = C.c_int
    lib.fs_simus_spectrum_static.argtypes = [D, P, P]
    I64 = C.POINTER(C.c_int64)
    lib.fs_simus_batch_workspace_bytes.restype = C.c_size_t
    lib.fs_simus_batch_workspace_bytes.argtypes = [D, C.c_int, C.c_int, I64]
    lib.fs_simus_spectrum_batch.restype = C.c_int
    lib.fs_simus_spectrum_batch.argtypes = [D, C.c_int, C.c_int, I64, P, P, C.c_size_t, P]
    lib.fs_tuning_set.restype = C.c_int
    lib.fs_tuning_set.argtypes = [D, C.c_int, C.POINTER(Tuning)]
    lib.fs_tuning_get.restype = C.c_int
    lib.fs_tuning_get.argtypes = [D, C.c_int, C.POINTER(Tuning)]
    lib.fs_rf_workspace_bytes.restype = C.c_size_t
    lib.fs_rf_workspace_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.fs_spectrum_to_rf.restype = C.c_int
    lib.fs_spectrum_to_rf.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P,
                                      C.c_size_t, P]
    lib.fs_rf_adjoint_workspace_bytes.restype = C.c_size_t
    lib.fs_rf_adjoint_workspace_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.fs_rf_adjoint.restype = C.c_int
    lib.fs_rf_adjoint.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, C.c_size_t, P]
    lib.fs_simus_adjoint_workspace_bytes.restype = C.c_size_t
    lib.fs_simus_adjoint_workspace_bytes.argtypes = [D, C.c_int]
    lib.fs_simus_adjoint.restype = C.c_int
    lib.fs_simus_adjoint.argtypes = [D, C.c_int, P, P, P, P, P, P, C.c_size_t, P]
    lib.fs_pfield_workspace_bytes.restype = C.c_size_t
    lib.fs_pfield_workspace_bytes.argtypes = [D]
    lib.fs_pfield.restype = C.c_int
    lib.fs_pfield.argtypes = [D, P, P, C.c_size_t, P]
    lib.fs_max_distance.restype = C.c_int
    lib.fs_max_distance.argtypes = [P, P, C.c_longlong, P, P, C.c_int, P, P]
    F = C.POINTER(Iir)
    lib.fs_rf2iq_workspace_bytes.restype = C.c_size_t
    lib.fs_rf2iq_workspace_bytes.argtypes = [C.c_int, C.c_int, F]
    lib.fs_rf2iq.restype = C.c_int
    lib.fs_rf2iq.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, F, C.c_int, P, P, C.c_size_t, P]
    lib.fs_das.restype = C.c_int
    lib.fs_das.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, C.c_longlong, C.c_double,
                           C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, P, P]
    lib.fs_set_kernel_timing.restype = C.c_int
    lib.fs_set_kernel_timing.argtypes = [C.c_int]
    lib.fs_kernel_time_ms.restype = C.c_double
    lib.fs_kernel_time_ms.argtypes = [C.POINTER(C.c_int)]
    _lib = lib
    return lib


def kernel_time_ms() -> tuple[float, int]:
    """Summed duration (ms) and count of the timed dominant-kernel launches since the last call."""
    n = C.c_int(0)
    ms = load().fs_kernel_time_ms(C.byref(n))
    return float(ms), int(n.value)


def check(status: int, what: str) -> None:
    """Raise the exception type matching a non-zero status."""
    if status == FS_OK:
        return
    msg = f"{what}: {load().fs_status_string(status).decode()} (status {status})"
    if status == FS_ERR_INVALID:
        raise ValueError(msg)
    if status == FS_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)
