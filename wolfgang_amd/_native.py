"""ctypes binding of the wolfgang_amd C ABI (include/wolfgang_amd.h).

The library is loaded from the package directory only. There is no fallback:
if ``libwolfgang_amd.so`` is missing or fails to load, every device operation
raises ``RuntimeError`` naming the build step to run.
"""

from __future__ import annotations

import ctypes as C
import pathlib
import threading

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libwolfgang_amd.so"

WG_OK = 0
WG_ERR_INVALID = 1
WG_ERR_RUNTIME = 2
WG_ERR_TYPE = 3
WG_ERR_OVERFLOW = 4

WG_DTYPE_COMPLEX64 = 0
WG_DTYPE_COMPLEX128 = 1

_u64p = C.POINTER(C.c_uint64)
_dblp = C.POINTER(C.c_double)
_size = C.c_size_t
_vp = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "wg_last_error": (C.c_char_p, []),
    "wg_abi_version": (C.c_int, []),
    "wg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "wg_device_info": (
        C.c_int,
        [C.c_int, C.c_char_p, _size, C.c_char_p, _size, C.POINTER(_size), C.POINTER(C.c_int)],
    ),
    "wg_runtime_versions": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "wg_pauli_sum_create": (C.c_int, [_size, _size, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "wg_pauli_sum_to_host": (C.c_int, [_vp, _vp, _vp, _vp]),
    "wg_pauli_sum_info": (
        C.c_int,
        [_vp, C.POINTER(_size), C.POINTER(_size), C.POINTER(_size), C.POINTER(C.c_int)],
    ),
    "wg_pauli_sum_device_buffers": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "wg_pauli_sum_destroy": (None, [_vp]),
    "wg_simplify": (C.c_int, [_vp, C.c_double, C.c_double, _vp, C.POINTER(_vp)]),
    "wg_matmul": (C.c_int, [_vp, _vp, C.c_int, _size, _vp, C.POINTER(_vp)]),
    "wg_commutes_with_host": (C.c_int, [_vp, _vp, _vp, _size, _size, _vp]),
    "wg_comm_matrix_create": (C.c_int, [_size, _size, C.c_int, C.POINTER(_vp)]),
    "wg_comm_matrix_destroy": (None, [_vp]),
    "wg_comm_matrix_info": (
        C.c_int,
        [_vp, C.POINTER(_size), C.POINTER(_size), C.POINTER(_size), C.POINTER(C.c_int)],
    ),
    "wg_comm_matrix_device_bits": (C.c_int, [_vp, C.POINTER(_vp)]),
    "wg_commutes_with_device": (C.c_int, [_vp, _vp, _vp, _size, _vp]),
    "wg_comm_matrix_to_host_u8": (C.c_int, [_vp, _vp, _size, _vp]),
    "wg_comm_matrix_to_host_packed": (C.c_int, [_vp, _vp, _size, _vp]),
    "wg_comm_matrix_from_host_u8": (C.c_int, [_vp, _vp, _size, _vp]),
    "wg_memory_cache_limit": (C.c_int, [_size, C.POINTER(_size)]),
    "wg_memory_cache_release": (C.c_int, [C.c_int]),
    "wg_memory_cache_info": (C.c_int, [C.c_int, C.POINTER(_size), C.POINTER(_size), C.POINTER(_size)]),
    "wg_comm_matrix_count": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _size, _vp]),
    "wg_expectation": (C.c_int, [_vp, _vp, C.c_int, _size, C.c_int, _vp, _dblp]),
    "wg_expectation_adjoint": (C.c_int, [_vp, _vp, _size, _vp, _vp, _vp, _vp, _vp, _dblp]),
    "wg_expectation_noalloc": (C.c_int, [_vp, _vp, _size, _vp, _vp]),
    "wg_expectation_z_counts": (C.c_int, [_vp, C.c_char_p, _vp, _vp, _size, _vp, _dblp]),
    "wg_expectation_plan_info": (
        C.c_int, [_vp, C.POINTER(_size), C.POINTER(_size), C.POINTER(_size), C.POINTER(C.c_int), _vp]
    ),
    "wg_debug_plan_passes": (C.c_int, [_size, _size, _vp, _vp, _vp, _size, C.POINTER(_size)]),
    "wg_debug_key_sweeps": (C.c_int, [_size, _size, _vp, _vp, _vp, _vp, C.POINTER(_size)]),
    "wg_debug_sigma_plan": (C.c_int, [_size, _size, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_size)]),
    "wg_debug_tiled_plan": (
        C.c_int, [_size, _size, _vp, _vp, _vp, _vp, C.POINTER(_size), C.POINTER(_size)]
    ),
    "wg_batch_create": (C.c_int, [_size, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "wg_batch_run": (C.c_int, [_vp, _vp]),
    "wg_batch_sync": (C.c_int, [_vp, _vp]),
    "wg_batch_info": (C.c_int, [_vp, _vp, C.POINTER(_size), C.POINTER(_size)]),
    "wg_batch_tiled_count": (C.c_int, [_vp, C.POINTER(_size)]),
    "wg_batch_destroy": (None, [_vp]),
    "wg_group_commuting": (C.c_int, [_vp, C.c_int, _vp, C.POINTER(_size), _vp, _vp]),
    "wg_product_batch_create": (C.c_int, [_size, _vp, _vp, _size, _size, _vp, C.POINTER(_vp)]),
    "wg_product_batch_run": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "wg_product_batch_info": (C.c_int, [_vp, C.POINTER(_size), C.POINTER(_size), C.POINTER(_size)]),
    "wg_product_batch_destroy": (None, [_vp]),
    "wg_profile_enable": (C.c_int, [C.c_int]),
    "wg_profile_reset": (C.c_int, []),
    "wg_profile_get": (C.c_int, [C.c_char_p, _dblp, C.POINTER(C.c_uint64)]),
}

_lock = threading.Lock()
_lib: C.CDLL | None = None
_load_error: str | None = None


def _share_hip_runtime_with_torch() -> None:
    """Make this library bind to the HIP runtime PyTorch uses, when PyTorch is
    installed. PyTorch ships its own libamdhip64.so.7 with the same SONAME as
    the system ROCm runtime; whichever is loaded first serves the whole process,
    and PyTorch cannot initialise the GPU on a runtime other than its own. Loading
    PyTorch's copy (RTLD_GLOBAL) before this library keeps one runtime shared
    by both, so device pointers and streams can cross between them."""
    import importlib.util

    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or spec.origin is None:
        return
    libdir = pathlib.Path(spec.origin).parent / "lib"
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        path = libdir / name
        if path.exists():
            try:
                C.CDLL(str(path), mode=C.RTLD_GLOBAL)
            except OSError:
                return


def load() -> C.CDLL:
    """Return the loaded library, raising RuntimeError if it cannot be loaded."""
    global _lib, _load_error
    with _lock:
        if _lib is not None:
            return _lib
        _share_hip_runtime_with_torch()
        if not LIB_PATH.exists():
            _load_error = (
                f"wolfgang_amd native library not found at {LIB_PATH}; build it with "
                "`python -m wolfgang_amd.build_native`"
            )
            raise RuntimeError(_load_error)
        try:
            lib = C.CDLL(str(LIB_PATH))
        except OSError as exc:
            _load_error = f"failed to load {LIB_PATH}: {exc}"
            raise RuntimeError(_load_error) from exc
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
        return lib


def last_error() -> str:
    msg = load().wg_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    """Translate a C ABI status into the reference's Python exception types."""
    if status == WG_OK:
        return
    msg = last_error()
    if status == WG_ERR_INVALID:
        raise ValueError(msg)
    if status == WG_ERR_TYPE:
        raise TypeError(msg)
    if status == WG_ERR_OVERFLOW:
        raise OverflowError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def profile_enable(enabled: bool = True) -> None:
    call("wg_profile_enable", int(bool(enabled)))


def profile_reset() -> None:
    call("wg_profile_reset")


def profile_get(kernel: str) -> tuple[float, int]:
    """(total milliseconds, launches) recorded for a kernel since the last reset."""
    ms = C.c_double()
    n = C.c_uint64()
    call("wg_profile_get", kernel.encode(), C.byref(ms), C.byref(n))
    return ms.value, n.value
