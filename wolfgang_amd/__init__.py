"""wolfgang_amd: the packed symplectic Pauli hot path on AMD Instinct MI355X (gfx950).

Drop-in for the reference's ``PauliSum`` / ``DevicePauliSum`` /
``DeviceCommutationMatrix`` surface (reference python/wolfgang_quantum/__init__.py:24-65).
Device operations run hand-written HIP kernels through the C ABI in
``include/wolfgang_amd.h``; there is no CPU fallback.
"""

from .pauli_sum import (
    DeviceCommutationMatrix,
    DevicePauliSum,
    PauliSum,
    checked_commutation_entries,
    checked_matmul_intermediate_terms,
    cuda_available,
    cuda_devices,
    hip_available,
    hip_devices,
    hip_status,
    metal_available,
    metal_devices,
    memory_cache_info,
    empty_cache,
    set_memory_cache_limit,
)

__all__ = [
    "PauliSum",
    "DevicePauliSum",
    "DeviceCommutationMatrix",
    "hip_available",
    "hip_devices",
    "hip_status",
    "cuda_available",
    "cuda_devices",
    "metal_available",
    "metal_devices",
    "checked_matmul_intermediate_terms",
    "checked_commutation_entries",
    "memory_cache_info",
    "empty_cache",
    "set_memory_cache_limit",
]

__version__ = "0.1.0"
