"""Synthetic workloads at the benchmark shapes (seeded, reproducible).

* ``random_pauli_masks``: each qubit independently non-identity with
  probability ``density``, then uniform over X/Y/Z (the 16-qubit / 2,000-term
  algebra workload uses density 0.25 as "I w.p. 0.75"; the 128-qubit algebra
  workload uses density 0.25), complex128 coefficients with re, im ~ N(0, 1).
* ``chemistry_masks``: hopping terms X_p Z_{p+1..q-1} X_q and Y_p Z..Z Y_q for a
  random pair p < q, plus Z_p and Z_p Z_q, in equal proportions; real
  coefficients N(0,1) * exp(-|q-p|/4) (factor 1 for Z_p).
* ``random_statevector``: re, im ~ N(0, 1), normalized in FP64.
"""

from __future__ import annotations

import math

import numpy as np

from .pauli_sum import PauliSum, _pack_bits


def random_pauli_masks(num_qubits: int, num_terms: int, density: float, seed: int,
                       coefficients: str = "complex") -> PauliSum:
    rng = np.random.default_rng(seed)
    active = rng.random((num_terms, num_qubits)) < density
    kind = rng.integers(0, 3, size=(num_terms, num_qubits))  # 0 X, 1 Y, 2 Z
    xb = active & (kind <= 1)
    zb = active & (kind >= 1)
    if coefficients == "complex":
        c = rng.normal(size=num_terms) + 1j * rng.normal(size=num_terms)
    else:
        c = rng.normal(size=num_terms).astype(np.complex128)
    return PauliSum._from_arrays(num_qubits, _pack_bits(xb, num_qubits), _pack_bits(zb, num_qubits), c)


def chemistry_masks(num_qubits: int, num_terms: int, seed: int) -> PauliSum:
    if num_qubits < 2:
        raise ValueError("the chemistry-like generator needs at least 2 qubits")
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 4, size=num_terms)  # 0 XX hop, 1 YY hop, 2 Z_p, 3 Z_p Z_q
    p = rng.integers(0, num_qubits, size=num_terms)
    q = rng.integers(0, num_qubits - 1, size=num_terms)
    q = np.where(q >= p, q + 1, q)
    lo, hi = np.minimum(p, q), np.maximum(p, q)
    idx = np.arange(num_qubits)
    between = (idx[None, :] > lo[:, None]) & (idx[None, :] < hi[:, None])
    at_lo = idx[None, :] == lo[:, None]
    at_hi = idx[None, :] == hi[:, None]
    pair = at_lo | at_hi
    xb = np.zeros((num_terms, num_qubits), dtype=bool)
    zb = np.zeros((num_terms, num_qubits), dtype=bool)
    hop_x = kind == 0
    hop_y = kind == 1
    zp = kind == 2
    zz = kind == 3
    xb[hop_x | hop_y] = pair[hop_x | hop_y]
    zb[hop_x] = between[hop_x]
    zb[hop_y] = (between | pair)[hop_y]
    zb[zp] = (idx[None, :] == p[:, None])[zp]
    zb[zz] = pair[zz]
    scale = np.where(zp, 1.0, np.exp(-np.abs(hi - lo) / 4.0))
    c = (rng.normal(size=num_terms) * scale).astype(np.complex128)
    return PauliSum._from_arrays(num_qubits, _pack_bits(xb, num_qubits), _pack_bits(zb, num_qubits), c)


def random_statevector(num_qubits: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    size = 1 << num_qubits
    psi = rng.normal(size=size) + 1j * rng.normal(size=size)
    return (psi / np.linalg.norm(psi)).astype(np.complex128)


def random_statevector_device(num_qubits: int, seed: int, device: int = 0):
    """Same distribution generated directly in HBM with torch (for 2^30 amplitudes)."""
    import torch

    g = torch.Generator(device=f"cuda:{device}")
    g.manual_seed(seed)
    psi = torch.randn(1 << num_qubits, dtype=torch.complex128, device=f"cuda:{device}", generator=g)
    psi /= torch.linalg.vector_norm(psi)
    return psi


def batch_shapes(count: int, seed: int, n_range=(12, 22), t_range=(100, 50_000)) -> list[tuple[int, int]]:
    """(num_qubits, num_terms) pairs: n ~ U{lo..hi}, T = round(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    n = rng.integers(n_range[0], n_range[1] + 1, size=count)
    t = np.rint(np.exp(rng.uniform(math.log(t_range[0]), math.log(t_range[1]), size=count))).astype(int)
    return [(int(a), int(b)) for a, b in zip(n, t)]


def product_shapes(count: int, seed: int, t_range=(500, 8_000), n_range=(64, 256)) -> list[tuple[int, int, int]]:
    """(num_qubits, lhs terms, rhs terms) for batched products: n ~ U{lo..hi},
    terms ~ U{lo..hi} per operand."""
    rng = np.random.default_rng([seed, 1])
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        tl = int(rng.integers(t_range[0], t_range[1] + 1))
        tr = int(rng.integers(t_range[0], t_range[1] + 1))
        out.append((n, tl, tr))
    return out
