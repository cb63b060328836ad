"""Regenerators for the reference's benchmark datasets, used to re-derive its
published known-answer digests (tests/golden/reference_kats.json).

The draw sequence follows the reference's generators so that the same seeds
give the same operators: benchmarks/bench_rocm_kernels.py:1007-1083
(deterministic_labels, unique_deterministic_labels, make_simplify_operator,
make_operator), :1152-1190 (campaign6 expectation / matmul operands) and the
digests at :804-813 (commutation) and :1104-1121, :1195-1203 (label hashes).
"""

from __future__ import annotations

import hashlib

import numpy as np

_PAULIS = np.asarray(["X", "Y", "Z"])


def deterministic_labels(num_qubits: int, num_terms: int, term_weight: int, seed: int) -> list[str]:
    rng = np.random.default_rng(seed)
    weight = min(term_weight, num_qubits)
    labels = []
    for _ in range(num_terms):
        chars = ["I"] * num_qubits
        for qubit in rng.choice(num_qubits, size=weight, replace=False):
            chars[num_qubits - 1 - int(qubit)] = str(rng.choice(_PAULIS))
        labels.append("".join(chars))
    return labels


def unique_deterministic_labels(num_qubits: int, unique_terms: int, term_weight: int, seed: int) -> list[str]:
    labels: list[str] = []
    seen: set[str] = set()
    attempt = 0
    while len(labels) < unique_terms:
        batch = deterministic_labels(
            num_qubits, max(unique_terms - len(labels), 32), term_weight, seed + attempt * 104729
        )
        for label in batch:
            if label in seen:
                continue
            seen.add(label)
            labels.append(label)
            if len(labels) == unique_terms:
                break
        attempt += 1
    return labels


def simplify_operator_labels(case: dict) -> tuple[list[str], list[complex]]:
    n, t, w, seed = case["num_qubits"], case["num_terms"], case["term_weight"], case["seed"]
    if case["mode"] == "cancelling_duplicates":
        label = unique_deterministic_labels(n, 1, w, seed)[0]
        return [label] * t, [1.0 if i % 2 == 0 else -1.0 for i in range(t)]
    unique = max(1, min(t, int(round(t * (1.0 - case["duplicate_rate"])))))
    pool = unique_deterministic_labels(n, unique, w, seed)
    labels = [pool[i % unique] for i in range(t)]
    rng = np.random.default_rng(seed + 17)
    coeffs = [complex(float(rng.normal()), float(rng.normal())) for _ in range(t)]
    return labels, coeffs


def commutation_operands(case: dict) -> tuple[list[str], list[str]]:
    n, w, seed = case["num_qubits"], case["term_weight"], case["seed"]
    lhs = deterministic_labels(n, case["lhs_terms"], w, seed)
    rhs = deterministic_labels(n, case["rhs_terms"], w, seed + 100000)
    return lhs, rhs


def campaign6_expectation(case: dict) -> tuple[list[str], list[complex], np.ndarray]:
    n, t, w, seed = case["num_qubits"], case["num_terms"], case["term_weight"], case["seed"]
    labels = deterministic_labels(n, t, w, seed)
    rng = np.random.default_rng(seed + 41)
    coeffs = [complex(float(rng.normal()), float(rng.normal())) for _ in range(t)]
    rng = np.random.default_rng(seed + 73)
    size = 2**n
    psi = rng.normal(size=size) + 1j * rng.normal(size=size)
    psi = np.asarray(psi / np.linalg.norm(psi), dtype=np.complex128)
    if case["dtype"] == "complex64":
        psi = psi.astype(np.complex64)
    return labels, coeffs, psi


def campaign6_matmul_operand(case: dict, side: str) -> tuple[list[str], list[complex]]:
    terms = case[f"{side}_terms"]
    unique = max(1, min(terms, int(round(terms * (1.0 - case["duplicate_rate"])))))
    offset = 0 if side == "lhs" else 100000
    pool = unique_deterministic_labels(case["num_qubits"], unique, case["term_weight"], case["seed"] + offset)
    labels = [pool[i % unique] for i in range(terms)]
    rng = np.random.default_rng(case["seed"] + offset + 17)
    coeffs = [complex(float(rng.normal()), float(rng.normal())) for _ in range(terms)]
    return labels, coeffs


def label_hash(labels: list[str], coeffs) -> str:
    digest = hashlib.sha256()
    for label, coeff in zip(labels, np.asarray(coeffs, dtype=np.complex128)):
        digest.update(label.encode("utf-8"))
        digest.update(b"\0")
        digest.update(np.float64(coeff.real).tobytes())
        digest.update(np.float64(coeff.imag).tobytes())
    return digest.hexdigest()
