"""Host-side API, guardrails and the C ABI library (no GPU needed)."""

import ctypes
import pathlib
import re

import numpy as np
import pytest

import wolfgang_amd
from wolfgang_amd import PauliSum, _native

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "wolfgang_amd.h").read_text()
    declared = set(re.findall(r"\b(wg_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in include/wolfgang_amd.h but not exported"
    assert declared <= set(_native.SIGNATURES), "ctypes signatures missing for some entry points"
    assert _native.load().wg_abi_version() == 1


def test_status_without_device_is_reported_not_raised():
    status = wolfgang_amd.hip_status()
    assert status["built"] is True
    if not status["runtime_available"]:
        assert status["skip_reason"]
        with pytest.raises(RuntimeError):
            PauliSum.from_labels(["X"]).to_device()


def test_labels_endianness_and_sparse_round_trip():
    # reference docs/architecture/semantic_contracts.md:278-293
    op = PauliSum.from_labels(["XYZ"], [1.0])
    assert op.to_sparse_list() == [("ZYX", [0, 1, 2], 1 + 0j)]
    assert op.x_words()[0, 0] == 0b110 and op.z_words()[0, 0] == 0b011
    for n in (64, 65, 128, 129):
        rng = np.random.default_rng(n)
        labels = ["".join(rng.choice(list("IXYZ"), size=n)) for _ in range(7)]
        coeffs = rng.normal(size=7) + 1j * rng.normal(size=7)
        op = PauliSum.from_labels(labels, coeffs)
        back, c = op.to_labels()
        assert back == labels
        np.testing.assert_array_equal(c, coeffs)
        again = PauliSum.from_sparse_list(op.to_sparse_list(), n)
        np.testing.assert_array_equal(again.x_words(), op.x_words())
        np.testing.assert_array_equal(again.z_words(), op.z_words())


def test_label_validation_messages():
    with pytest.raises(ValueError, match="cannot infer num_qubits"):
        PauliSum.from_labels([])
    with pytest.raises(ValueError, match="same length"):
        PauliSum.from_labels(["XX", "X"])
    with pytest.raises(ValueError, match="only I, X, Y, or Z"):
        PauliSum.from_labels(["XA"])
    with pytest.raises(ValueError, match="coefficient count must match label count"):
        PauliSum.from_labels(["X", "Y"], [1.0])
    with pytest.raises(ValueError, match="not a single string"):
        PauliSum.from_labels("XY")
    with pytest.raises(ValueError, match="duplicate qubit index"):
        PauliSum.from_sparse_list([("XX", [0, 0], 1.0)], 2)
    with pytest.raises(ValueError, match="out of range"):
        PauliSum.from_sparse_list([("X", [2], 1.0)], 2)


def test_add_and_scalar_multiply_preserve_order():
    # reference tests/test_phase5_arithmetic.py (add/scale keep term order)
    op = PauliSum.from_labels(["X", "Z"], [1.0 + 2.0j, -3.0])
    labels, coeffs = (op + op).to_labels()
    assert labels == ["X", "Z", "X", "Z"]
    labels, coeffs = ((1.5j) * op).to_labels()
    np.testing.assert_array_equal(coeffs, [(1.0 + 2.0j) * 1.5j, -3.0 * 1.5j])
    zero = 0.0 * op
    assert zero.num_terms == 2
    with pytest.raises(ValueError, match="same num_qubits"):
        op + PauliSum.from_labels(["XX"])


def test_guardrails_reject_before_allocation():
    # reference tests/test_phase5_arithmetic.py:176-192, test_phase6:96-112
    with pytest.raises(ValueError, match="max_intermediate_terms"):
        PauliSum(1, 3).matmul(PauliSum(1, 4), max_intermediate_terms=11)
    with pytest.raises(ValueError, match="intermediate term count overflows"):
        PauliSum._checked_matmul_size_for_testing(1 << 63, 2, (1 << 63))
    with pytest.raises(ValueError, match="max_commutation_matrix_entries"):
        PauliSum(1, 3).commutes_with(PauliSum(1, 4), max_commutation_matrix_entries=11)
    with pytest.raises(ValueError, match="entry count overflows"):
        PauliSum._checked_commutation_size_for_testing(1 << 63, 2, (1 << 63))
    with pytest.raises(ValueError, match="non-negative integer"):
        PauliSum(1, 1).matmul(PauliSum(1, 1), max_intermediate_terms=-1)
    with pytest.raises(ValueError, match="tolerances must be non-negative"):
        PauliSum.from_labels(["X"]).simplify(atol=-1e-12)
    assert PauliSum._checked_matmul_size_for_testing(3, 4, 12) == 12


def test_backend_selector_messages():
    op = PauliSum.from_labels(["X"])
    with pytest.raises(RuntimeError, match="built without CUDA"):
        op.to_device(backend="cuda")
    with pytest.raises(ValueError, match="backend must be None"):
        op.to_device(backend="tpu")


def test_product_package_does_not_import_the_oracle():
    for path in (ROOT / "wolfgang_amd").rglob("*.py"):
        text = path.read_text()
        assert "import oracle" not in text and "from oracle" not in text, path
    for path in (ROOT / "wolfgang_amd" / "csrc").glob("*"):
        assert "oracle" not in path.read_text().lower(), path


def test_group_commuting_argument_messages():
    # Reference tests/test_phase6_commutation_grouping.py:187-194 and
    # bindings/python/pauli_sum_py.cpp:105-118 (validated before any device work).
    op = PauliSum.from_labels(["X"], [1.0])
    with pytest.raises(ValueError, match="mode"):
        op.group_commuting(mode="diagonal")
    with pytest.raises(ValueError, match="strategy"):
        op.group_commuting(strategy="smallest_first")
    with pytest.raises(ValueError, match="max_terms_for_graph must be a non-negative integer"):
        op.group_commuting(max_terms_for_graph=-1)
    assert PauliSum.empty(2).group_commuting(mode="full") == []
