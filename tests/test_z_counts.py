"""expectation_z_counts: the oracle pinned by the reference's known answers
(tests/test_phase8_expectation.py:140-186 of the reference: a direct
computation, the dense-bitstring endianness, the validation messages), by
closed forms (all counts on one bitstring; uniform counts over every
bitstring), and by the reference's own code (oracle/_ref); then the device
path against the oracle."""

import itertools

import numpy as np
import pytest

from wolfgang_amd.pauli_sum import PauliSum, _parse_z_counts
from wolfgang_amd.workloads import random_pauli_masks

LABELS = ["ZI", "IZ", "ZZ", "II"]
COEFFS = [1.25, -0.5 + 0.25j, 2.0, 0.75]
COUNTS = {"00": 7, "01": 5, "10": 3, "11": 11}


def direct(labels, coeffs, counts):
    # Label character i and bitstring character i are the same qubit; a Z
    # meeting a '1' flips the sign.
    total = sum(counts.values())
    e = 0j
    for lab, c in zip(labels, coeffs):
        s = 0.0
        for bits, n in counts.items():
            odd = sum(1 for p, b in zip(lab, bits) if p == "Z" and b == "1") % 2
            s += -n if odd else n
        e += c * s / total
    return e


def args(op, counts):
    return (op.num_qubits, op.x_words(), op.z_words(), op.coeffs(), list(counts), list(counts.values()))


def diagonal_op(n, t, seed):
    op = random_pauli_masks(n, t, 0.4, seed)
    return PauliSum.from_packed(n, np.zeros_like(op.x_words()), op.z_words(), op.coeffs())


def random_counts(n, rows, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, size=(rows, n))
    keys = ["".join("1" if v else "0" for v in row) for row in bits]
    counts = {}
    for k in keys:
        counts[k] = counts.get(k, 0) + int(rng.integers(1, 50))
    return counts


def test_oracle_reference_fixture(orc):
    op = PauliSum.from_labels(LABELS, COEFFS)
    got = orc.expectation_z_counts(*args(op, COUNTS))
    assert abs(got - direct(LABELS, COEFFS, COUNTS)) <= 1e-12


def test_oracle_endianness(orc):
    counts = {"01": 10}
    assert orc.expectation_z_counts(*args(PauliSum.from_labels(["IZ"], [1.0]), counts)) == -1.0
    assert orc.expectation_z_counts(*args(PauliSum.from_labels(["ZI"], [1.0]), counts)) == 1.0


def test_oracle_closed_forms(orc):
    n = 6
    op = diagonal_op(n, 40, 11)
    z = op.z_words()[:, 0]
    c = op.coeffs()
    # All counts on one bitstring b: E = sum_t c_t (-1)^{|z_t & b|}.
    for b in (0, 0b101101, 0b111111):
        bits = format(b, f"0{n}b")
        want = sum(ct * (-1) ** bin(int(zt) & b).count("1") for zt, ct in zip(z, c))
        assert abs(orc.expectation_z_counts(*args(op, {bits: 3})) - want) <= 1e-12 * np.abs(c).sum()
    # Uniform counts over every bitstring: only z = 0 terms survive.
    uniform = {"".join(p): 2 for p in itertools.product("01", repeat=n)}
    want = c[z == 0].sum()
    assert abs(orc.expectation_z_counts(*args(op, uniform)) - want) <= 1e-12 * np.abs(c).sum()


@pytest.mark.parametrize("n,terms", [(10, 50), (70, 30), (200, 20)])
def test_oracle_matches_reference(orc, ref_lib, n, terms):
    op = diagonal_op(n, terms, 100 + n)
    counts = random_counts(n, 300, n)
    assert orc.expectation_z_counts(*args(op, counts)) == ref_lib.expectation_z_counts(*args(op, counts))


@pytest.mark.parametrize("labels,counts,match", [
    (["X"], {"0": 1}, "diagonal"),
    (["Y"], {"0": 1}, "diagonal"),
    (["ZZ"], {}, "total count"),
    (["ZZ"], {"0": 1}, "length"),
    (["ZZ"], {"02": 1}, "0 or 1"),
    (["ZZ"], {"00": -1}, "non-negative"),
    (["ZZ"], {"00": float("inf")}, "finite"),
    (["ZZ"], {"00": float("nan")}, "finite"),
])
def test_oracle_validation(orc, labels, counts, match):
    op = PauliSum.from_labels(labels, [1.0])
    with pytest.raises(ValueError, match=match):
        orc.expectation_z_counts(*args(op, counts))


def test_parse_counts_mapping():
    b, v = _parse_z_counts({"01": 3, "10": 2.5})
    assert b == ["01", "10"] and v.tolist() == [3.0, 2.5]
    with pytest.raises(ValueError, match="mapping"):
        _parse_z_counts([("01", 1)])
    with pytest.raises(ValueError, match="strings"):
        _parse_z_counts({1: 1})
    with pytest.raises(ValueError, match="numeric"):
        _parse_z_counts({"00": object()})
    with pytest.raises(ValueError, match="numeric"):
        _parse_z_counts({"00": "3"})


# ------------------------------------------------------------------ device


@pytest.mark.gpu
def test_device_reference_fixture(gpu):
    op = PauliSum.from_labels(LABELS, COEFFS)
    assert abs(op.expectation_z_counts(COUNTS) - direct(LABELS, COEFFS, COUNTS)) <= 1e-12
    assert PauliSum.from_labels(["IZ"], [1.0]).expectation_z_counts({"01": 10}) == pytest.approx(-1.0)
    assert PauliSum.from_labels(["ZI"], [1.0]).expectation_z_counts({"01": 10}) == pytest.approx(1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("n,terms,rows", [(20, 300, 3000), (100, 200, 2000), (130, 64, 500),
                                          (200, 40, 700), (300, 25, 400)])
def test_device_matches_oracle(gpu, orc, n, terms, rows):
    op = diagonal_op(n, terms, 7 * n)
    counts = random_counts(n, rows, n + 1)
    want = orc.expectation_z_counts(*args(op, counts))
    got = op.to_device().expectation_z_counts(counts)
    assert abs(got - want) <= 1e-12 * np.abs(op.coeffs()).sum()


@pytest.mark.gpu
def test_device_edge_cases(gpu):
    assert PauliSum.empty(3).expectation_z_counts({"010": 4}) == 0
    zero_q = PauliSum.from_packed(0, np.zeros((2, 0), np.uint64), np.zeros((2, 0), np.uint64),
                                  np.array([1.5, -0.25j]))
    assert zero_q.expectation_z_counts({"": 3}) == 1.5 - 0.25j


@pytest.mark.gpu
@pytest.mark.parametrize("labels,counts,match", [
    (["X"], {"0": 1}, "diagonal"),
    (["IY"], {"00": 1}, "diagonal"),
    (["ZZ"], {}, "total count"),
    (["ZZ"], {"0": 1}, "length"),
    (["ZZ"], {"02": 1}, "0 or 1"),
    (["ZZ"], {"00": -1}, "non-negative"),
    (["ZZ"], {"00": float("inf")}, "finite"),
    (["ZZ"], {"00": object()}, "numeric"),
])
def test_device_validation(gpu, labels, counts, match):
    with pytest.raises(ValueError, match=match):
        PauliSum.from_labels(labels, [1.0]).expectation_z_counts(counts)
