"""GPU parity for the algebra hot path: commutation bits, products, simplify.
Bit-exact against the CPU oracle (masks, phases, order and coefficient bits)."""

import json
import pathlib

import numpy as np
import pytest

import kat_datasets as kd
from conftest import assert_bit_exact, packed
from wolfgang_amd import DeviceCommutationMatrix, PauliSum
from wolfgang_amd.workloads import random_pauli_masks

pytestmark = pytest.mark.gpu
KATS = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_kats.json").read_text())


def xz(op):
    return op.x_words(), op.z_words()


# ------------------------------------------------------------ commutation

@pytest.mark.parametrize("n", [1, 8, 63, 64, 65, 128, 130, 256, 300])
@pytest.mark.parametrize("shape", [(1, 1), (1, 77), (77, 1), (63, 64), (64, 65), (130, 200)])
def test_commutation_bits_match_oracle(gpu, orc, n, shape):
    L, R = shape
    a = random_pauli_masks(n, L, 0.3, 10 * n + L)
    b = random_pauli_masks(n, R, 0.3, 10 * n + R + 1)
    want = orc.commutes_with(n, xz(a), xz(b)).astype(bool)
    got = a.commutes_with(b)
    if L == 1 and R == 1:
        assert got is bool(want[0, 0])
    elif L == 1 or R == 1:
        np.testing.assert_array_equal(got, want.reshape(-1))
    else:
        np.testing.assert_array_equal(got, want)
    da, db = a.to_device(), b.to_device()
    m = da.commutes_with_device(db)
    np.testing.assert_array_equal(m.to_host(), want)
    assert m.count_commuting() == int(want.sum())
    np.testing.assert_array_equal(m.count_commuting(axis=1), want.sum(axis=1))
    np.testing.assert_array_equal(m.count_commuting(axis=0), want.sum(axis=0))
    np.testing.assert_array_equal(m.conflict_degrees(axis=1), (~want).sum(axis=1))
    assert m.conflict_degrees() == int((~want).sum())
    packed_bits = m.to_host_packed()
    unpacked = np.unpackbits(packed_bits.view(np.uint8), bitorder="little").reshape(L, -1)[:, :R]
    np.testing.assert_array_equal(unpacked.astype(bool), want)


@pytest.mark.parametrize("case", KATS["commutation"], ids=lambda c: c["name"])
def test_commutation_reference_counts(gpu, case):
    lhs, rhs = kd.commutation_operands(case)
    dl, dr = PauliSum.from_labels(lhs).to_device(), PauliSum.from_labels(rhs).to_device()
    m = dl.commutes_with_device(dr)
    assert m.count_commuting() == case["commuting_count"]
    assert m.conflict_degrees() == case["conflict_count"]


def test_commutation_edges_and_errors(gpu):
    empty = PauliSum.empty(5).to_device()
    some = random_pauli_masks(5, 4, 0.5, 1).to_device()
    assert empty.commutes_with(some).shape == (0, 4) or empty.commutes_with(some).size == 0
    m = DeviceCommutationMatrix.empty((3, 3))
    with pytest.raises(ValueError, match="output shape does not match"):
        some.commutes_with_device(some, output=m)
    with pytest.raises(ValueError, match="same num_qubits"):
        some.commutes_with(random_pauli_masks(6, 4, 0.5, 1).to_device())
    with pytest.raises(ValueError, match="max_commutation_matrix_entries"):
        some.commutes_with(some, max_commutation_matrix_entries=15)
    out = np.zeros(16, dtype=bool)
    some.commutes_with_into(some, out)
    np.testing.assert_array_equal(out.reshape(4, 4), some.commutes_with(some))
    # Zero-qubit operators: everything commutes.
    z0 = PauliSum(0, 3)
    assert z0.commutes_with(z0).all()


def test_commutation_overwrites_poisoned_output(gpu, orc):
    """The reference's poison check (benchmarks/bench_metal_kernels.py:946):
    a matrix filled from the host through the test hook is fully overwritten
    by commutes_with_device(output=...)."""
    from wolfgang_amd.pauli_sum import _copy_device_commutation_matrix_from_host_for_testing as poison

    a = random_pauli_masks(100, 70, 0.3, 5)
    b = random_pauli_masks(100, 130, 0.3, 6)
    da, db = a.to_device(), b.to_device()
    m = DeviceCommutationMatrix.empty((70, 130))
    rng = np.random.default_rng(7)
    flags = rng.integers(0, 3, size=(70, 130), dtype=np.uint8)
    poison(m, flags)
    np.testing.assert_array_equal(m.to_host(), flags != 0)
    assert m.count_commuting() == int((flags != 0).sum())
    # Bits past the last column stay zero: the per-row counts see only real columns.
    np.testing.assert_array_equal(m.count_commuting(axis=1), (flags != 0).sum(axis=1))
    poison(m, np.ones((70, 130), dtype=np.uint8))
    da.commutes_with_device(db, output=m)
    np.testing.assert_array_equal(m.to_host(), orc.commutes_with(100, xz(a), xz(b)).astype(bool))
    with pytest.raises(ValueError, match="testing copy size does not match matrix entries"):
        poison(m, np.ones(5, dtype=np.uint8))


# ------------------------------------------------------------ products

@pytest.mark.parametrize("n", [1, 16, 64, 65, 128, 200])
@pytest.mark.parametrize("simplify", [False, True])
def test_matmul_bit_exact(gpu, orc, n, simplify):
    a = random_pauli_masks(n, 53, 0.3, 7 * n + 1)
    b = random_pauli_masks(n, 41, 0.3, 7 * n + 2)
    want = orc.matmul(n, packed(a)[1:], packed(b)[1:], simplify=simplify)
    got = a.matmul(b, simplify=simplify)
    assert_bit_exact(got, want)


@pytest.mark.parametrize("n", [20, 50, 150, 256])
def test_matmul_one_sweep_sort_bit_exact(gpu, orc, n):
    # 1,024 x 1,024 = 2^20 products: the smallest product sort that runs the
    # one-sweep radix passes, whose digit counts come from the operands' key
    # words (packed x << 32 | z at n = 20, x0 at 50; two and four mask words
    # at 150 and 256).
    a = random_pauli_masks(n, 1024, 0.3, 11 * n + 1)
    b = random_pauli_masks(n, 1024, 0.3, 11 * n + 2)
    assert_bit_exact(a @ b, orc.matmul(n, packed(a)[1:], packed(b)[1:]))


@pytest.mark.parametrize("n", [1, 65])
def test_duplicate_heavy_matmul_hash_path_bit_exact(gpu, orc, n):
    # reference tests/test_phase5_arithmetic.py:148-173: 128 x 128 products of
    # 2-term pools take the hash-accumulation path.
    pool_a = random_pauli_masks(n, 2, 0.6, 91)
    pool_b = random_pauli_masks(n, 2, 0.6, 92)
    idx = np.arange(128) % 2
    ca = np.array([complex((i % 17) - 8, (i % 5) - 2) for i in range(128)])
    cb = np.array([complex((i % 11) - 5, (i % 7) - 3) for i in range(128)])
    a = PauliSum.from_packed(n, pool_a.x_words()[idx], pool_a.z_words()[idx], ca)
    b = PauliSum.from_packed(n, pool_b.x_words()[idx], pool_b.z_words()[idx], cb)
    assert_bit_exact(a @ b, orc.matmul(n, packed(a)[1:], packed(b)[1:]))
    # Real coefficients: -0.0 imaginary parts make the accumulation seed visible.
    a = PauliSum.from_packed(n, pool_a.x_words()[idx], pool_a.z_words()[idx], ca.real.astype(complex))
    b = PauliSum.from_packed(n, pool_b.x_words()[idx], pool_b.z_words()[idx], -cb.real.astype(complex))
    assert_bit_exact(a @ b, orc.matmul(n, packed(a)[1:], packed(b)[1:]))


@pytest.mark.parametrize(
    ("lhs", "rhs", "label", "phase"),
    [("X", "Y", "Z", 1j), ("Y", "X", "Z", -1j), ("Y", "Z", "X", 1j),
     ("Z", "Y", "X", -1j), ("Z", "X", "Y", 1j), ("X", "Z", "Y", -1j)],
)
def test_phase_table_on_device(gpu, lhs, rhs, label, phase):
    out = PauliSum.from_labels([lhs], [2.0]) @ PauliSum.from_labels([rhs], [3.0])
    labels, coeffs = out.to_labels()
    assert labels == [label] and coeffs[0] == 6.0 * phase


def test_matmul_order_fixture_and_guards(gpu):
    lhs = PauliSum.from_labels(["X", "Z"], [1.0, 2.0])
    rhs = PauliSum.from_labels(["Y", "I"], [3.0, 4.0])
    labels, coeffs = lhs.matmul(rhs, simplify=False).to_labels()
    assert labels == ["Z", "X", "X", "Z"]
    np.testing.assert_array_equal(coeffs, [3j, 4, -6j, 8])
    with pytest.raises(ValueError, match="max_intermediate_terms"):
        lhs.to_device().matmul(rhs.to_device(), max_intermediate_terms=3)
    assert (PauliSum.empty(3) @ PauliSum.from_labels(["XYZ"])).num_terms == 0


@pytest.mark.parametrize("case", KATS["matmul"], ids=lambda c: c["name"])
def test_matmul_reference_cases(gpu, orc, case):
    lhs = PauliSum.from_labels(*kd.campaign6_matmul_operand(case, "lhs"))
    rhs = PauliSum.from_labels(*kd.campaign6_matmul_operand(case, "rhs"))
    got = lhs.matmul(rhs, simplify=case["simplify"])
    assert got.num_terms == case["output_terms"]
    assert_bit_exact(got, orc.matmul(case["num_qubits"], packed(lhs)[1:], packed(rhs)[1:],
                                     simplify=case["simplify"]))


# ------------------------------------------------------------ simplify

@pytest.mark.parametrize("case", KATS["simplify"], ids=lambda c: c["name"])
def test_simplify_reference_canonical_hash(gpu, orc, case):
    labels, coeffs = kd.simplify_operator_labels(case)
    op = PauliSum.from_labels(labels, coeffs)
    got = op.to_device().simplify().to_host()
    out_labels, out_coeffs = got.to_labels()
    assert len(out_labels) == case["output_terms"]
    assert kd.label_hash(out_labels, out_coeffs) == case["canonical_label_hash"]
    assert_bit_exact(got, orc.simplify(*packed(op)))


@pytest.mark.parametrize("n", [1, 5, 24, 33, 64, 65, 130, 257])
@pytest.mark.parametrize("terms", [1, 100, 5000])
def test_simplify_bit_exact_random(gpu, orc, n, terms):
    pool = random_pauli_masks(n, max(1, terms // 3), 0.3, n * 31 + terms)
    rng = np.random.default_rng(n + terms)
    pick = rng.integers(0, pool.num_terms, size=terms)
    c = rng.normal(size=terms) + 1j * rng.normal(size=terms)
    c[rng.random(terms) < 0.1] = 1e-13
    op = PauliSum.from_packed(n, pool.x_words()[pick], pool.z_words()[pick], c)
    for atol, rtol in ((1e-12, 0.0), (0.0, 0.0), (0.5, 0.0), (0.0, 0.1)):
        assert_bit_exact(op.to_device().simplify(atol, rtol).to_host(), orc.simplify(*packed(op), atol, rtol))


@pytest.mark.parametrize("n", [40, 64, 130, 256])
@pytest.mark.parametrize("terms", [3000, 70000])
def test_simplify_word_ties_bit_exact(gpu, orc, n, terms):
    # Keys wider than one word sort word-major; here every word takes only a
    # few values, so long runs tie on the leading words and are ordered by the
    # later ones (and equal keys keep construction order for their sums).
    rng = np.random.default_rng(n * 7 + terms)
    W = (n + 63) // 64
    choices = []
    for w in range(W):
        width = min(64, n - 64 * w)
        top = np.uint64(1) << np.uint64(width - 1)
        mask = np.uint64((1 << width) - 1)
        choices.append(np.array([0, 1, top, (top | np.uint64(5)) & mask], dtype=np.uint64))
    x = np.stack([choices[w][rng.integers(0, 4, size=terms)] for w in range(W)], axis=1)
    z = np.stack([choices[w][rng.integers(0, 3, size=terms)] for w in range(W)], axis=1)
    c = rng.normal(size=terms) + 1j * rng.normal(size=terms)
    op = PauliSum.from_packed(n, x, z, c)
    assert_bit_exact(op.to_device().simplify(0.0, 0.0).to_host(), orc.simplify(*packed(op), 0.0, 0.0))
    # Mostly distinct keys with a tied minority.
    pool = random_pauli_masks(n, terms, 0.25, n + terms)
    xs, zs = pool.x_words().copy(), pool.z_words().copy()
    xs[: terms // 4, 0] = xs[terms // 4: terms // 2, 0]
    op2 = PauliSum.from_packed(n, xs, zs, c)
    assert_bit_exact(op2.to_device().simplify().to_host(), orc.simplify(*packed(op2)))


@pytest.mark.parametrize("n", [48, 128, 200])
@pytest.mark.parametrize("terms", [100_000, 1_200_000])
def test_simplify_few_ties_bit_exact(gpu, orc, n, terms):
    # Mostly distinct keys with a handful of ties: exact duplicates (their
    # coefficients summed in construction order, some cancelling below the
    # threshold) and keys equal in word 0 only. The tied items are refined
    # on their own list and the sums find their segments through the
    # positions that repeat their predecessor; 1.2 M terms also take the
    # one-sweep radix passes.
    rng = np.random.default_rng(n + terms)
    pool = random_pauli_masks(n, terms, 0.25, 31 * n + terms)
    xs, zs = pool.x_words().copy(), pool.z_words().copy()
    c = rng.normal(size=terms) + 1j * rng.normal(size=terms)
    dst = rng.choice(terms, size=120, replace=False)
    src, dst = dst[:60], dst[60:]
    xs[dst[:40]] = xs[src[:40]]          # exact duplicates
    zs[dst[:40]] = zs[src[:40]]
    c[dst[:10]] = -c[src[:10]]           # ... ten of which cancel
    xs[dst[40:], 0] = xs[src[40:], 0]    # equal word 0 only
    if zs.shape[1] == 1:
        zs[dst[40:], 0] ^= np.uint64(1)
    op = PauliSum.from_packed(n, xs, zs, c)
    assert_bit_exact(op.to_device().simplify().to_host(), orc.simplify(*packed(op)))


def test_simplify_signed_zero_seed_matches_oracle(gpu, orc):
    for n in (1, 65):
        pool = random_pauli_masks(n, 4, 0.6, 3)
        idx = np.arange(4096) % 4
        c = (np.arange(4096) % 13 - 6).astype(complex) * np.where(idx == 1, -1, 1)
        op = PauliSum.from_packed(n, pool.x_words()[idx], pool.z_words()[idx], c)
        assert orc.simplify_hash_path(n, op.x_words(), op.z_words())
        assert_bit_exact(op.simplify(0.0, 0.0), orc.simplify(*packed(op), 0.0, 0.0))
        few = PauliSum.from_packed(n, op.x_words()[:50], op.z_words()[:50], -np.abs(c[:50]).astype(complex) * 0 - 0.0)
        assert_bit_exact(few.simplify(0.0, 0.0), orc.simplify(*packed(few), 0.0, 0.0))


def test_simplify_edges(gpu):
    assert PauliSum.empty(4).simplify().num_terms == 0
    out = PauliSum.from_labels(["XX", "ZZ"], [1.0e-13, 0.0]).simplify()
    assert out.num_terms == 0 and out.num_qubits == 2
    out = PauliSum(0, 3).simplify(0.0, 0.0)
    assert out.num_terms == 0  # three zero identity terms sum to 0.0 -> dropped at atol 0
    ident = PauliSum._from_arrays(0, np.zeros((3, 0), np.uint64), np.zeros((3, 0), np.uint64),
                                  np.array([1.0, 2.0, 3.0], complex))
    assert ident.simplify().coeffs()[0] == 6.0
    with pytest.raises(ValueError, match="tolerances"):
        PauliSum.from_labels(["X"]).to_device().simplify(atol=float("nan"))
    # Zero-qubit products (no mask words at all), simplified and not.
    a = PauliSum._from_arrays(0, np.zeros((3, 0), np.uint64), np.zeros((3, 0), np.uint64),
                              np.array([1.0, 2.0j, -1.0]))
    b = PauliSum._from_arrays(0, np.zeros((2, 0), np.uint64), np.zeros((2, 0), np.uint64),
                              np.array([0.5, 1.0]))
    assert a.matmul(b, simplify=False).num_terms == 6
    prod = a @ b
    assert prod.num_terms == 1 and prod.coeffs()[0] == pytest.approx(1.5 * (2.0j))
