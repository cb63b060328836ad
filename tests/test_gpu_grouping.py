"""Greedy commuting groups on the device against the oracle restatement and
the reference's fixtures (reference tests/test_phase6_commutation_grouping.py)."""

import numpy as np
import pytest

from wolfgang_amd import PauliSum
from wolfgang_amd.workloads import chemistry_masks, random_pauli_masks

pytestmark = pytest.mark.gpu


def labels_of(groups):
    return [g.to_labels()[0] for g in groups]


def test_grouping_reference_fixture(gpu):
    op = PauliSum.from_labels(["II", "XI", "XX", "IZ"], [1.0, 2.0, 3.0, 4.0])
    groups = op.group_commuting(mode="qwc")
    assert labels_of(groups) == [["XX", "XI", "II"], ["IZ"]]
    np.testing.assert_array_equal(groups[0].coeffs(), np.asarray([3.0, 2.0, 1.0], dtype=np.complex128))
    np.testing.assert_array_equal(groups[1].coeffs(), np.asarray([4.0], dtype=np.complex128))
    op = PauliSum.from_labels(["XI", "IX", "XX", "YY", "ZI", "IZ", "II"], [1.0, -2.0, 3.0, 4.0, 5.0, -6.0, 7.0])
    for mode in ("qwc", "full"):
        a = labels_of(op.group_commuting(mode=mode))
        assert a == labels_of(op.group_commuting(mode=mode))
        assert a == labels_of(op.group_commuting(mode=mode, max_terms_for_graph=0))
        assert sorted(sum(a, [])) == sorted(op.to_labels()[0])
    assert PauliSum.empty(3).group_commuting() == []


@pytest.mark.parametrize("mode", ["qwc", "full"])
@pytest.mark.parametrize("n,terms", [(1, 5), (6, 300), (33, 2000), (64, 1500), (70, 800), (130, 600), (200, 300)])
def test_grouping_matches_oracle(gpu, orc, mode, n, terms):
    pool = random_pauli_masks(n, max(1, terms // 2), 0.3, n * 17 + terms)
    rng = np.random.default_rng(n * 3 + terms)
    pick = rng.integers(0, pool.num_terms, size=terms)
    c = rng.normal(size=terms) + 1j * rng.normal(size=terms)
    op = PauliSum.from_packed(n, pool.x_words()[pick], pool.z_words()[pick], c)
    want = orc.group_commuting(n, op.x_words(), op.z_words(), mode)
    got = op.group_commuting(mode=mode)
    assert len(got) == len(want)
    for g, idx in zip(got, want):
        np.testing.assert_array_equal(g.x_words(), op.x_words()[idx])
        np.testing.assert_array_equal(g.z_words(), op.z_words()[idx])
        np.testing.assert_array_equal(g.coeffs().view(np.uint64), op.coeffs()[idx].view(np.uint64))


def test_grouping_chemistry_full(gpu, orc):
    op = chemistry_masks(16, 6000, 11)
    want = orc.group_commuting(16, op.x_words(), op.z_words(), "full")
    got = op.group_commuting(mode="full")
    assert [g.num_terms for g in got] == [len(w) for w in want]
    for g, idx in zip(got, want):
        np.testing.assert_array_equal(g.x_words(), op.x_words()[idx])


def _masks(n, letters):
    W = (n + 63) // 64
    x = np.zeros(W, np.uint64)
    z = np.zeros(W, np.uint64)
    for q, c in letters.items():
        if c in "XY":
            x[q // 64] |= np.uint64(1 << (q % 64))
        if c in "ZY":
            z[q // 64] |= np.uint64(1 << (q % 64))
    return x, z


def _pauli_product(n, factors):
    x = np.zeros(n, bool)
    z = np.zeros(n, bool)
    for f in factors:
        for q, c in f.items():
            x[q] ^= c in "XY"
            z[q] ^= c in "ZY"
    return {q: "Y" if x[q] and z[q] else ("X" if x[q] else "Z") for q in range(n) if x[q] or z[q]}


def scan_case(mode):
    """17 heavy terms h_0..h_16 fill the first 1,024 visiting positions (copies,
    heaviest first), each opening its own group g_k. The next positions hold
    u_15..u_0, where u_k conflicts with h_j for j < k and not with h_k (so it
    joins g_k), and finally t and t2, which every g_k accepts before their batch
    but which conflict with every u_k: t's first 15 candidate groups are
    all taken by conflicting u's, so it reaches g_16 only by scanning on; t2 also
    conflicts with t and opens a group. Returns (n, x, z)."""
    if mode == "qwc":
        n, d = 59, 17
        shared, pad = range(18, 22), range(22, 59)
        heavy = []
        for k in range(17):
            h = {k: "X"}
            v = k
            for q in shared:  # distinct full-support words: the h_k conflict pairwise
                h[q] = "XYZ"[v % 3]
                v //= 3
            h.update({q: "X" for q in list(pad)[: 20 + 16 - k]})
            heavy.append(h)
        us = []
        for k in range(16):
            u = {j: "Z" for j in range(k)}
            u[k] = "X"
            u[d] = "Z"
            us.append(u)
    else:
        n, d, pad = 70, 9, range(10, 70)

        def majorana(k):  # pairwise anticommuting strings on qubits 0..8
            h = {q: "Z" for q in range(k // 2)}
            h[k // 2] = "X" if k % 2 == 0 else "Y"
            return h

        heavy = []
        for k in range(17):
            h = majorana(k)
            h.update({q: "X" for q in list(pad)[: 60 - 2 * k]})
            heavy.append(h)
        us = []
        for k in range(16):
            # An even product anticommutes with majorana(j) iff j is a factor.
            u = _pauli_product(n, [majorana(j) for j in list(range(k)) + ([17] if k % 2 else [])])
            u[d] = "Z"
            us.append(u)
    copies = [64] + [60] * 16
    terms = [h for k, h in enumerate(heavy) for _ in range(copies[k])] + us + [{d: "X"}, {d: "Y"}]
    xs, zs = zip(*(_masks(n, t) for t in terms))
    return n, np.array(xs), np.array(zs)


@pytest.mark.parametrize("mode", ["qwc", "full"])
def test_grouping_scan_past_candidates(gpu, orc, mode):
    n, x, z = scan_case(mode)
    T = x.shape[0]
    op = PauliSum.from_packed(n, x, z, np.ones(T, np.complex128))
    want = orc.group_commuting(n, op.x_words(), op.z_words(), mode)
    got = op.group_commuting(mode=mode)
    assert len(want) == 18
    # t joins the 17th heavy term's group; t2 opens the last group alone.
    assert T - 2 in list(want[16]) and list(want[17]) == [T - 1]
    assert [g.num_terms for g in got] == [len(w) for w in want]
    for g, idx in zip(got, want):
        np.testing.assert_array_equal(g.x_words(), op.x_words()[idx])
        np.testing.assert_array_equal(g.z_words(), op.z_words()[idx])


@pytest.mark.parametrize("mode,n", [("qwc", 24), ("full", 12)])
def test_grouping_many_groups_global_stamps(gpu, orc, mode, n):
    # More than 32,768 terms: group ids may exceed the resolve kernel's LDS
    # stamp table, so its stamps live in device memory.
    op = random_pauli_masks(n, 33_000, 0.3, 4100 + n)
    want = orc.group_commuting(n, op.x_words(), op.z_words(), mode)
    got = op.group_commuting(mode=mode)
    assert [g.num_terms for g in got] == [len(w) for w in want]
    for g, idx in zip(got, want):
        np.testing.assert_array_equal(g.x_words(), op.x_words()[idx])
        np.testing.assert_array_equal(g.z_words(), op.z_words()[idx])


@pytest.mark.parametrize("mode", ["qwc", "full"])
def test_grouping_benchmark_operator(gpu, orc, mode):
    # The 16,384-term, 128-qubit operator the grouping benchmark times.
    op = random_pauli_masks(128, 16_384, 0.25, 2004)
    want = orc.group_commuting(128, op.x_words(), op.z_words(), mode)
    got = op.group_commuting(mode=mode)
    assert [g.num_terms for g in got] == [len(w) for w in want]
    for g, idx in zip(got, want):
        np.testing.assert_array_equal(g.x_words(), op.x_words()[idx])
        np.testing.assert_array_equal(g.z_words(), op.z_words()[idx])


@pytest.mark.parametrize("mode", ["qwc", "full"])
@pytest.mark.parametrize("terms", [3, 9000])
def test_grouping_zero_qubits(gpu, orc, mode, terms):
    # No mask words: every term is the identity and all share one group.
    op = PauliSum.from_labels([""] * terms, np.arange(1, terms + 1, dtype=np.float64))
    want = orc.group_commuting(0, op.x_words(), op.z_words(), mode)
    got = op.group_commuting(mode=mode)
    assert len(want) == 1 and len(got) == 1
    np.testing.assert_array_equal(got[0].coeffs(), op.coeffs()[want[0]])
