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
