"""The library's device-memory block cache: freed buffers are reused for
requests of the same size class, results are unchanged, and the cache honours
its limit and gives memory back on request."""

import numpy as np
import pytest

import wolfgang_amd
from conftest import assert_bit_exact, packed
from wolfgang_amd.workloads import random_pauli_masks

pytestmark = pytest.mark.gpu


def test_freed_blocks_are_reused_and_results_unchanged(gpu, orc):
    wolfgang_amd.empty_cache(0)
    a = random_pauli_masks(100, 300, 0.3, 21)
    b = random_pauli_masks(100, 200, 0.3, 22)
    want = orc.matmul(100, (a.x_words(), a.z_words(), a.coeffs()), (b.x_words(), b.z_words(), b.coeffs()))
    da, db = a.to_device(), b.to_device()
    first = da.matmul(db)
    assert_bit_exact(first.to_host(), want)
    live_before = wolfgang_amd.memory_cache_info(0)["live_bytes"]
    del first
    info = wolfgang_amd.memory_cache_info(0)
    assert info["cached_bytes"] > 0 and info["cached_blocks"] >= 3
    assert info["live_bytes"] < live_before
    # Same shapes again: the output's three buffers come from the cache, and a
    # buffer reused from an earlier result must not leak old values.
    for _ in range(3):
        again = da.matmul(db)
        assert_bit_exact(again.to_host(), want)
        del again
    after = wolfgang_amd.memory_cache_info(0)
    assert after["cached_bytes"] >= info["cached_bytes"]


def test_limit_and_release(gpu):
    ops = [random_pauli_masks(64, 5000, 0.3, 30 + i).to_device() for i in range(4)]
    del ops
    assert wolfgang_amd.memory_cache_info(0)["cached_bytes"] > 0
    prev = wolfgang_amd.set_memory_cache_limit(1 << 20)
    try:
        assert wolfgang_amd.memory_cache_info(0)["cached_bytes"] <= 1 << 20
        # Blocks larger than the limit go straight back to the runtime.
        big = random_pauli_masks(64, 100_000, 0.3, 40).to_device()
        del big
        assert wolfgang_amd.memory_cache_info(0)["cached_bytes"] <= 1 << 20
    finally:
        wolfgang_amd.set_memory_cache_limit(prev)
    wolfgang_amd.empty_cache(0)
    assert wolfgang_amd.memory_cache_info(0)["cached_bytes"] == 0
    assert wolfgang_amd.memory_cache_info(0)["cached_blocks"] == 0


def test_expectation_after_reuse_matches_oracle(gpu, orc):
    # Scratch and plan buffers recycled between operators of different sizes.
    from wolfgang_amd.workloads import random_statevector

    for seed, (n, t) in enumerate([(12, 900), (14, 300), (12, 900), (13, 2000)]):
        op = random_pauli_masks(n, t, 0.3, 100 + seed)
        psi = random_statevector(n, 200 + seed)
        e = op.to_device().expectation_statevector(psi)
        want = orc.expectation_compensated(*packed(op), psi)
        assert abs(e - want) <= 1e-12 * float(np.abs(op.coeffs()).sum())
