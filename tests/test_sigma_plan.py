"""CPU replay of the tiled sigma sweep's key records (wg_debug_sigma_plan)
against the oracle: the records' sign decomposition (uniform, lane and
register parts) and the constant sign folded into the coefficients must give
sigma = H psi and H^dagger psi of the oracle (1e-12 relative), for random,
chemistry-like and diagonal-heavy operators and for a pass of 300 keys on one
x-mask. No GPU needed: the replay runs the host planner and evaluates each
record on the CPU."""

import ctypes as C

import numpy as np
import pytest

from conftest import packed
from wolfgang_amd import PauliSum, _native
from wolfgang_amd.workloads import chemistry_masks, random_pauli_masks, random_statevector


def replay(op, psi):
    n, T = op.num_qubits, op.num_terms
    x = np.ascontiguousarray(op.x_words()[:, 0])
    z = np.ascontiguousarray(op.z_words()[:, 0])
    c = np.ascontiguousarray(op.coeffs())
    sig = np.zeros(1 << n, dtype=np.complex128)
    sigd = np.zeros(1 << n, dtype=np.complex128)
    nr = C.c_size_t()
    _native.call("wg_debug_sigma_plan", n, T, x.ctypes.data, z.ctypes.data, c.ctypes.data, psi.ctypes.data,
                 sig.ctypes.data, sigd.ctypes.data, C.byref(nr))
    return sig, sigd, nr.value


def one_x_many_z(n, keys, seed):
    """`keys` distinct z-masks on one x-mask (one long pass), both parities."""
    rng = np.random.default_rng(seed)
    x = np.full((keys, 1), (1 << 3) | (1 << (n - 2)) | (1 << 9), dtype=np.uint64)
    z = np.unique(rng.integers(0, 1 << n, size=3 * keys, dtype=np.uint64))[:keys].reshape(-1, 1)
    c = rng.normal(size=keys) + 1j * rng.normal(size=keys)
    return PauliSum.from_packed(n, x, z, c)


@pytest.mark.parametrize("case", ["random18", "chemistry19", "diag_heavy18", "long_pass18", "random20",
                                  "random_wide18"])
def test_sigma_records_replay_matches_oracle(orc, case):
    if case == "random18":
        op = random_pauli_masks(18, 300, 0.3, 81)
    elif case == "chemistry19":
        op = chemistry_masks(19, 600, 82)
    elif case == "diag_heavy18":
        r = random_pauli_masks(18, 200, 0.4, 83)
        xs = r.x_words().copy()
        xs[::2] = 0  # half the terms diagonal: the diagonal items ride in pass 0
        op = PauliSum.from_packed(18, xs, r.z_words(), r.coeffs())
    elif case == "long_pass18":
        op = one_x_many_z(18, 300, 84)
    elif case == "random_wide18":
        # More than 1,024 distinct x-masks: the greedy coset planner.
        op = random_pauli_masks(18, 2500, 0.3, 86)
    else:
        op = random_pauli_masks(20, 120, 0.25, 85)
    n = op.num_qubits
    psi = random_statevector(n, 90)
    sig, sigd, records = replay(op, psi)
    threads = 8
    want = orc.apply(*packed(op), psi, threads=threads)
    want_d = orc.apply(n, op.x_words(), op.z_words(), np.conj(op.coeffs()), psi, threads=threads)
    scale = float(np.abs(op.coeffs()).sum())
    assert np.max(np.abs(sig - want)) <= 1e-12 * scale
    assert np.max(np.abs(sigd - want_d)) <= 1e-12 * scale
    keys, _ = orc.distinct_keys(n, op.x_words(), op.z_words())
    assert keys <= records < keys + 4 * (keys + 1)
