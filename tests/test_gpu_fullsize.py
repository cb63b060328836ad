"""GPU parity at the stated size of every workload configuration (MI355X).

Each case builds the workload exactly as the benchmark does (same generators,
seeds and sizes) and checks the device results against the CPU oracle, run on
the host's cores:

* 16 qubits, 2,000 terms (each qubit I w.p. 0.75): H @ H with simplify,
  bit-exact (uint64 views of masks and coefficients); the 2,000^2 commutation
  matrix, bit-exact; the expectation within 1e-12 * sum|c| * ||psi||^2 of the
  oracle's compensated sum.
* 128 qubits: the 65,536^2 packed commutation matrix bit-for-bit against the
  oracle's row-threaded kernel, its dense uint8 view and its row / column /
  total counts; the 8,192^2 products with simplify, bit-exact.
* 30 qubits, 100,000 chemistry-like terms: per-key <P_k> of the full operator
  (owner sweep) on a stratified sample -- at least one key of every tiled pass
  and every owner-bundle kind, plus diagonal keys (Walsh-Hadamard sweep) --
  within 1e-12 * ||psi||^2 of the oracle's compensated values; E equals the
  per-term values summed with the coefficients, and E is within 1e-12 * sum|c|
  of the oracle's x-grouped double-double evaluation of every key.
* 24 qubits, 20,000 random terms: central finite differences (h = 1e-6) of a
  real loss through the autograd function on 4 coefficients and 4 amplitudes
  (real and imaginary parts), the same gradients against the oracle's
  scalar-path values (<P_k> and sigma rows), and sigma / sigma^dagger rows.
* The 256-instance batch (n in 12..22, 100..50,000 terms): a subset of
  instances against the oracle and the plain per-instance path; the smallest
  pairs of the 64-product batch bit-exact against the oracle.

Tolerances are stated per check. Host threads for the oracle: min(16, cores).
"""

import math
import os

import numpy as np
import pytest

from conftest import assert_bit_exact, packed
from wolfgang_amd import _native
from wolfgang_amd.workloads import (
    batch_shapes,
    chemistry_masks,
    product_shapes,
    random_pauli_masks,
    random_statevector,
    random_statevector_device,
)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
RTOL = 1e-12


def say(msg):
    print(f"[fullsize] {msg}", flush=True)


def scale(c, psi_norm2=1.0):
    return float(np.abs(c).sum()) * psi_norm2


def test_c1_plain_config_in_full(gpu, orc):
    n, T = 16, 2_000
    op = random_pauli_masks(n, T, 0.25, 1001)
    psi = random_statevector(n, 1002)
    d = op.to_device()
    say("C1 H@H + simplify")
    prod = d.matmul(d).to_host()
    assert_bit_exact(prod, orc.matmul(n, (op.x_words(), op.z_words(), op.coeffs()),
                                      (op.x_words(), op.z_words(), op.coeffs())))
    say("C1 commutation 2000^2")
    m = d.commutes_with(d)
    want = orc.commutes_with(n, (op.x_words(), op.z_words()), (op.x_words(), op.z_words()),
                             max_entries=T * T, variant=2, threads=THREADS)
    np.testing.assert_array_equal(m, want.astype(bool))
    say("C1 expectation")
    e = d.expectation_statevector(psi)
    ref = orc.expectation_compensated(*packed(op), psi, threads=THREADS)
    assert abs(e - ref) <= RTOL * scale(op.coeffs())


@pytest.mark.timeout(300)
def test_c2_commutation_65536_squared_bit_exact(gpu, orc):
    n, T = 128, 65_536
    op = random_pauli_masks(n, T, 0.25, 2001)
    d = op.to_device()
    say("C2 commutation 65536^2 on the device")
    m = d.commutes_with_device(d, max_commutation_matrix_entries=T * T)
    got = m.to_host_packed()
    say("C2 commutation oracle (row-threaded)")
    want = orc.commutes_packed(n, (op.x_words(), op.z_words()), (op.x_words(), op.z_words()),
                               threads=THREADS)
    assert got.shape == want.shape == (T, T // 64)
    assert np.array_equal(got, want), "packed commutation bits differ from the oracle"
    rows = np.bitwise_count(want).sum(axis=1, dtype=np.uint64)
    say("C2 counts")
    assert m.count_commuting() == int(rows.sum())
    np.testing.assert_array_equal(m.count_commuting(axis=1), rows)
    # The operator against itself gives a symmetric matrix: column counts
    # equal row counts.
    np.testing.assert_array_equal(m.count_commuting(axis=0), rows)
    np.testing.assert_array_equal(m.conflict_degrees(axis=1), T - rows)
    del got
    say("C2 dense uint8 view (4 GiB)")
    dense = m.to_host()
    bits = np.unpackbits(want.view(np.uint8), axis=1, bitorder="little")
    assert np.array_equal(dense.view(np.uint8), bits), "dense view differs from the packed bits"


@pytest.mark.timeout(300)
def test_c2_products_8192_squared_simplify_bit_exact(gpu, orc):
    n, T = 128, 8_192
    a = random_pauli_masks(n, T, 0.25, 2002)
    b = random_pauli_masks(n, T, 0.25, 2003)
    say("C2 products 8192^2 + simplify on the device")
    got = a.to_device().matmul(b.to_device(), simplify=True, max_intermediate_terms=T * T).to_host()
    say("C2 products oracle")
    orc.set_threads(THREADS)
    try:
        want = orc.matmul(n, (a.x_words(), a.z_words(), a.coeffs()), (b.x_words(), b.z_words(), b.coeffs()),
                          simplify=True, max_terms=T * T)
    finally:
        orc.set_threads(1)
    assert_bit_exact(got, want)


def _key_sample(op, per_cell=2, diag=8):
    """Term indices of a stratified key sample: `per_cell` keys of every
    (tiled pass, owner-bundle kind) cell and `diag` diagonal keys."""
    import ctypes as C

    x = np.ascontiguousarray(op.x_words()[:, 0])
    z = np.ascontiguousarray(op.z_words()[:, 0])
    T = x.shape[0]
    kp = np.zeros(T, dtype=np.int32)
    kk = np.zeros(T, dtype=np.int32)
    nk = C.c_size_t()
    _native.call("wg_debug_key_sweeps", op.num_qubits, T, x.ctypes.data, z.ctypes.data, kp.ctypes.data,
                 kk.ctypes.data, C.byref(nk))
    K = nk.value
    # First term of each key (keys are numbered in first-appearance order).
    keys_seen, term_of_key = {}, []
    for t in range(T):
        kxz = (int(x[t]), int(z[t]))
        if kxz not in keys_seen:
            keys_seen[kxz] = len(term_of_key)
            term_of_key.append(t)
    assert len(term_of_key) == K
    cells = {}
    for k in range(K):
        cells.setdefault((int(kp[k]), int(kk[k])), []).append(k)
    picked = []
    for (p, kind), ks in sorted(cells.items()):
        if p < 0:
            picked += ks[:diag]
        else:
            picked += ks[:per_cell]
    passes = {p for (p, _) in cells if p >= 0}
    return np.array([term_of_key[k] for k in picked], dtype=np.int64), len(passes), len(cells)


@pytest.mark.timeout(300)
def test_c3_expectation_30_qubits_stratified_keys(gpu, orc):
    import torch

    n, T = 30, 100_000
    op = chemistry_masks(n, T, 3001)
    d = op.to_device()
    say("C3 psi (2^30 amplitudes) on the device")
    psi = random_statevector_device(n, 3002, 0)
    tv = torch.empty(T, dtype=torch.complex128, device=psi.device)
    stream = torch.cuda.current_stream().cuda_stream
    say("C3 owner sweep with per-term values")
    e_adj = d.expectation_adjoint(psi.data_ptr(), psi.numel(), tv.data_ptr(), None, None, stream=stream)
    e = d.expectation_statevector(psi)
    assert e == e_adj, "the per-term path and the plain path give different bits"
    info = d.expectation_plan_info()
    assert info["tiled"]
    sample, passes, cells = _key_sample(op)
    assert passes == info["sweeps"], "the sample misses a pass"
    say(f"C3 oracle on {sample.size} sampled terms ({cells} pass/kind cells), psi to host")
    psi_h = psi.cpu().numpy()
    del psi
    want = orc.term_values(n, op.x_words()[sample], op.z_words()[sample], psi_h, threads=THREADS)
    got = tv.cpu().numpy()
    # <P_k> is real for the Hermitian P_k; the device writes (value, 0).
    assert np.all(got.imag == 0.0)
    err = np.abs(got.real[sample] - want)
    say(f"C3 max per-key error {err.max():.3e}")
    assert err.max() <= RTOL, f"per-key error {err.max()}"
    # E is the coefficient-weighted sum of the per-term values (tolerance:
    # 1e-12 of sum |c|).
    c = op.coeffs()
    tot = complex(math.fsum((c.real * got.real).tolist()), math.fsum((c.imag * got.real).tolist()))
    assert abs(e - tot) <= RTOL * scale(c)
    say("C3 full energy: every key by the oracle's x-grouped evaluation")
    want_e = orc.expectation_grouped(n, op.x_words(), op.z_words(), c, psi_h, threads=THREADS, pieces=8)
    say(f"C3 energy error {abs(e - want_e):.3e}")
    assert abs(e - want_e) <= RTOL * scale(c)


@pytest.mark.timeout(300)
def test_c4_autograd_24_qubits_finite_differences(gpu, orc):
    import torch

    from wolfgang_amd.autograd import expectation_value

    n, T = 24, 20_000
    op = random_pauli_masks(n, T, 0.25, 4001)
    d = op.to_device()
    psi0 = random_statevector_device(n, 4002, 0)
    c0 = torch.tensor(op.coeffs(), device=psi0.device)
    psi = psi0.clone().requires_grad_(True)
    c = c0.clone().requires_grad_(True)
    w = 0.3  # loss = Re E + w Im E, so grad_output g = 1 + i w
    say("C4 forward + backward through autograd")
    E = expectation_value(d, psi, c)
    (E.real + w * E.imag).backward()
    g_psi, g_c = psi.grad.detach(), c.grad.detach()

    def loss(pv, cv):
        with torch.no_grad():
            ev = expectation_value(d, pv, cv)
        return ev.real.item() + w * ev.imag.item()

    h = 1e-6
    rng = np.random.default_rng(4003)
    ks = rng.choice(T, 4, replace=False)
    js = rng.choice(1 << n, 4, replace=False)
    say("C4 finite differences")
    for k in ks:
        for unit, part in ((1.0, "real"), (1j, "imag")):
            cp, cm = c0.clone(), c0.clone()
            cp[k] += h * unit
            cm[k] -= h * unit
            fd = (loss(psi0, cp) - loss(psi0, cm)) / (2 * h)
            an = getattr(g_c[k], part).item()
            # E is linear in c: the difference is exact up to the rounding of
            # the coefficient-weighted sum.
            assert abs(fd - an) <= 1e-8 + 1e-6 * abs(an), (k, part, fd, an)
    for j in js:
        for unit, part in ((1.0, "real"), (1j, "imag")):
            pp, pm = psi0.clone(), psi0.clone()
            pp[j] += h * unit
            pm[j] -= h * unit
            fd = (loss(pp, c0) - loss(pm, c0)) / (2 * h)
            an = getattr(g_psi[j], part).item()
            # E is quadratic in psi: central differences are exact up to the
            # rounding of E (~1e-15 absolute here) divided by 2h.
            assert abs(fd - an) <= 1e-7 + 1e-6 * abs(an), (j, part, fd, an)
    say("C4 scalar-path gradients (oracle)")
    psi_h = psi0.cpu().numpy()
    coeffs = op.coeffs()
    g = 1.0 + 1j * w
    tv = orc.term_values(n, op.x_words()[ks], op.z_words()[ks], psi_h, threads=THREADS)
    np.testing.assert_allclose(g_c[ks].cpu().numpy(), g * np.conj(tv), rtol=0, atol=1e-12)
    rows = np.sort(js).astype(np.uint64)
    sig = orc.apply_rows(n, op.x_words(), op.z_words(), coeffs, psi_h, rows, threads=THREADS)
    sig_dag = orc.apply_rows(n, op.x_words(), op.z_words(), np.conj(coeffs), psi_h, rows, threads=THREADS)
    want = np.conj(g) * sig + g * sig_dag
    sc = scale(coeffs) * float(np.abs(psi_h).max())
    np.testing.assert_allclose(g_psi[rows.astype(np.int64)].cpu().numpy(), want, rtol=0, atol=RTOL * sc)
    say("C4 sigma and sigma^dagger rows")
    sigma = torch.empty_like(psi0)
    sigma_d = torch.empty_like(psi0)
    tvd = torch.empty(T, dtype=torch.complex128, device=psi0.device)
    stream = torch.cuda.current_stream().cuda_stream
    d.expectation_adjoint(psi0.data_ptr(), psi0.numel(), tvd.data_ptr(), sigma.data_ptr(), sigma_d.data_ptr(),
                          stream=stream)
    rows = np.sort(rng.choice(1 << n, 48, replace=False)).astype(np.uint64)
    idx = torch.from_numpy(rows.astype(np.int64)).to(psi0.device)
    np.testing.assert_allclose(
        sigma[idx].cpu().numpy(),
        orc.apply_rows(n, op.x_words(), op.z_words(), coeffs, psi_h, rows, threads=THREADS),
        rtol=0, atol=RTOL * sc)
    np.testing.assert_allclose(
        sigma_d[idx].cpu().numpy(),
        orc.apply_rows(n, op.x_words(), op.z_words(), np.conj(coeffs), psi_h, rows, threads=THREADS),
        rtol=0, atol=RTOL * sc)
    kt = rng.choice(T, 32, replace=False)
    np.testing.assert_allclose(
        tvd[torch.from_numpy(kt).to(psi0.device)].cpu().numpy().real,
        orc.term_values(n, op.x_words()[kt], op.z_words()[kt], psi_h, threads=THREADS), rtol=0, atol=RTOL)


@pytest.mark.timeout(300)
def test_c5_batch_256_instances_subset_vs_oracle(gpu, orc):
    import torch

    from wolfgang_amd.batch import BatchedExpectation, plain_forward_adjoint

    shapes = batch_shapes(256, 5001)
    ops, psis = [], []
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(5002)
    for i, (n, T) in enumerate(shapes):
        ops.append(random_pauli_masks(n, T, 0.25, 5003 + i))
        p = torch.randn(1 << n, dtype=torch.complex128, device="cuda:0", generator=gen)
        psis.append(p / torch.linalg.vector_norm(p))
    dops = [op.to_device() for op in ops]
    say("C5 batch creation (with autotuning) and one replay")
    b = BatchedExpectation(dops, psis, sigma=True, sigma_dagger=True, use_cache=False)
    b.run()
    torch.cuda.synchronize()
    e_batch = b.energies.cpu().numpy().copy()
    tv_b = [t.cpu().numpy() for t in b.term_values]
    rng = np.random.default_rng(5004)
    work = np.array([T * (1 << n) for n, T in shapes], dtype=np.float64)
    order = np.argsort(work)
    tiled = [i for i in order if shapes[i][0] >= 18]
    pick = list(order[:3]) + list(order[120:122]) + tiled[:2] + tiled[-1:]
    for i in pick:
        n, T = shapes[i]
        op = ops[i]
        psi_h = psis[i].cpu().numpy()
        c = op.coeffs()
        sc = scale(c)
        say(f"C5 instance {i}: n={n}, T={T}")
        rows = np.sort(rng.choice(1 << n, 24, replace=False)).astype(np.uint64)
        idx = torch.from_numpy(rows.astype(np.int64)).to("cuda:0")
        np.testing.assert_allclose(
            b.sigma[i][idx].cpu().numpy(), orc.apply_rows(n, op.x_words(), op.z_words(), c, psi_h, rows,
                                                          threads=THREADS), rtol=0, atol=RTOL * sc)
        np.testing.assert_allclose(
            b.sigma_dagger[i][idx].cpu().numpy(),
            orc.apply_rows(n, op.x_words(), op.z_words(), np.conj(c), psi_h, rows, threads=THREADS),
            rtol=0, atol=RTOL * sc)
        kt = rng.choice(T, min(T, 24), replace=False)
        np.testing.assert_allclose(
            tv_b[i].real[kt], orc.term_values(n, op.x_words()[kt], op.z_words()[kt], psi_h, threads=THREADS),
            rtol=0, atol=RTOL)
        if T * (1 << n) <= 2e9:
            ref = orc.expectation_compensated(*packed(op), psi_h, threads=THREADS)
        else:
            # E against the device's own per-term values (each sampled term is
            # checked against the oracle above).
            tvr = tv_b[i].real
            ref = complex(math.fsum((c.real * tvr).tolist()), math.fsum((c.imag * tvr).tolist()))
        assert abs(e_batch[i] - ref) <= RTOL * sc, (i, e_batch[i], ref)
    say("C5 plain per-instance path on the same buffers")
    plain_forward_adjoint(b)
    torch.cuda.synchronize()
    e_plain = b.energies.cpu().numpy()
    sc_all = np.array([scale(op.coeffs()) for op in ops])
    assert np.all(np.abs(e_plain - e_batch) <= RTOL * sc_all)


def test_c5_product_batch_smallest_pairs_vs_oracle(gpu, orc):
    from wolfgang_amd.batch import ProductBatch, plain_products

    shapes = product_shapes(64, 5001)
    order = sorted(range(len(shapes)), key=lambda i: shapes[i][1] * shapes[i][2])[:6]
    lhs, rhs, host = [], [], []
    for i in order:
        n, tl, tr = shapes[i]
        a = random_pauli_masks(n, tl, 0.25, 6000 + 2 * i)
        b = random_pauli_masks(n, tr, 0.25, 6001 + 2 * i)
        host.append((n, a, b))
        lhs.append(a.to_device())
        rhs.append(b.to_device())
    say(f"C5 product batch of the {len(order)} smallest pairs")
    pb = ProductBatch(lhs, rhs, max_intermediate_terms=1 << 40)
    res = pb.run(simplify=True)
    plain = plain_products(lhs, rhs, simplify=True, max_intermediate_terms=1 << 40)
    orc.set_threads(THREADS)
    try:
        for k, (n, a, b) in enumerate(host):
            got = res[k].to_host()
            p = plain[k].to_host()
            np.testing.assert_array_equal(got.x_words(), p.x_words())
            np.testing.assert_array_equal(got.coeffs().view(np.uint64), p.coeffs().view(np.uint64))
            if k < 2:
                assert_bit_exact(got, orc.matmul(n, (a.x_words(), a.z_words(), a.coeffs()),
                                                 (b.x_words(), b.z_words(), b.coeffs()), simplify=True,
                                                 max_terms=1 << 40))
    finally:
        orc.set_threads(1)
