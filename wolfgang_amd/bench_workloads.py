"""Secondary benchmark workloads for bench.py (one JSON line each).

* commutation: 128-qubit, 65,536-term operator (each qubit non-identity w.p.
  0.25) against itself -> 65,536^2 packed commutation matrix on the device.
* matmul: two 8,192-term 128-qubit operators -> 67,108,864 products, then the
  sort-reduce simplify (atol 1e-12).
* batch: 256 (statevector, Hamiltonian) pairs, n ~ U{12..22}, terms
  log-uniform in [100, 50,000] (random masks, density 0.25, complex
  coefficients): forward + adjoint in one graph replay, against the
  per-instance path on the same buffers.
"""

from __future__ import annotations

import json
import time

import numpy as np


def _timed(fn, steps: int, warmup: int) -> float:
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def commutation(args) -> dict:
    from . import _native
    from .workloads import random_pauli_masks

    n, T = 128, 65_536
    op = random_pauli_masks(n, T, 0.25, 2001)
    d = op.to_device()
    m = d.commutes_with_device(d, max_commutation_matrix_entries=T * T)
    _native.profile_enable(True)
    _native.profile_reset()
    dt = _timed(lambda: d.commutes_with_device(d, T * T, output=m), args.steps, args.warmup)
    ms, launches = _native.profile_get("commute_packed")
    _native.profile_enable(False)
    kernel_s = ms / 1000.0 / max(1, launches)
    alg_bytes = 2 * (T * 2 * 2 * 8) + T * T // 8
    count = m.count_commuting()
    return {
        "metric": "commutation pairs/s (65,536^2, 128 qubits, packed bit output)",
        "value": T * T / dt,
        "unit": "pairs/s",
        "ms_per_step": dt * 1000,
        "kernel_ms": kernel_s * 1000,
        "config": {"workload": "commutation", "num_qubits": n, "num_terms": T, "seed": 2001,
                   "commuting_count": count},
        "roofline": {"bound": "integer VALU (HBM floor shown)", "achieved": alg_bytes / kernel_s / 1e9,
                     "peak": 8000.0, "unit": "GB/s", "frac": alg_bytes / kernel_s / 1e9 / 8000.0,
                     "algorithmic_bytes": alg_bytes},
    }


def matmul(args) -> dict:
    from . import _native
    from .workloads import random_pauli_masks

    n, T = 128, 8_192
    a = random_pauli_masks(n, T, 0.25, 2002).to_device()
    b = random_pauli_masks(n, T, 0.25, 2003).to_device()
    P = T * T
    _native.profile_enable(True)
    _native.profile_reset()
    out = {}

    def step():
        out["r"] = a.matmul(b, simplify=True, max_intermediate_terms=P)

    dt = _timed(step, args.steps, args.warmup)
    calls = max(1, args.steps + args.warmup)
    U = out["r"].num_terms
    # Algorithmic HBM bytes per launch of each timed kernel: the key pass
    # writes a (word-0 key, index) record per product (the operands stay in
    # L2); a radix scatter pass reads and writes every 12-byte record; the
    # emit reads the sorted indices and segment starts and writes each output
    # term's two 2-word masks and its coefficient.
    per_launch_bytes = {"product_keys": 12 * P, "radix_scatter": 24 * P,
                        "product_emit": 8 * P + 48 * U}
    kernels = {}
    for name, nbytes in per_launch_bytes.items():
        ms, launches = _native.profile_get(name)
        if launches:
            avg = ms / launches
            kernels[name] = {"ms_per_step": ms / calls, "launches_per_step": launches / calls,
                             "avg_ms": avg, "GBs": nbytes / (avg / 1000.0) / 1e9}
    _native.profile_enable(False)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    achieved = kernels[dom]["GBs"]
    return {
        "metric": "term-pair products/s with simplify (8,192 x 8,192, 128 qubits)",
        "value": P / dt,
        "unit": "products/s",
        "ms_per_step": dt * 1000,
        "config": {"workload": "matmul+simplify", "num_qubits": n, "lhs_terms": T, "rhs_terms": T,
                   "products": P, "output_terms": U},
        "kernels": kernels,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": 8000.0,
                     "unit": "GB/s", "frac": achieved / 8000.0,
                     "algorithmic_bytes_per_launch": per_launch_bytes[dom]},
    }


def batch_speedup(batched_steps: int = 3, plain_steps: int = 2, warmup: int = 1,
                  breakdown: bool = False) -> dict:
    """256 (statevector, Hamiltonian) pairs, n ~ U{12..22}, terms log-uniform
    in [100, 50,000] (random masks, density 0.25, complex coefficients):
    forward + adjoint (E, per-term values, sigma and sigma^dagger) in one graph
    replay against the plain per-instance path on the same buffers. The
    energies of the two paths are compared relative to sum |c| per instance."""
    import torch

    from .batch import BatchedExpectation, plain_forward_adjoint
    from .workloads import batch_shapes, random_pauli_masks

    shapes = batch_shapes(256, 5001)
    ops, psis, l1 = [], [], []
    g = torch.Generator(device="cuda:0")
    g.manual_seed(5002)
    for i, (n, T) in enumerate(shapes):
        op = random_pauli_masks(n, T, 0.25, 5003 + i)
        l1.append(float(np.abs(op.coeffs()).sum()))
        ops.append(op.to_device())
        p = torch.randn(1 << n, dtype=torch.complex128, device="cuda:0", generator=g)
        psis.append(p / torch.linalg.vector_norm(p))
    t0 = time.perf_counter()
    b = BatchedExpectation(ops, psis, sigma=True, sigma_dagger=True)
    setup_s = time.perf_counter() - t0
    batched = _timed(b.run, batched_steps, warmup)
    e_batched = b.energies.clone()
    # The same batch with every launch inside the one graph (one hipGraphLaunch
    # per run), for comparison with the autotuned choice.
    forced = list(b.tunables[:4]) + [1]
    if b.tiled_instances and b.tunables[4] == 0:
        g_all = BatchedExpectation(ops, psis, sigma=True, sigma_dagger=True, tunables=forced, use_cache=False)
        all_in_graph_ms = _timed(g_all.run, batched_steps, warmup) * 1000
        del g_all
    else:
        all_in_graph_ms = batched * 1000
    plain = _timed(lambda: plain_forward_adjoint(b), plain_steps, warmup)
    diff = torch.abs(b.energies - e_batched).cpu().numpy()
    pairs = sum(T * (1 << n) for n, T in shapes)
    extra = {"per_qubit_count": _per_instance_breakdown(b, shapes)} if breakdown else {}
    return {
        **extra,
        "tiled_instances": b.tiled_instances,
        "speedup": plain / batched,
        "batched_ms": batched * 1000,
        "batched_ms_all_in_graph": all_in_graph_ms,
        "plain_ms": plain * 1000,
        "max_abs_energy_difference": float(diff.max()),
        "max_rel_energy_difference": float(np.max(diff / np.asarray(l1))),
        "term_amp_pairs_per_s_batched": pairs / batched,
        "setup_s": setup_s,
        "tunables": b.tunables,
        "units": b.units,
        "steps": {"batched": batched_steps, "plain": plain_steps, "warmup": warmup},
        "config": {"instances": len(shapes), "seed": 5001, "qubits": [12, 22], "terms": [100, 50000],
                   "generator": "random masks, density 0.25, complex coefficients",
                   "sum_terms": sum(T for _, T in shapes), "outputs": "E, <P_k>, sigma, sigma^dagger"},
    }


def _per_instance_breakdown(b, shapes) -> dict:
    """Plain-path time of every instance on its own (synchronized before and
    after), grouped by qubit count, with whether its plan is tiled."""
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    rows = {}
    for i, (op, psi) in enumerate(zip(b.ops, b.psis)):
        n, T = shapes[i]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.expectation_adjoint(psi.data_ptr(), psi.numel(), b.term_values[i].data_ptr(), b.sigma[i].data_ptr(),
                               b.sigma_dagger[i].data_ptr(), stream=stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        r = rows.setdefault(n, {"instances": 0, "tiled": 0, "plain_ms": 0.0, "term_amp_pairs": 0})
        r["instances"] += 1
        r["tiled"] += int(op.expectation_plan_info()["tiled"])
        r["plain_ms"] += dt * 1000
        r["term_amp_pairs"] += T * (1 << n)
    return {str(n): rows[n] for n in sorted(rows)}


def batch(args) -> dict:
    res = batch_speedup(args.steps, max(1, args.steps // 2), args.warmup, breakdown=True)
    return {
        "metric": "batched fwd+adjoint speedup vs plain per-instance GPU path",
        "value": res["speedup"],
        "unit": "x",
        **{k: v for k, v in res.items() if k not in ("speedup", "config")},
        "config": {"workload": "batch", **res["config"]},
    }


def products(args) -> dict:
    """64 batched Hamiltonian products of uneven sizes (500-8,000 terms per
    operand, 64-256 qubits, random density-0.25 terms with complex
    coefficients), simplified: the batched path (persistent queue-fed product
    kernel per graph-captured chunk, then per-pair simplify) against the plain
    per-pair matmul path on the same operands."""
    import torch

    from .batch import ProductBatch, plain_products
    from .workloads import product_shapes, random_pauli_masks

    shapes = product_shapes(64, 5001)
    lhs, rhs = [], []
    for i, (n, tl, tr) in enumerate(shapes):
        lhs.append(random_pauli_masks(n, tl, 0.25, 6000 + 2 * i).to_device())
        rhs.append(random_pauli_masks(n, tr, 0.25, 6001 + 2 * i).to_device())
    total = sum(tl * tr for _, tl, tr in shapes)
    t0 = time.perf_counter()
    pb = ProductBatch(lhs, rhs, max_intermediate_terms=1 << 40, budget_bytes=16 << 30)
    setup_s = time.perf_counter() - t0
    # The outputs hold about as many terms as there are products (~100 GB);
    # keep only the three smallest pairs' results for the bit-exact check.
    keep = sorted(range(len(shapes)), key=lambda i: shapes[i][1] * shapes[i][2])[:3]
    out = {"b": {}, "p": {}, "terms": 0}

    def batched():
        res = pb.run(simplify=True)
        out["terms"] = sum(r.num_terms for r in res)
        out["b"] = {i: res[i] for i in keep}
        del res

    def plain():
        out["p"] = {}
        for i, (a, b) in enumerate(zip(lhs, rhs)):
            r = plain_products([a], [b], simplify=True, max_intermediate_terms=1 << 40)[0]
            if i in keep:
                out["p"][i] = r
            del r

    t_b = _timed(batched, args.steps, args.warmup)
    t_p = _timed(plain, max(1, args.steps // 2), 1)
    for i in keep:
        a, b = out["b"][i].to_host(), out["p"][i].to_host()
        if not (np.array_equal(a.x_words(), b.x_words()) and np.array_equal(a.z_words(), b.z_words())
                and np.array_equal(a.coeffs().view(np.uint64), b.coeffs().view(np.uint64))):
            raise RuntimeError(f"batched product {i} differs from its matmul")
    return {
        "metric": "batched Hamiltonian products + simplify: speedup vs plain per-pair matmul",
        "value": t_p / t_b,
        "unit": "x",
        "batched_ms": t_b * 1000,
        "plain_ms": t_p * 1000,
        "products_per_s_batched": total / t_b,
        "setup_s": setup_s,
        "chunks": pb.chunks,
        "config": {"workload": "products", "pairs": len(shapes), "seed": 5001,
                   "terms": [500, 8000], "qubits": [64, 256], "total_products": total,
                   "output_terms": out["terms"]},
    }


def grouping(args) -> dict:
    """Greedy largest-first commuting groups (full commutation and qubit-wise)
    of a 16,384-term, 128-qubit random operator (density 0.25) on the device."""
    import torch

    from .workloads import random_pauli_masks

    n, T = 128, 16_384
    op = random_pauli_masks(n, T, 0.25, 2004)
    d = op.to_device()
    res = {}
    for mode in ("full", "qwc"):
        out = {}

        def step():
            out["g"] = d.group_commuting(mode)

        res[mode] = {"ms": _timed(step, max(1, args.steps // 2), 1) * 1000, "groups": len(out["g"])}
    return {
        "metric": "greedy commuting groups, 16,384 terms x 128 qubits (full mode time)",
        "value": res["full"]["ms"],
        "unit": "ms",
        "higher_is_better": False,
        "config": {"workload": "grouping", "num_qubits": n, "num_terms": T, "seed": 2004},
        "results": res,
    }


def adjoint(args) -> dict:
    """Forward + fused adjoint (sigma = H psi and per-term <P_k>) on a 24-qubit
    state with a 20,000-term Hamiltonian: the random density-0.25 generator
    with complex coefficients (primary) and the chemistry-like generator
    (secondary)."""
    import torch

    from . import _native
    from .workloads import chemistry_masks, random_pauli_masks, random_statevector_device

    n, T = 24, 20_000
    psi = random_statevector_device(n, 4002, 0)
    out = {}
    for label, op in (("random_density_0.25", random_pauli_masks(n, T, 0.25, 4001)),
                      ("chemistry", chemistry_masks(n, T, 4001))):
        d = op.to_device(0)
        x = op.x_words()[:, 0]
        gx = int(np.unique(x).size)
        tv = torch.empty(T, dtype=torch.complex128, device="cuda:0")
        sigma = torch.empty_like(psi)
        stream = torch.cuda.current_stream().cuda_stream

        def fwd():
            d.expectation_statevector(psi)

        def fwd_adj():
            d.expectation_adjoint(psi.data_ptr(), psi.numel(), tv.data_ptr(), sigma.data_ptr(), None,
                                  stream=stream)

        from .autograd import expectation_value

        c_t = torch.tensor(op.coeffs(), device="cuda:0")

        def fwd_bwd():
            # One training step through the autograd function: E, then dE
            # with respect to both psi and the coefficients.
            p = psi.detach().requires_grad_(True)
            cc = c_t.detach().requires_grad_(True)
            e = expectation_value(d, p, cc)
            (e.real + 0.3 * e.imag).backward()

        t_fwd = _timed(fwd, args.steps, args.warmup)
        t_train = _timed(fwd_bwd, args.steps, args.warmup)
        _native.profile_enable(True)
        _native.profile_reset()
        t_adj = _timed(fwd_adj, args.steps, args.warmup)
        calls = args.steps + args.warmup
        # sigma: the tiled sweep (n >= 18, one launch per pass) or the direct
        # sweep (smaller states).
        tiled_ms, tiled_launches = _native.profile_get("sigma_tiled")
        sig_ms = (tiled_ms + _native.profile_get("sigma_direct")[0]) / calls
        _native.profile_enable(False)
        amp = 1 << n
        passes = tiled_launches // calls
        # Bytes the tiled sigma sweep must move per call: every pass reads its
        # psi tiles and reads and writes sigma once (16 B each per amplitude),
        # except the first, which writes sigma without reading it.
        sig_bytes = (48 * passes - 16) * amp if passes else None
        out[label] = {
            "distinct_x_groups": gx,
            "plan": d.expectation_plan_info(),
            "forward_ms": t_fwd * 1000,
            "forward_plus_adjoint_ms": t_adj * 1000,
            "autograd_forward_backward_ms": t_train * 1000,
            "sigma_kernels_ms": sig_ms,
            "sigma_launches_per_call": passes,
            "sigma_bytes_per_call": sig_bytes,
            "sigma_GBs": sig_bytes / (sig_ms / 1000) / 1e9 if sig_bytes and sig_ms else None,
            "forward_term_amp_pairs_per_s": T * amp / t_fwd,
            "forward_definition_equivalent_GBs": gx * amp * 16 / t_fwd / 1e9,
        }
    prim = out["random_density_0.25"]
    return {
        "metric": "forward + adjoint (sigma = H psi, per-term <P_k>) time, 24 qubits, 20,000 terms",
        "value": prim["forward_plus_adjoint_ms"],
        "unit": "ms",
        "higher_is_better": False,
        "config": {"workload": "adjoint", "num_qubits": n, "num_terms": T, "op_seed": 4001,
                   "psi_seed": 4002},
        "results": out,
        "roofline": {"bound": "hbm", "kernel": "sigma_tiled", "peak": 8000.0, "unit": "GB/s",
                     "achieved": prim["sigma_GBs"],
                     "frac": prim["sigma_GBs"] / 8000.0 if prim["sigma_GBs"] else None,
                     "traffic": prim["sigma_bytes_per_call"] / prim["sigma_launches_per_call"]
                     if prim["sigma_launches_per_call"] else None,
                     "definition": "psi read plus sigma read and written once per pass (48 B per "
                                   "amplitude; 32 B on the first pass), per launch; the rocprofv3 "
                                   "FETCH_SIZE/WRITE_SIZE counters of the launches agree"},
    }


def compute(args) -> dict:
    """The JSON line of the selected workload."""
    fn = {"commutation": commutation, "matmul": matmul, "batch": batch, "adjoint": adjoint,
          "products": products, "grouping": grouping}[args.workload]
    line = fn(args)
    line.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "dtype": "f64",
                 "data": "synthetic"})
    return line


def run(args) -> int:
    print(json.dumps(compute(args)), flush=True)
    return 0
