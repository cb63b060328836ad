"""Extract the reference's published known-answer digests into reference_kats.json.

Reads the reference's benchmark summaries (docs/benchmarks/data/*/summary.json)
and its case tables' parameters; run only where the reference tree is mounted.
The output is data (shapes, seeds, counts, sha256 digests), committed beside
this script so that tests never read the reference at run time.

Sources:
  commutation counts  docs/benchmarks/data/rocm_mi300x_campaign2_2026-04-30/summary.json
                      (computed from the CPU result, benchmarks/bench_rocm_kernels.py:1366,1403)
  simplify digests    docs/benchmarks/data/rocm_mi300x_campaign3_2026-04-30/summary.json
                      (CPU `expected`, benchmarks/bench_rocm_kernels.py:1104-1121,1733)
  campaign6 digests   docs/benchmarks/data/rocm_mi300x_campaign6_2026-04-30/summary.json
                      (operator label hashes; results are device values, tolerance only)
Case parameters: benchmarks/bench_rocm_kernels.py:44-98,281-345,459-588.
"""

import json
import pathlib
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
DATA = REF / "docs" / "benchmarks" / "data"
OUT = pathlib.Path(__file__).resolve().parent / "reference_kats.json"

COMMUTATION_CASES = {
    # name: (num_qubits, lhs_terms, rhs_terms, term_weight, seed)
    "campaign2_small_dense_output": (8, 128, 128, 2, 9522),
    "campaign2_mid_dense_output": (64, 2048, 2048, 6, 9523),
    "campaign2_large_dense_output": (128, 4096, 4096, 8, 9524),
    "campaign2_compact_mid": (64, 2048, 2048, 6, 9623),
    "campaign2_compact_large": (128, 4096, 4096, 8, 9624),
    "campaign2_profiler_dense_pairs": (128, 4096, 4096, 8, 9725),
}
SIMPLIFY_CASES = {
    # name: (num_qubits, num_terms, term_weight, duplicate_rate, seed, mode)
    "campaign3_smoke_one_word": (8, 128, 2, 0.25, 39421, "random"),
    "campaign3_duplicate_heavy": (24, 32768, 4, 0.875, 39422, "random"),
    "campaign3_duplicate_light": (24, 32768, 4, 0.0625, 39423, "random"),
    "campaign3_all_zero": (24, 4096, 4, 1.0, 39424, "cancelling_duplicates"),
    "campaign3_profiler_duplicate_heavy": (24, 32768, 4, 0.875, 39427, "random"),
    "campaign3_strategy_duplicate_heavy": (24, 32768, 4, 0.875, 39428, "random"),
    "campaign3_wide_two_word": (70, 8192, 6, 0.25, 39425, "random"),
    "campaign3_generic_multiword": (130, 4096, 8, 0.25, 39426, "random"),
}
CAMPAIGN6_EXPECTATION = {
    # name: (num_qubits, num_terms, term_weight, dtype, seed)
    "campaign6_expectation_two_qubit_complex128": (2, 5, 2, "complex128", 69421),
    "campaign6_expectation_ten_qubit_complex64": (10, 256, 4, "complex64", 69422),
    "campaign6_profiler_expectation": (14, 1024, 6, "complex128", 69441),
}
CAMPAIGN6_MATMUL = {
    # name: (num_qubits, lhs_terms, rhs_terms, term_weight, duplicate_rate, simplify, seed)
    "campaign6_matmul_one_word_simplify_true": (24, 256, 256, 4, 0.25, True, 69431),
    "campaign6_matmul_two_word_simplify_false": (70, 64, 64, 6, 0.0, False, 69432),
    "campaign6_profiler_matmul": (70, 128, 128, 6, 0.25, True, 69442),
}


def rows(campaign: str):
    summary = json.loads((DATA / campaign / "summary.json").read_text())
    out = {}
    for row in summary["rows"]:
        out.setdefault(row["case"], row["correctness_digest"])
    return out


def main() -> None:
    c2 = rows("rocm_mi300x_campaign2_2026-04-30")
    c3 = rows("rocm_mi300x_campaign3_2026-04-30")
    c6 = rows("rocm_mi300x_campaign6_2026-04-30")
    kats = {"commutation": [], "simplify": [], "expectation": [], "matmul": []}
    for name, (n, l, r, w, seed) in COMMUTATION_CASES.items():
        d = c2[name]
        kats["commutation"].append(
            dict(name=name, num_qubits=n, lhs_terms=l, rhs_terms=r, term_weight=w, seed=seed,
                 commuting_count=d["commuting_count"], conflict_count=d["conflict_count"])
        )
    for name, (n, t, w, dup, seed, mode) in SIMPLIFY_CASES.items():
        d = c3[name]
        kats["simplify"].append(
            dict(name=name, num_qubits=n, num_terms=t, term_weight=w, duplicate_rate=dup, seed=seed,
                 mode=mode, output_terms=d["output_terms"], coefficient_l1=d["coefficient_l1"],
                 canonical_label_hash=d["canonical_label_hash"])
        )
    for name, (n, t, w, dtype, seed) in CAMPAIGN6_EXPECTATION.items():
        d = c6[name]
        kats["expectation"].append(
            dict(name=name, num_qubits=n, num_terms=t, term_weight=w, dtype=dtype, seed=seed,
                 operator_label_hash=d["label_hash"], result_real=d["result_real"],
                 result_imag=d["result_imag"])
        )
    for name, (n, l, r, w, dup, simp, seed) in CAMPAIGN6_MATMUL.items():
        d = c6[name]
        kats["matmul"].append(
            dict(name=name, num_qubits=n, lhs_terms=l, rhs_terms=r, term_weight=w, duplicate_rate=dup,
                 simplify=simp, seed=seed, output_terms=d["output_terms"],
                 coefficient_l1=d["coefficient_l1"])
        )
    OUT.write_text(json.dumps(kats, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
