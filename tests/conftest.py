import pathlib
import subprocess
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "tests" / "golden"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (oracle/liboracle.so); built with make if missing."""
    import oracle

    if not oracle.ORACLE_LIB.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return oracle.oracle()


@pytest.fixture(scope="session")
def ref_lib():
    """The reference's own scalar path compiled in place, or skip where absent."""
    import oracle

    r = oracle.reference()
    if r is None:
        pytest.skip("oracle/_ref was not built (needs the reference tree)")
    return r


@pytest.fixture(scope="session")
def gpu():
    import wolfgang_amd

    status = wolfgang_amd.hip_status()
    if not status["runtime_available"]:
        pytest.skip(status["skip_reason"] or "no HIP device")
    return status


def packed(op):
    return op.num_qubits, op.x_words(), op.z_words(), op.coeffs()


def assert_bit_exact(actual, expected_packed):
    """actual: PauliSum; expected: (n, x, z, c) from the oracle."""
    n, x, z, c = expected_packed
    assert actual.num_qubits == n
    assert actual.num_terms == c.shape[0]
    np.testing.assert_array_equal(actual.x_words(), x)
    np.testing.assert_array_equal(actual.z_words(), z)
    np.testing.assert_array_equal(actual.coeffs().view(np.uint64), c.view(np.uint64))
