"""The plain complex64 CPU baseline (oracle/cpu_sv.c) that bench.py times:
it must compute the same states as the FP64 oracle (to FP32 rounding) and
sample by the same rule, or its timing says nothing."""
import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits

KINDS_1Q = ["X", "Y", "Z", "H", "Phase", "RotX", "RotY", "RotZ"]
KINDS_C = ["CX", "CY", "CZ", "CH", "CPhase", "CRotX", "CRotY", "CRotZ"]


def _random_gates(n, count, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(count):
        r = rng.integers(5)
        if r == 0:
            gates.append(oracle.gate(KINDS_1Q[rng.integers(len(KINDS_1Q))], int(rng.integers(n)),
                                     angle=rng.uniform(-3, 3)))
        elif r == 1:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(oracle.gate(KINDS_C[rng.integers(len(KINDS_C))], int(a), int(b), angle=rng.uniform(-3, 3)))
        elif r == 2:
            gates.append(oracle.gate("U3", int(rng.integers(n)), angle=rng.uniform(0, 6), phi=rng.uniform(0, 6),
                                     lam=rng.uniform(0, 6)))
        elif r == 3:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(oracle.gate("Swap", int(a), int(b)))
        else:
            a, b, c = rng.choice(n, 3, replace=False)
            gates.append(oracle.gate(["CCX", "CSwap"][rng.integers(2)], int(a), int(b), int(c)))
    return gates


@pytest.mark.parametrize("n,threads", [(3, 1), (7, 4), (12, 8)])
def test_cpu_baseline_matches_oracle(n, threads):
    gates = _random_gates(n, 200, seed=n)
    v = oracle.gaussian_state(n, 3)
    want = oracle.run(v, n, gates)
    got = oracle.cpu_sv_run(v.astype(np.complex64), n, gates, threads=threads)
    assert oracle.infidelity(got, want) < 1e-10
    assert np.abs(got - want).max() < 2e-6


@pytest.mark.parametrize("n", [10, 20])
def test_cpu_baseline_workloads(n):
    """The measured workloads' gate sets (brickwork U3/CZ; random layers H/X/RZ/RY/CX)."""
    for gates in (circuits.brickwork(n, 3, seed=1), circuits.random_layered(n, 3, seed=1)):
        want = oracle.run(oracle.basis_state(n, 0), n, gates, threads=4)
        got = oracle.cpu_sv_run(np.zeros(1 << n, np.complex64) + np.eye(1, 1 << n, dtype=np.complex64)[0], n,
                                gates, threads=4)
        assert oracle.infidelity(got, want) < 1e-10


def test_cpu_baseline_sampling_rule():
    n = 9
    v = oracle.gaussian_state(n, 8).astype(np.complex64)
    rng = np.random.default_rng(2)
    u = np.concatenate([rng.uniform(size=500), [0.0, np.nextafter(1.0, 0.0)]])
    got = oracle.cpu_sv_sample(v, n, u, threads=4)
    # brute force of the rule: first index whose FP64 inclusive CDF exceeds u
    cdf = np.cumsum(np.abs(v.astype(np.complex128)) ** 2)
    want = np.minimum(np.searchsorted(cdf, u, side="right"), (1 << n) - 1)
    np.testing.assert_array_equal(got, want)
