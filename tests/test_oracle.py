"""Pins the FP64 oracle: the reference's own known-answer tests
(test/tests.cpp), the explicit-operator method the reference uses, and
closed forms / invariants that a dropped term, a wrong sign or a transposed
qubit order would break."""
import math

import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits
from tests.helpers import (ALL_KINDS_4Q, cfequal, fixture_gates, init_qamps, kind_gate,
                           load_kat, qstate, random_state)

KAT = load_kat()


@pytest.mark.parametrize("case", KAT["gate_cases"], ids=lambda c: c["name"])
def test_kat_gates(case):
    """test/tests.cpp:414-907 gate known answers."""
    n = case["n"]
    sv = oracle.product_state(n, init_qamps(case["init"]))
    out = oracle.run(sv, n, fixture_gates(case["gates"], n))
    for idx, re, im in case["expect"]:
        got = out[idx]
        if case["tol"] == 0:
            assert abs(got - complex(re, im)) < 1e-12, (idx, got)
        else:
            assert cfequal(got, complex(re, im), case["tol"]), (idx, got, re, im)


@pytest.mark.parametrize("eq", KAT["equalities"], ids=lambda c: c["name"])
def test_kat_circuit_equalities(eq):
    """test/tests.cpp:909-961: ControlGate/Gate equal their flattened circuits."""
    n = eq["n"]
    U1 = oracle.unitary(n, fixture_gates(eq["lhs"], n))
    U2 = oracle.unitary(n, fixture_gates(eq["rhs"], n))
    np.testing.assert_allclose(U1, U2, atol=1e-12)


@pytest.mark.parametrize("case", KAT["ch_cases"], ids=lambda c: c["name"])
def test_kat_ch(case):
    """test/tests.cpp:963-1002 (asserted with 1e-12 in FP64; the reference's
    exact float equality cannot hold: 0.70710677^2 != 0.7071068^2)."""
    n = case["n"]
    sv = oracle.product_state(n, init_qamps(case["init"]))
    out = oracle.run(sv, n, fixture_gates(case["gates"], n))
    ref = oracle.product_state(n, init_qamps(case["expect_product"]))
    np.testing.assert_allclose(out, ref, atol=1e-12)


@pytest.mark.parametrize("case", KAT["prob_cases"], ids=lambda c: c["name"])
def test_kat_probabilities(case):
    """test/tests.cpp:1141-1194."""
    n = case["n"]
    sv = oracle.product_state(n, init_qamps(case["init"]))
    for q, p in enumerate(case["p1"]):
        got = oracle.prob1(sv, n, q)
        if case["tol"] == 0:
            assert got == p
        else:
            assert abs(got - p) <= max(abs(got), p) * case["tol"]


@pytest.mark.parametrize("case", KAT["measure_cases"], ids=lambda c: c["name"])
def test_kat_peek_measure_all(case):
    """test/tests.cpp:281-301: a basis state samples to its own index for any u."""
    n = case["n"]
    sv = oracle.product_state(n, init_qamps(case["init"]))
    u = np.array([0.0, 0.25, 0.5, 0.999999], dtype=np.float64)
    assert (oracle.sample_c128(sv, n, u) == case["peek_all"]).all()
    assert (oracle.sample_fp64_cdf(sv, n, u) == case["peek_all"]).all()


@pytest.mark.parametrize("spec", ALL_KINDS_4Q, ids=lambda s: s[0])
@pytest.mark.parametrize("n", [4, 5])
def test_index_level_equals_reference_method(spec, n):
    """The index-level gate equals the reference's explicit operator (Kronecker
    product I (x) U (x) I or the expanded CSR rows) times the state."""
    g = kind_gate(spec)
    sv = random_state(n, 11 + n)
    op = oracle.dense_operator(n, g)
    np.testing.assert_allclose(oracle.run(sv, n, [g]), op @ sv, atol=1e-13)
    np.testing.assert_allclose(op.conj().T @ op, np.eye(1 << n), atol=1e-13)


def test_kron_operator_against_numpy():
    """The dense one-qubit operator is numpy's kron(I, U, I) with qubit 0 most significant."""
    n = 4
    for q in range(n):
        g = oracle.gate("U3", q, angle=0.4, phi=1.3, lam=-0.2)
        U = oracle.matrix_1q(g)
        ref = np.kron(np.kron(np.eye(1 << q), U), np.eye(1 << (n - q - 1)))
        np.testing.assert_allclose(oracle.dense_operator(n, g), ref, atol=1e-15)


def test_u3_identity():
    """U3(t, p, l) = Phase(p) RotY(t) Phase(l) exactly (and the OpenQASM form)."""
    n = 3
    sv = random_state(n, 5)
    t, p, l = 0.9, -2.2, 1.4
    a = oracle.run(sv, n, [oracle.gate("U3", 1, angle=t, phi=p, lam=l)])
    b = oracle.run(sv, n, [oracle.gate("Phase", 1, angle=l), oracle.gate("RotY", 1, angle=t),
                           oracle.gate("Phase", 1, angle=p)])
    np.testing.assert_allclose(a, b, atol=1e-14)
    tf, pf, lf = (float(np.float32(x)) for x in (t, p, l))
    U = np.array([[math.cos(tf / 2), -np.exp(1j * lf) * math.sin(tf / 2)],
                  [np.exp(1j * pf) * math.sin(tf / 2), np.exp(1j * (pf + lf)) * math.cos(tf / 2)]])
    np.testing.assert_allclose(oracle.matrix_1q(oracle.gate("U3", 0, angle=t, phi=p, lam=l)), U, atol=1e-15)


def test_pauli_and_clifford_identities():
    n = 3
    sv = random_state(n, 7)
    H, X, Z = (lambda q: oracle.gate("H", q)), (lambda q: oracle.gate("X", q)), (lambda q: oracle.gate("Z", q))
    np.testing.assert_allclose(oracle.run(sv, n, [H(1), X(1), H(1)]), oracle.run(sv, n, [Z(1)]), atol=1e-14)
    np.testing.assert_allclose(oracle.run(sv, n, [X(2), X(2)]), sv, atol=0)
    # CX(c, t) = H(t) CZ(c, t) H(t)
    np.testing.assert_allclose(oracle.run(sv, n, [oracle.gate("CX", 0, 2)]),
                               oracle.run(sv, n, [H(2), oracle.gate("CZ", 0, 2), H(2)]), atol=1e-14)
    # Y = i X Z
    np.testing.assert_allclose(oracle.run(sv, n, [oracle.gate("Y", 0)]),
                               1j * oracle.run(sv, n, [Z(0), X(0)]), atol=1e-15)
    # Swap = three CX
    np.testing.assert_allclose(oracle.run(sv, n, [oracle.gate("Swap", 0, 2)]),
                               oracle.run(sv, n, [oracle.gate("CX", 0, 2), oracle.gate("CX", 2, 0),
                                                  oracle.gate("CX", 0, 2)]), atol=0)
    # Or(a, b, t) = X a; X b; X t; CCX(a, b, t); X a; X b  (Or::to_string, src/quantum.cpp:1662-1677)
    np.testing.assert_allclose(oracle.run(sv, n, [oracle.gate("Or", 0, 1, 2)]),
                               oracle.run(sv, n, [X(0), X(1), X(2), oracle.gate("CCX", 0, 1, 2), X(0), X(1)]),
                               atol=0)


def test_permutation_basis_states_exact():
    """Permutation gates map basis states to basis states exactly."""
    n = 5
    rng = np.random.default_rng(3)
    gates = []
    for _ in range(40):
        kind = rng.choice(["X", "CX", "CCX", "Swap", "CSwap", "Or"])
        q = rng.choice(n, size=3, replace=False)
        gates.append(oracle.gate(str(kind), int(q[0]), int(q[1]), int(q[2])))
    for idx in range(1 << n):
        out = oracle.run(oracle.basis_state(n, idx), n, gates)
        assert np.count_nonzero(out) == 1 and out.max() == 1.0


@pytest.mark.parametrize("n", [3, 6, 9])
def test_qft_closed_form(n):
    """fourier_transform(n) + final swaps = bitrev(ifft(bitrev(v))) * sqrt(N)."""
    N = 1 << n
    v = random_state(n, 100 + n)
    gates = circuits.fourier_transform(n, final_swaps=True)
    out = oracle.run(v, n, gates)
    rev = np.array([int(format(i, f"0{n}b")[::-1], 2) for i in range(N)])
    ref = np.fft.ifft(v[rev])[rev] * math.sqrt(N)
    # float32 angles pi/2^k limit agreement to ~1e-7
    assert oracle.infidelity(out, ref) < 1e-12
    np.testing.assert_allclose(out, ref, atol=3e-7)
    noswap = oracle.run(v, n, circuits.fourier_transform(n, final_swaps=False))
    np.testing.assert_allclose(noswap, np.fft.ifft(v[rev]) * math.sqrt(N), atol=3e-7)


def test_generate_statevector_is_kron():
    n = 4
    qa = np.stack([qstate([0.6, 0, 0, 0.8]), qstate("-"), qstate("1"), qstate([1, 2, 3, -1])])
    ref = qa[0]
    for q in range(1, n):
        ref = np.kron(ref, qa[q])
    np.testing.assert_allclose(oracle.product_state(n, qa), ref, atol=1e-15)


def test_prob_and_expectation():
    n = 6
    sv = random_state(n, 9)
    p = np.abs(sv) ** 2
    idx = np.arange(1 << n)
    ez = 0.0
    for q in range(n):
        m = 1 << (n - 1 - q)
        p1 = p[(idx & m) != 0].sum()
        assert abs(oracle.prob1(sv, n, q) - p1) < 1e-14
        ez += (1 - p1) - p1
    assert abs(oracle.expect_sum_z(sv, n) - ez) < 1e-13
    assert abs(oracle.norm2(sv, n) - 1.0) < 1e-14


def _brute_sample(amps, u):
    """first i with sum_{j<=i} |a_j|^2 > u, by a plain loop in Python floats."""
    out = []
    p = [float(a.real) ** 2 + float(a.imag) ** 2 for a in amps]
    for x in u:
        acc = 0.0
        k = None
        for i, pi in enumerate(p):
            acc += pi
            if acc > x:
                k = i
                break
        if k is None:
            k = max(i for i, pi in enumerate(p) if pi > 0)
        out.append(k)
    return np.array(out, dtype=np.uint32)


def test_sampling_rule_matches_brute_force():
    n = 7
    rng = np.random.default_rng(21)
    sv = random_state(n, 21).astype(np.complex64)
    u = rng.uniform(size=5000)
    fixed = oracle.sample_c64(sv, n, u)
    brute = _brute_sample(sv, u)
    # The fixed-point cumulative sum differs from the float loop only within
    # ~1e-15 of a boundary; with these uniforms none falls there.
    assert (fixed == brute).all()
    assert (oracle.sample_fp64_cdf(sv.astype(np.complex128), n, u) == brute).all()


def test_sampling_edges():
    n = 4
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[3] = 0.6
    sv[9] = 0.8j
    u = np.array([0.0, 0.3599999, 0.3600001, 0.9999999, 1.0, 1.5])  # float32(0.6)^2 = 0.36000003
    got = oracle.sample_c64(sv, n, u)
    assert list(got[:4]) == [3, 3, 9, 9]
    # u at or above the total: the last index with non-zero probability
    assert list(got[4:]) == [9, 9]


def test_sampling_distribution():
    """Histogram of many shots lies within the reference's 99.9% band
    (test/tests.cpp:343-406, z = 3.291)."""
    n = 5
    qa = np.stack([qstate([1, 0, 1, 0]), qstate([-2, 0, 3, 0]), qstate("0"), qstate("1"), qstate([0, 0.6, 0.8, 0])])
    sv = oracle.product_state(n, qa).astype(np.complex64)
    reps = 10000
    u = np.random.default_rng(5).uniform(size=reps)
    counts = np.bincount(oracle.sample_c64(sv, n, u), minlength=1 << n)
    p = np.abs(sv.astype(np.complex128)) ** 2
    for i in range(1 << n):
        sd = math.sqrt(p[i] * (1 - p[i]))
        d = math.sqrt(reps) * 3.291 * sd
        assert reps * p[i] - d - 1 <= counts[i] <= reps * p[i] + d + 1


def test_gaussian_state_normalised():
    n = 8
    sv = oracle.gaussian_state(n, 1234)
    assert abs(np.linalg.norm(sv) - 1.0) < 1e-14
    raw = np.array([oracle.lib().orc_gaussian_component(1234, i, 0) for i in range(1 << n)])
    assert abs(raw.mean()) < 0.2 and 0.8 < raw.std() < 1.2


def test_refmethod_matches_oracle():
    """The reference-method timing restatement (dense N x N products) computes
    the same state as the index-level oracle for every gate kind."""
    n = 6
    rng = np.random.default_rng(11)
    gates = []
    for spec in ALL_KINDS_4Q:
        kind, q, *ang = spec
        qs = [int(x) for x in rng.choice(n, size=len(q), replace=False)]
        gates.append(kind_gate((kind, tuple(qs), ang[0]) if ang else (kind, tuple(qs))))
    U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    gates.append(oracle.gate("Circuit", 2, k=2, u=U))
    gates.append(oracle.gate("ControlCircuit", 0, 3, k=2, u=U))
    sv = random_state(n, 5)
    np.testing.assert_allclose(oracle.refmethod_run(sv, n, gates, 2), oracle.run(sv, n, gates), atol=1e-12)
