"""Circuit generators for the workloads this engine is measured on.

Gates are plain dicts ``{"kind", "q": (q0, q1, q2), "angle", "phi", "lam",
"k", "u"}`` with the reference constructors' argument order (see
:mod:`afquantumsim_amd.ops`).  Angles are rounded to float32 because the
reference's gate classes store ``float`` angles (include/quantum.h:1015).
"""
from __future__ import annotations

import math

import numpy as np

PI_F = float(np.float32(math.pi))  # aqs::pi, include/quantum.h:121


def gate(kind, q0=0, q1=0, q2=0, angle=0.0, phi=0.0, lam=0.0, k=0, u=None):
    return {"kind": kind, "q": (int(q0), int(q1), int(q2)),
            "angle": float(np.float32(angle)), "phi": float(np.float32(phi)),
            "lam": float(np.float32(lam)), "k": int(k),
            "u": None if u is None else np.ascontiguousarray(u, dtype=np.complex128)}


def random_layered(n: int, depth: int, seed: int = 1):
    """20-qubit random circuit workload: per layer, each qubit draws one of
    {H, X, RZ(t), RY(t)} with t ~ U[0, 2pi), then CX on a random perfect
    matching of the qubits (random control/target order)."""
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(depth):
        for q in range(n):
            c = int(rng.integers(4))
            t = float(rng.uniform(0.0, 2.0 * math.pi))
            if c == 0:
                gates.append(gate("H", q))
            elif c == 1:
                gates.append(gate("X", q))
            elif c == 2:
                gates.append(gate("RotZ", q, angle=t))
            else:
                gates.append(gate("RotY", q, angle=t))
        perm = rng.permutation(n)
        for p in range(n // 2):
            a, b = int(perm[2 * p]), int(perm[2 * p + 1])
            gates.append(gate("CX", a, b))
    return gates


def brickwork(n: int, depth: int, seed: int = 2):
    """Brickwork workload: per layer U3(t, p, l) on every qubit with angles
    ~ U[0, 2pi), then CZ on (0,1),(2,3),... in even layers and (1,2),(3,4),...
    in odd layers."""
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(depth):
        for q in range(n):
            t, p, l = rng.uniform(0.0, 2.0 * math.pi, size=3)
            gates.append(gate("U3", q, angle=t, phi=p, lam=l))
        start = layer % 2
        for a in range(start, n - 1, 2):
            gates.append(gate("CZ", a, a + 1))
    return gates


def fourier_transform(n: int, final_swaps: bool = True):
    """aqs::fourier_transform (src/quantum_algo.cpp:103-114): for i = n-1..0,
    H(i) then CPhase(j, i, pi / 2^(i-j)) for j < i.  The reference emits no
    final swaps; with ``final_swaps`` Swap(i, n-1-i) for i < n/2 is appended."""
    gates = []
    for i in range(n - 1, -1, -1):
        gates.append(gate("H", i))
        for j in range(i):
            ang = np.float32(PI_F) / np.float32(1 << (i - j))
            gates.append(gate("CPhase", j, i, angle=float(ang)))
    if final_swaps:
        for i in range(n // 2):
            gates.append(gate("Swap", i, n - 1 - i))
    return gates


def inverse_fourier_transform(n: int):
    """aqs::inverse_fourier_transform (src/quantum_algo.cpp:116-129)."""
    gates = []
    for i in range(n):
        for j in range(i - 1, -1, -1):
            ang = -np.float32(PI_F) / np.float32(1 << (i - j))
            gates.append(gate("CPhase", j, i, angle=float(ang)))
        gates.append(gate("H", i))
    return gates


def ry_rz_ladder(n: int, layers: int, rng: np.random.Generator):
    """Variational ansatz: per layer RY(t) on every qubit, RZ(t) on every
    qubit (t ~ U[-pi, pi]), then CX(i, i+1) for i < n-1."""
    gates = []
    for _ in range(layers):
        for q in range(n):
            gates.append(gate("RotY", q, angle=float(rng.uniform(-math.pi, math.pi))))
        for q in range(n):
            gates.append(gate("RotZ", q, angle=float(rng.uniform(-math.pi, math.pi))))
        for q in range(n - 1):
            gates.append(gate("CX", q, q + 1))
    return gates


def variational_batch(count: int, nmin: int = 10, nmax: int = 16, lmin: int = 4,
                      lmax: int = 12, seed: int = 4):
    """Batched ansatz workload: ``count`` states with n ~ U{nmin..nmax} and
    L ~ U{lmin..lmax} layers each.  Returns a list of (n, gates)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        L = int(rng.integers(lmin, lmax + 1))
        out.append((n, ry_rz_ladder(n, L, rng)))
    return out
