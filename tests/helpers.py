"""Shared helpers for the parity tests (test-only)."""
from __future__ import annotations

import json
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

R2 = 0.70710678118654752440


def load_kat():
    with open(os.path.join(GOLDEN, "kat_gates.json")) as f:
        return json.load(f)


def qstate(v):
    """A QState value from the fixture: '0', '1', '+', '-' or [re0, im0, re1, im1],
    normalised as QState::force_normalize does (src/quantum.cpp:145-157), in FP64."""
    if v == "0":
        a = np.array([1.0, 0.0], dtype=np.complex128)
    elif v == "1":
        a = np.array([0.0, 1.0], dtype=np.complex128)
    elif v == "+":
        a = np.array([R2, R2], dtype=np.complex128)
    elif v == "-":
        a = np.array([R2, -R2], dtype=np.complex128)
    else:
        a = np.array([complex(v[0], v[1]), complex(v[2], v[3])], dtype=np.complex128)
    return a / np.linalg.norm(a)


def init_qamps(init):
    return np.stack([qstate(v) for v in init])


def fixture_gates(gl, n):
    """Fixture gate entries -> oracle gate dicts (circuit gates get their
    sub-unitary from the oracle itself)."""
    out = []
    for e in gl:
        if e["kind"] in ("Circuit", "ControlCircuit"):
            sub = fixture_gates(e["sub"], e["sub_n"])
            U = oracle.unitary(e["sub_n"], sub)
            if e["kind"] == "Circuit":
                out.append(oracle.gate("Circuit", e["begin"], k=e["sub_n"], u=U))
            else:
                out.append(oracle.gate("ControlCircuit", e["control"], e["begin"], k=e["sub_n"], u=U))
        else:
            q = e["q"]
            out.append(oracle.gate(e["kind"], q[0], q[1], q[2], angle=e.get("angle", 0.0)))
    return out


def fequal(a, b, diff):
    """test/tests.cpp:19-22: relative tolerance (an expected 0 needs an exact 0)."""
    return abs(a - b) <= max(abs(a), abs(b)) * diff


def cfequal(a, b, diff):
    return fequal(a.real, b.real, diff) and fequal(a.imag, b.imag, diff)


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(np.complex128)


ALL_KINDS_4Q = [
    ("X", (1,)), ("Y", (2,)), ("Z", (3,)), ("H", (0,)), ("Phase", (2,), 0.7),
    ("Swap", (0, 3)), ("RotX", (1,), 1.1), ("RotY", (3,), -0.4), ("RotZ", (0,), 2.5),
    ("CX", (2, 0)), ("CY", (0, 3)), ("CZ", (1, 2)), ("CH", (3, 1)), ("CPhase", (0, 2), 0.3),
    ("CSwap", (1, 0, 3)), ("CRotX", (2, 3), 0.9), ("CRotY", (3, 0), 1.7), ("CRotZ", (0, 1), -1.2),
    ("CCX", (0, 3, 1)), ("Or", (2, 1, 0)), ("U3", (2,), 0.8),
]


def kind_gate(spec):
    kind, q = spec[0], spec[1]
    ang = spec[2] if len(spec) > 2 else 0.0
    q = tuple(q) + (0,) * (3 - len(q))
    if kind == "U3":
        return oracle.gate("U3", q[0], angle=ang, phi=0.3, lam=-1.9)
    return oracle.gate(kind, q[0], q[1], q[2], angle=ang)
