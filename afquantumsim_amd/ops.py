"""Lowering of the reference's gate classes to aqs_op passes.

Each reference gate (include/quantum.h:857-1589) becomes one op of the C ABI
(include/aqs_hip.h).  Two matrix flavours:

* ``exact=False`` (default, the complex64 product path): entries computed
  the way the reference computes them -- float32 cos/sin of the float angle
  (e.g. RotX, src/quantum.cpp:683-688) and the float constant 0.70710678118f
  for H (src/quantum.cpp:46-49);
* ``exact=True`` (the complex128 reference path): FP64 trigonometry of the
  same float angles and 1/sqrt(2) in double, i.e. the exact operator.

U3 has no reference counterpart; its matrix is built in FP64 in both cases.
"""
from __future__ import annotations

import math

import numpy as np

from . import hip

F = np.float32
H_F = float(F(0.70710678118))
H_D = 0.70710678118654752440


def _cs_half(angle, exact):
    if exact:
        a = float(F(angle)) / 2.0
        return math.cos(a), math.sin(a)
    a = F(angle) / F(2.0)
    return float(np.cos(a)), float(np.sin(a))


def _eiphi(angle, exact):
    if exact:
        a = float(F(angle))
        return complex(math.cos(a), math.sin(a))
    a = F(angle)
    return complex(float(np.cos(a)), float(np.sin(a)))


def _mat1(t, m, controls=()):
    return {"kind": hip.OP_MAT1, "targets": (t,), "controls": tuple(controls), "m": np.asarray(m, dtype=np.complex128)}


def _diag1(t, d0, d1, controls=()):
    return {"kind": hip.OP_DIAG1, "targets": (t,), "controls": tuple(controls),
            "m": np.array([d0, 0, 0, d1], dtype=np.complex128)}


def _phase(targets, ph, controls=()):
    return {"kind": hip.OP_PHASE, "targets": tuple(targets), "controls": tuple(controls),
            "m": np.array([ph, 0, 0, 0], dtype=np.complex128)}


def _x(t, controls=(), flags=0):
    return {"kind": hip.OP_X, "targets": (t,), "controls": tuple(controls), "flags": flags}


def matrix_rotx(angle, exact=False):
    c, s = _cs_half(angle, exact)
    return [c, -1j * s, -1j * s, c]


def matrix_roty(angle, exact=False):
    c, s = _cs_half(angle, exact)
    return [c, -s, s, c]


def matrix_rotz_diag(angle, exact=False):
    c, s = _cs_half(angle, exact)
    return complex(c, -s), complex(c, s)


def matrix_u3(theta, phi, lam):
    t, p, l = float(F(theta)), float(F(phi)), float(F(lam))
    c, s = math.cos(t / 2), math.sin(t / 2)
    return [c, -complex(math.cos(l), math.sin(l)) * s, complex(math.cos(p), math.sin(p)) * s,
            complex(math.cos(p + l), math.sin(p + l)) * c]


def matrix_h(exact=False):
    h = H_D if exact else H_F
    return [h, h, h, -h]


Y_MAT = [0, -1j, 1j, 0]


def lower(g, exact: bool = False) -> list:
    """One gate dict (see afquantumsim_amd.circuits.gate) -> list of ops."""
    k = g["kind"]
    q0, q1, q2 = g["q"]
    a = g.get("angle", 0.0)
    if k == "Barrier":
        return []
    if k == "X":
        return [_x(q0)]
    if k == "Y":
        return [_mat1(q0, Y_MAT)]
    if k == "Z":
        return [_phase((q0,), -1.0)]
    if k == "H":
        return [_mat1(q0, matrix_h(exact))]
    if k == "Phase":
        return [_phase((q0,), _eiphi(a, exact))]
    if k == "Swap":
        return [{"kind": hip.OP_SWAP, "targets": (q0, q1)}]
    if k == "RotX":
        return [_mat1(q0, matrix_rotx(a, exact))]
    if k == "RotY":
        return [_mat1(q0, matrix_roty(a, exact))]
    if k == "RotZ":
        return [_diag1(q0, *matrix_rotz_diag(a, exact))]
    if k == "U3":
        return [_mat1(q0, matrix_u3(a, g.get("phi", 0.0), g.get("lam", 0.0)))]
    if k == "CX":
        return [_x(q1, (q0,))]
    if k == "CY":
        return [_mat1(q1, Y_MAT, (q0,))]
    if k == "CZ":
        return [_phase((q1,), -1.0, (q0,))]
    if k == "CH":
        return [_mat1(q1, matrix_h(exact), (q0,))]
    if k == "CPhase":
        return [_phase((q1,), _eiphi(a, exact), (q0,))]
    if k == "CSwap":
        return [{"kind": hip.OP_SWAP, "targets": (q1, q2), "controls": (q0,)}]
    if k == "CRotX":
        return [_mat1(q1, matrix_rotx(a, exact), (q0,))]
    if k == "CRotY":
        return [_mat1(q1, matrix_roty(a, exact), (q0,))]
    if k == "CRotZ":
        return [_diag1(q1, *matrix_rotz_diag(a, exact), controls=(q0,))]
    if k == "CCX":
        ctl = (q0,) if q0 == q1 else (q0, q1)
        return [_x(q2, ctl)]
    if k == "Or":
        ctl = (q0,) if q0 == q1 else (q0, q1)
        return [_x(q2, ctl, hip.CTRL_ANY)]
    if k == "Circuit":
        kk = g["k"]
        return [{"kind": hip.OP_MATK, "targets": tuple(range(q0, q0 + kk)), "u": g["u"]}]
    if k == "ControlCircuit":
        kk = g["k"]
        return [{"kind": hip.OP_MATK, "targets": tuple(range(q1, q1 + kk)), "controls": (q0,), "u": g["u"]}]
    raise ValueError(f"unknown gate kind {k}")


def lower_all(gates, exact: bool = False) -> list:
    out = []
    for g in gates:
        out.extend(lower(g, exact))
    return out


def touched_fraction(op, n: int) -> float:
    """Fraction of the 2^n amplitudes an op reads and writes (algorithmic
    bytes = 16 B x 2^n x this fraction)."""
    kind = op["kind"]
    nc = len(set(op.get("controls", ())))
    if kind == hip.OP_PHASE:
        bits = len(set(op["targets"]) | set(op.get("controls", ())))
        return 0.5 ** bits
    if kind == hip.OP_SWAP:
        return 0.5 * 0.5 ** nc
    if op.get("flags", 0) & hip.CTRL_ANY:
        return 1.0 - 0.5 ** nc
    return 0.5 ** nc
