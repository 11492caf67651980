"""Writes kat_gates.json: the known-answer tests of the reference's
test/tests.cpp transcribed as data (inputs + expected amplitudes).

Each case records the reference line range it comes from, the per-qubit
initial states handed to QSimulator::generate_statevector (QState values
before normalisation), the gate list, and the expected amplitudes with the
reference's tolerance (relative 1e-4 through fequal, test/tests.cpp:19-27;
an expected exact value is marked tol 0).

Circuit-equality cases (ControlGate/Gate/CH, test/tests.cpp:909-1002) are
recorded as a pair of gate lists whose unitaries must agree.

Run:  python tests/golden/make_kat.py   (no inputs; deterministic)
"""
import json
import math
import os

import numpy as np

f32 = lambda x: float(np.float32(x))
PI = f32(math.pi)
ATAN075 = f32(math.atan(np.float32(0.75)))
ATAN1 = f32(math.atan(np.float32(1.0)))
ACOS06x2 = f32(np.float32(math.acos(np.float32(0.6))) * np.float32(2.0))
S8 = 0.3535533906

Z0 = "0"
O1 = "1"
MINUS = "-"
A = [0.6, 0.0, 0.0, 0.8]   # QState({0.6,0},{0,0.8})
B = [0.0, 0.8, 0.6, 0.0]   # QState({0,0.8},{0.6,0})


def g(kind, *q, angle=0.0):
    q = list(q) + [0] * (3 - len(q))
    return {"kind": kind, "q": q, "angle": angle}


CASES = []


def case(name, lines, init, gates, expect, tol=1e-4):
    CASES.append({"name": name, "lines": lines, "n": 4, "init": init,
                  "gates": gates, "expect": [[i, re, im] for i, re, im in expect],
                  "tol": tol})


# test/tests.cpp:424-439  X gate
case("X", "424-439", [Z0, Z0, Z0, O1], [g("X", 2), g("X", 3)], [(0b0010, 1.0, 0.0)], 0)
# :441-459  Y gate
case("Y", "441-459", [Z0, Z0, Z0, O1], [g("Y", 2), g("Y", 3)], [(0b0010, 1.0, 0.0)], 0)
# :461-479  Z gate
case("Z", "461-479", [Z0, Z0, Z0, O1], [g("Z", 2), g("Z", 3)], [(0b0001, -1.0, 0.0)], 0)
# :481-491  Not gate (alias of X)
case("Not", "481-491", [Z0, Z0, Z0, O1], [g("X", 2), g("X", 3)], [(0b0010, 1.0, 0.0)], 0)
# :493-519  Hadamard
case("H", "493-519", [Z0, Z0, Z0, O1], [g("H", 2), g("H", 3)],
     [(0b0000, 0.5, 0.0), (0b0001, -0.5, 0.0), (0b0010, 0.5, 0.0), (0b0011, -0.5, 0.0)])
# :521-547  Phase
case("Phase", "521-547", [Z0, Z0, A, Z0],
     [g("Phase", 2, angle=ATAN075), g("Phase", 3, angle=ATAN1)],
     [(0b0000, 0.6, 0.0), (0b0001, 0.0, 0.0), (0b0010, -0.48, 0.64), (0b0011, 0.0, 0.0)])
# :549-565  RotX
case("RotX", "549-565", [Z0, Z0, MINUS, Z0], [g("RotX", 2, angle=f32(PI / 2))],
     [(0b0000, 0.5, 0.5), (0b0001, 0.0, 0.0), (0b0010, -0.5, -0.5), (0b0011, 0.0, 0.0)])
# :567-583  RotY
case("RotY", "567-583", [Z0, Z0, O1, Z0], [g("RotY", 2, angle=ACOS06x2)],
     [(0b0000, -0.8, 0.0), (0b0001, 0.0, 0.0), (0b0010, 0.6, 0.0), (0b0011, 0.0, 0.0)])
# :585-601  RotZ
case("RotZ", "585-601", [Z0, Z0, MINUS, Z0], [g("RotZ", 2, angle=f32(-PI / 2))],
     [(0b0000, 0.5, 0.5), (0b0001, 0.0, 0.0), (0b0010, -0.5, 0.5), (0b0011, 0.0, 0.0)])
XOR_EXP = [(0b0100, 0.0, 0.48), (0b0101, 0.36, 0.0), (0b0110, -0.64, 0.0), (0b0111, 0.0, 0.48)]
# :603-619  Xor (alias of CX)
case("Xor", "603-619", [Z0, O1, A, A], [g("CX", 0, 2), g("CX", 1, 3)], XOR_EXP)
# :621-637  Swap
case("Swap", "621-637", [Z0, O1, A, B], [g("Swap", 0, 2), g("Swap", 1, 3)],
     [(0b0001, 0.0, 0.48), (0b0101, 0.36, 0.0), (0b1001, -0.64, 0.0), (0b1101, 0.0, 0.48)])
# :639-655  CNot
case("CNot", "639-655", [Z0, O1, A, A], [g("CX", 0, 2), g("CX", 1, 3)], XOR_EXP)
# :657-685  CX (and vector insertion)
case("CX", "657-685", [[1.0, 0.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], A, A],
     [g("CX", 0, 2), g("CX", 1, 3)], XOR_EXP)
# :687-703  CY
case("CY", "687-703", [Z0, O1, A, A], [g("CY", 0, 2), g("CY", 1, 3)],
     [(0b0100, 0.48, 0.0), (0b0101, 0.0, 0.36), (0b0110, 0.0, 0.64), (0b0111, -0.48, 0.0)])
# :705-720  CZ
case("CZ", "705-720", [Z0, O1, A, A], [g("CZ", 0, 2), g("CZ", 1, 3)],
     [(0b0100, 0.36, 0.0), (0b0101, 0.0, -0.48), (0b0110, 0.0, 0.48), (0b0111, 0.64, 0.0)])
# :722-738  CPhase
case("CPhase", "722-738", [Z0, O1, A, A],
     [g("CPhase", 0, 2, angle=ATAN075), g("CPhase", 1, 3, angle=ATAN075)],
     [(0b0100, 0.36, 0.0), (0b0101, -0.288, 0.384), (0b0110, 0.0, 0.48), (0b0111, -0.512, -0.384)])
# :740-760  CRotX
case("CRotX", "740-760", [Z0, O1, MINUS, MINUS],
     [g("CRotX", 0, 2, angle=f32(PI / 2)), g("CRotX", 1, 3, angle=f32(PI / 2))],
     [(0b0100, S8, S8), (0b0101, -S8, -S8), (0b0110, -S8, -S8), (0b0111, S8, S8)])
# :762-779  CRotY
case("CRotY", "762-779", [Z0, O1, O1, O1],
     [g("CRotY", 0, 2, angle=ACOS06x2), g("CRotY", 1, 3, angle=ACOS06x2)],
     [(0b0100, 0.0, 0.0), (0b0101, 0.0, 0.0), (0b0110, -0.8, 0.0), (0b0111, 0.6, 0.0)])
# :781-798  CRotZ
case("CRotZ", "781-798", [Z0, O1, MINUS, MINUS],
     [g("CRotZ", 0, 2, angle=f32(-PI / 2)), g("CRotZ", 1, 3, angle=f32(-PI / 2))],
     [(0b0100, S8, S8), (0b0101, -S8, S8), (0b0110, -S8, -S8), (0b0111, S8, -S8)])
# :800-825  Or (three initial states with the same circuit)
OR_G = [g("Or", 0, 1, 2), g("Or", 1, 0, 3)]
case("Or_a", "800-813", [Z0, O1, Z0, Z0], OR_G, [(0b0111, 1.0, 0.0)])
case("Or_b", "815-819", [O1, O1, Z0, Z0], OR_G, [(0b1111, 1.0, 0.0)])
case("Or_c", "821-825", [Z0, Z0, Z0, Z0], OR_G, [(0b0000, 1.0, 0.0)])
# :827-852  And (alias of CCNot)
AND_G = [g("CCX", 0, 1, 2), g("CCX", 1, 0, 3)]
case("And_a", "827-840", [Z0, O1, Z0, Z0], AND_G, [(0b0100, 1.0, 0.0)])
case("And_b", "842-846", [O1, O1, Z0, Z0], AND_G, [(0b1111, 1.0, 0.0)])
case("And_c", "848-852", [Z0, Z0, Z0, Z0], AND_G, [(0b0000, 1.0, 0.0)])
# :854-880  CSwap
case("CSwap_a", "854-869", [Z0, O1, A, B], [g("CSwap", 0, 2, 3)],
     [(0b0100, 0.0, 0.48), (0b0101, 0.36, 0.0), (0b0110, -0.64, 0.0), (0b0111, 0.0, 0.48)])
case("CSwap_b", "871-880", [Z0, O1, A, B], [g("CSwap", 1, 2, 3)],
     [(0b0100, 0.0, 0.48), (0b0101, -0.64, 0.0), (0b0110, 0.36, 0.0), (0b0111, 0.0, 0.48)])
# :882-907  CCNot
case("CCNot_a", "882-895", [Z0, O1, Z0, Z0], AND_G, [(0b0100, 1.0, 0.0)])
case("CCNot_b", "897-901", [O1, O1, Z0, Z0], AND_G, [(0b1111, 1.0, 0.0)])
case("CCNot_c", "903-907", [Z0, Z0, Z0, Z0], AND_G, [(0b0000, 1.0, 0.0)])

# Circuit equalities (unitaries of two gate lists must agree).
SUB_CTRL = [g("H", 0), g("CX", 0, 1), g("Z", 0), g("Swap", 0, 1)]      # :911-915
SUB_GATE = [g("H", 0), g("CPhase", 0, 1, angle=1.0), g("X", 1), g("Swap", 0, 1)]  # :946-949
EQUALITIES = [
    {"name": "ControlGate_0_2", "lines": "909-928", "n": 4,
     "lhs": [{"kind": "ControlCircuit", "control": 0, "begin": 2, "sub_n": 2, "sub": SUB_CTRL}],
     "rhs": [g("CH", 0, 2), g("CCX", 0, 2, 3), g("CZ", 0, 2), g("CSwap", 0, 2, 3)]},
    {"name": "ControlGate_2_0", "lines": "930-941", "n": 4,
     "lhs": [{"kind": "ControlCircuit", "control": 2, "begin": 0, "sub_n": 2, "sub": SUB_CTRL}],
     "rhs": [g("CH", 2, 0), g("CCX", 0, 2, 1), g("CZ", 2, 0), g("CSwap", 2, 0, 1)]},
    {"name": "Gate_1", "lines": "943-961", "n": 4,
     "lhs": [{"kind": "Circuit", "begin": 1, "sub_n": 2, "sub": SUB_GATE}],
     "rhs": [g("H", 1), g("CPhase", 1, 2, angle=1.0), g("X", 2), g("Swap", 1, 2)]},
]

# CH (test/tests.cpp:963-1002): simulated state vs a product state.
R2 = f32(math.sqrt(np.float32(0.5)))
CH_CASES = [
    {"name": "CH_inactive", "lines": "965-985", "n": 4,
     "init": [Z0, Z0, O1, Z0], "gates": [g("CH", 0, 1), g("CH", 3, 2)],
     "expect_product": [Z0, Z0, O1, Z0]},
    {"name": "CH_active", "lines": "986-1002", "n": 4,
     "init": [O1, Z0, O1, O1], "gates": [g("CH", 0, 1), g("CH", 3, 2)],
     "expect_product": [O1, [R2, 0.0, R2, 0.0], [R2, 0.0, -R2, 0.0], O1]},
]

# Per-qubit probabilities (test/tests.cpp:1141-1194).
PROB_CASES = [
    {"name": "prob_zero", "lines": "1146-1154", "n": 4, "init": [Z0, Z0, Z0, Z0],
     "p1": [0.0, 0.0, 0.0, 0.0], "tol": 0},
    {"name": "prob_last_one", "lines": "1156-1166", "n": 4, "init": [Z0, Z0, Z0, [0.0, 0.0, 1.0, 0.0]],
     "p1": [0.0, 0.0, 0.0, 1.0], "tol": 0},
    {"name": "prob_last_plus", "lines": "1168-1177", "n": 4, "init": [Z0, Z0, Z0, [1.0, 0.0, 1.0, 0.0]],
     "p1": [0.0, 0.0, 0.0, 0.5], "tol": 1e-4},
    {"name": "prob_first_A", "lines": "1179-1191", "n": 4, "init": [A, Z0, Z0, [1.0, 0.0, 1.0, 0.0]],
     "p1": [0.64, 0.0, 0.0, 0.5], "tol": 1e-4},
]

# Measurement (test/tests.cpp:274-336): deterministic parts.
MEAS_CASES = [
    {"name": "peek_all_zero", "lines": "281-288", "n": 3, "init": [Z0, Z0, Z0], "peek_all": 0},
    {"name": "peek_all_101", "lines": "290-301", "n": 3, "init": [O1, Z0, O1], "peek_all": 0b101},
]


# Host QState known answers (test/tests.cpp:42-245).  "in" is the pair of
# amplitudes handed to the constructor (re0, im0, re1, im1), before the
# constructor's FP32 normalisation; "op" is what is applied; "expect" the
# resulting pair; tol 0 = exact equality, else fequal's relative tolerance.
QSTATE_CASES = [
    {"name": "default", "lines": "42-58", "op": "default", "in": None, "expect": [1.0, 0.0, 0.0, 0.0], "tol": 0},
    {"name": "ctor", "lines": "42-58", "op": "none", "in": [0.6, 0.0, 0.0, 0.8], "expect": [0.6, 0.0, 0.0, 0.8],
     "tol": 0},
    {"name": "copy", "lines": "42-73", "op": "copy", "in": [0.6, 0.0, 0.0, 0.8], "expect": [0.6, 0.0, 0.0, 0.8],
     "tol": 0},
    {"name": "normalize_34", "lines": "78-97", "op": "none", "in": [3.0, 0.0, 4.0, 0.0],
     "expect": [0.6, 0.0, 0.8, 0.0], "tol": 0},
    {"name": "X_op", "lines": "195-206", "op": "X", "in": [0.5, -0.5, 0.7, 0.1], "expect": [0.7, 0.1, 0.5, -0.5],
     "tol": 0},
    {"name": "Y_op", "lines": "195-212", "op": "Y", "in": [0.5, -0.5, 0.7, 0.1], "expect": [0.1, -0.7, 0.5, 0.5],
     "tol": 0},
    {"name": "Z_op", "lines": "195-217", "op": "Z", "in": [0.5, -0.5, 0.7, 0.1], "expect": [0.5, -0.5, -0.7, -0.1],
     "tol": 0},
    # H_op is compared with QState({1.2,-0.4},{-0.2,-0.6}) after its normalisation
    {"name": "H_op", "lines": "219-226", "op": "H", "in": [0.5, -0.5, 0.7, 0.1],
     "expect_normalized": [1.2, -0.4, -0.2, -0.6], "tol": 1e-4},
]
# Probabilities of a QState (test/tests.cpp:102-131): in -> (P(0), P(1)).
QSTATE_PROB_CASES = [
    {"in": [1.0, 0.0, 0.0, 0.0], "p0": 1.0, "p1": 0.0, "tol": 0},
    {"in": [0.0, 1.0, 0.0, 0.0], "p0": 1.0, "p1": 0.0, "tol": 0},
    {"in": [0.0, 0.0, 1.0, 0.0], "p0": 0.0, "p1": 1.0, "tol": 0},
    {"in": [0.0, 0.0, 0.0, 1.0], "p0": 0.0, "p1": 1.0, "tol": 0},
    {"in": [0.0, 1.0, 0.0, 1.0], "p0": 0.5, "p1": 0.5, "tol": 1e-4},
    {"in": [0.6, 0.0, 0.0, 0.8], "p0": 0.36, "p1": 0.64, "tol": 1e-4},
]

# QSimulator construction (test/tests.cpp:251-267): every qubit set to `qubit`,
# amplitude `index` of the generated state.
QSIM_CASES = [
    {"name": "all_same", "lines": "251-267", "n": 3, "qubit": [0.6, 0.0, 0.0, 0.8], "index": 0,
     "expect": [0.216, 0.0], "tol": 1e-4},
]

# Collapse (test/tests.cpp:303-336).  measure_all leaves the basis state it
# returned; measure(q) keeps the outcome's half, renormalised: "after_one" /
# "after_zero" list amplitudes for either outcome.
COLLAPSE_CASES = [
    {"name": "measure_all", "lines": "303-313", "n": 3, "init": [[0.0, 1.0, -1.0, 0.0], Z0, [0.6, 0.0, 0.0, 0.8]],
     "op": "measure_all", "support": [0b000, 0b001, 0b100, 0b101]},
    {"name": "measure_q0", "lines": "315-336", "n": 3, "init": [[0.0, 1.0, -1.0, 0.0], Z0, [0.6, 0.0, 0.0, 0.8]],
     "op": "measure", "qubit": 0,
     "after_one": [[0b100, -0.6, 0.0], [0b101, 0.0, -0.8]],
     "after_zero": [[0b000, 0.0, 0.6], [0b001, -0.8, 0.0]], "tol": 1e-4},
]

# Statistical profiles (test/tests.cpp:136-190, 343-406): counts over `reps`
# shots must fall in the 99.9% band (z = 3.291) around reps * p.
PROFILE_CASES = [
    {"name": "qsim_one_qubit_set", "lines": "349-381", "n": 5, "reps": 10000,
     "init": [Z0, Z0, Z0, O1, Z0]},
    {"name": "qsim_mixed", "lines": "383-403", "n": 5, "reps": 10000,
     "init": [[1.0, 0.0, 1.0, 0.0], [-2.0, 0.0, 3.0, 0.0], Z0, O1, [0.0, 0.6, 0.8, 0.0]]},
]


# Builders of quantum_gates.h (test/tests.cpp:1011-1135) as data.  A builder
# node: {"builder": name, args...}; "Circuit"/"ControlCircuit" nodes wrap a
# sub-circuit (gate list or builder) at a position.
SPECIAL_PEEK = [
    # NControl_Gate(6, {0,2,4,5}, 3, X) embedded at 0; init per qubit, expected peek_measure_all
    {"name": "ncontrol_6", "lines": "1015-1047", "n": 6,
     "circuit": [{"kind": "Circuit", "begin": 0,
                  "builder": {"builder": "NControl_list", "qubits": 6, "controls": [0, 2, 4, 5], "target": 3,
                              "gate": "X"}}],
     "runs": [{"init": [O1] * 6, "peek_all": 0b111011},
              {"init": [O1, O1, O1, Z0, O1, O1], "peek_all": 0b111111},
              {"init": [O1, Z0, O1, Z0, O1, O1], "peek_all": 0b101111},
              {"init": [O1, Z0, O1, O1, O1, O1], "peek_all": 0b101011}]},
]
SPECIAL_EQUALITIES = [
    {"name": "ncontrol_one_below", "lines": "1049-1060", "n": 4,
     "lhs": [{"kind": "Circuit", "begin": 0,
              "builder": {"builder": "NControl_list", "qubits": 4, "controls": [3], "target": 0, "gate": "X"}}],
     "rhs": [g("CX", 3, 0)], "tol": 0},
    {"name": "ncontrol_above_below", "lines": "1062-1071", "n": 4,
     "lhs": [{"kind": "Circuit", "begin": 0,
              "builder": {"builder": "NControl_list", "qubits": 4, "controls": [0, 3], "target": 2, "gate": "X"}}],
     "rhs": [g("CCX", 0, 3, 2)], "tol": 0},
    {"name": "control_group", "lines": "1074-1084", "n": 4,
     "lhs": [{"kind": "Circuit", "begin": 0,
              "builder": {"builder": "Control_Group", "qubits": 4, "control": 1, "targets": [0, 2, 3], "gate": "X"}}],
     "rhs": [g("CX", 1, 0), g("CX", 1, 2), g("CX", 1, 3)], "tol": 0},
    {"name": "group", "lines": "1086-1094", "n": 4,
     "lhs_builder": {"builder": "Group", "qubits": 4, "targets": [0, 2, 3], "gate": "X", "compile": True},
     "rhs": [g("X", 0), g("X", 2), g("X", 3)], "tol": 0},
    {"name": "rewire", "lines": "1096-1108", "n": 4,
     "lhs_builder": {"builder": "Rewire", "qubits": 4, "positions": [2, 1, 3, 0], "compile": True,
                     "sub_n": 4, "sub": [g("H", 0), g("CX", 0, 2), g("H", 1), g("Z", 3)]},
     "rhs": [g("H", 2), g("CX", 2, 3), g("H", 1), g("Z", 0)], "tol": 0},
    {"name": "adjoint", "lines": "1110-1131", "n": 4,
     "lhs_builder": {"builder": "Adjoint", "sub_n": 4,
                     "sub": [g("Swap", 1, 2), g("CPhase", 2, 1, angle=f32(-PI / 3)),
                             g("CRotX", 1, 3, angle=f32(PI / 4)), g("X", 0),
                             g("CRotY", 3, 0, angle=f32(-PI / 2))]},
     "rhs": [g("CRotY", 3, 0, angle=f32(PI / 2)), g("X", 0), g("CRotX", 1, 3, angle=f32(-PI / 4)),
             g("CPhase", 2, 1, angle=f32(PI / 3)), g("Swap", 1, 2)], "tol": 1e-5},
]


def main():
    out = {"source": "test/tests.cpp of arrayfire/afQuantumSim (transcribed)",
           "qubit_order": "qubit 0 is the most significant bit",
           "gate_cases": CASES, "equalities": EQUALITIES, "ch_cases": CH_CASES,
           "prob_cases": PROB_CASES, "measure_cases": MEAS_CASES,
           "qstate_cases": QSTATE_CASES, "qstate_prob_cases": QSTATE_PROB_CASES, "qsim_cases": QSIM_CASES,
           "collapse_cases": COLLAPSE_CASES, "profile_cases": PROFILE_CASES,
           "special_peek": SPECIAL_PEEK, "special_equalities": SPECIAL_EQUALITIES}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_gates.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, len(CASES), "gate cases")


if __name__ == "__main__":
    main()
