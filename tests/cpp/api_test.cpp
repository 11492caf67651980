// End-to-end tests of the aqs:: C++ API (libaqs.so over libaqs_hip.so).
//
//   api_test run  <kat.json>   known answers of the reference's test suite,
//                              read from tests/golden/kat_gates.json, plus the
//                              checks below; prints "ALL PASSED"
//   api_test dump <out.json>   runs API-only paths (builders, set_basis,
//                              linear_entanglement_varstate, measure_all) and
//                              writes their states / matrices for the Python
//                              tests to compare with numpy and the oracle
//
// Tolerances: fixture cases carry the reference's own (relative 1e-4 through
// its fequal, test/tests.cpp:19-27; exact where it asserts ==).  The
// reference's exact equalities between two different float computations
// (CH, test/tests.cpp:986-1002; circuit matrices, :909-961) are checked to
// 1e-6 absolute: 0.70710677^2 and 0.7071068^2 differ in IEEE FP32.
#include <cmath>
#include <cstdio>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "aqs/quantum.h"
#include "aqs/quantum_algo.h"
#include "aqs/quantum_gates.h"
#include "mini_json.h"

using namespace aqs;
using J = mini_json::Value;

namespace {

int g_checks = 0;
std::string g_where;

void require(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) {
        std::fprintf(stderr, "FAILED [%s] %s\n", g_where.c_str(), what.c_str());
        std::exit(1);
    }
}

// |a - b| <= max(|a|, |b|) * rel; rel = 0 means a == b.
bool rel_close(float a, float b, float rel) {
    if (rel == 0.f) return a == b;
    return std::fabs(a - b) <= std::fmax(std::fabs(a), std::fabs(b)) * rel;
}
bool crel_close(cfloat a, cfloat b, float rel) { return rel_close(a.real(), b.real(), rel) && rel_close(a.imag(), b.imag(), rel); }

float max_abs_diff(const std::vector<cfloat>& a, const std::vector<cfloat>& b) {
    if (a.size() != b.size()) return INFINITY;
    float m = 0.f;
    for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::abs(a[i] - b[i]));
    return m;
}
float max_abs_diff(const CircuitMatrix& a, const CircuitMatrix& b) {
    if (a.dim != b.dim) return INFINITY;
    return max_abs_diff(a.data, b.data);
}

template <typename Ex>
bool throws(const std::function<void()>& f, const std::string& message = "") {
    try {
        f();
    } catch (const Ex& e) {
        return message.empty() || message == e.what();
    } catch (...) {
        return false;
    }
    return false;
}

// ---- fixtures -> API objects ----------------------------------------------

QState qstate_of(const J& v) {
    if (v.type == J::Type::String) {
        const std::string& s = v.s();
        if (s == "0") return QState::zero();
        if (s == "1") return QState::one();
        if (s == "-") return QState::minus();
        if (s == "+") return QState::plus();
        throw std::runtime_error("unknown qubit state " + s);
    }
    return QState({v[0].f(), v[1].f()}, {v[2].f(), v[3].f()});
}

std::vector<QState> qstates_of(const J& list) {
    std::vector<QState> out;
    for (const auto& v : list.arr) out.push_back(qstate_of(v));
    return out;
}

// The one-qubit circuit a builder takes (X::gate(), Z::gate(), ...).
const QCircuit& named_gate(const std::string& name) {
    if (name == "X") return X::gate();
    if (name == "Y") return Y::gate();
    if (name == "Z") return Z::gate();
    if (name == "H") return H::gate();
    throw std::runtime_error("unknown named gate " + name);
}

void add_nodes(QCircuit& qc, const J& nodes);
QCircuit build_builder(const J& b);

QCircuit circuit_of(uint32_t n, const J& nodes) {
    QCircuit qc(n);
    add_nodes(qc, nodes);
    return qc;
}

void add_gate(QCircuit& qc, const J& g) {
    const std::string& k = g["kind"].s();
    if (k == "Circuit" || k == "ControlCircuit") {
        const QCircuit sub = g.has("builder") ? build_builder(g["builder"]) : circuit_of(g["sub_n"].u(), g["sub"]);
        if (k == "Circuit")
            qc << Gate(sub, g["begin"].u());
        else
            qc << ControlGate(sub, g["control"].u(), g["begin"].u());
        return;
    }
    const uint32_t a = g["q"][0].u(), b = g["q"][1].u(), c = g["q"][2].u();
    const float t = g["angle"].f();
    if (k == "X") qc << X{a};
    else if (k == "Y") qc << Y{a};
    else if (k == "Z") qc << Z{a};
    else if (k == "H") qc << H{a};
    else if (k == "Phase") qc << Phase{a, t};
    else if (k == "RotX") qc << RotX{a, t};
    else if (k == "RotY") qc << RotY{a, t};
    else if (k == "RotZ") qc << RotZ{a, t};
    else if (k == "CX") qc << CX{a, b};
    else if (k == "CY") qc << CY{a, b};
    else if (k == "CZ") qc << CZ{a, b};
    else if (k == "CH") qc << CH{a, b};
    else if (k == "CPhase") qc << CPhase{a, b, t};
    else if (k == "CRotX") qc << CRotX{a, b, t};
    else if (k == "CRotY") qc << CRotY{a, b, t};
    else if (k == "CRotZ") qc << CRotZ{a, b, t};
    else if (k == "Swap") qc << Swap{a, b};
    else if (k == "CSwap") qc << CSwap{a, b, c};
    else if (k == "CCX") qc << CCNot{a, b, c};
    else if (k == "Or") qc << Or{a, b, c};
    else throw std::runtime_error("unknown gate kind " + k);
}

void add_nodes(QCircuit& qc, const J& nodes) {
    for (const auto& g : nodes.arr) add_gate(qc, g);
}

std::vector<uint32_t> uints(const J& list) {
    std::vector<uint32_t> out;
    for (const auto& v : list.arr) out.push_back(v.u());
    return out;
}

QCircuit build_builder(const J& b) {
    const std::string& k = b["builder"].s();
    const bool compile = b.has("compile") && b["compile"].boolean();
    if (k == "NControl_list")
        return NControl_Gate(b["qubits"].u(), uints(b["controls"]), b["target"].u(), named_gate(b["gate"].s()), compile);
    if (k == "NControl_range")
        return NControl_Gate(b["qubits"].u(), b["control_begin"].u(), b["control_count"].u(), b["target"].u(),
                             named_gate(b["gate"].s()), compile);
    if (k == "Control_Group")
        return Control_Group_Gate(b["qubits"].u(), b["control"].u(), uints(b["targets"]), named_gate(b["gate"].s()),
                                  compile);
    if (k == "Group") return Group_Gate(b["qubits"].u(), uints(b["targets"]), named_gate(b["gate"].s()), compile);
    if (k == "Rewire") {
        QCircuit sub = circuit_of(b["sub_n"].u(), b["sub"]);
        sub.compile();
        return Rewire_Gate(b["qubits"].u(), uints(b["positions"]), sub, compile);
    }
    if (k == "Adjoint") {
        QCircuit sub = circuit_of(b["sub_n"].u(), b["sub"]);
        sub.compile();
        return Adjoint_Gate(sub);
    }
    throw std::runtime_error("unknown builder " + k);
}

QSimulator simulator_of(const J& init) {
    const auto states = qstates_of(init);
    return QSimulator((uint32_t)states.size(), states);
}

// ---- fixture sections ------------------------------------------------------

// Gate known answers: each case runs uncompiled (one pass per gate) and
// compiled (fused passes); listed amplitudes must match, and when they carry
// all the probability every other amplitude must be zero.
void run_gate_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        for (bool compiled : {false, true}) {
            g_where = c["name"].s() + (compiled ? " (compiled)" : "") + " test/tests.cpp:" + c["lines"].s();
            QSimulator qs = simulator_of(c["init"]);
            QCircuit qc = circuit_of(c["n"].u(), c["gates"]);
            if (compiled) qc.compile();
            qs.simulate(qc);
            const auto sv = qs.statevector().host();
            const float tol = c["tol"].f();
            std::vector<bool> listed(sv.size(), false);
            double listed_p = 0.0;
            for (const auto& e : c["expect"].arr) {
                const uint32_t idx = e[0].u();
                const cfloat want{e[1].f(), e[2].f()};
                listed[idx] = true;
                listed_p += std::norm(std::complex<double>(want));
                std::ostringstream msg;
                msg << "amplitude " << idx << " = " << sv[idx] << ", expected " << want;
                require(crel_close(sv[idx], want, tol) || (tol > 0 && std::abs(sv[idx] - want) < 1e-7f), msg.str());
            }
            if (std::fabs(listed_p - 1.0) < 1e-4)
                for (size_t i = 0; i < sv.size(); ++i)
                    if (!listed[i]) require(std::abs(sv[i]) <= 1e-6f, "unlisted amplitude " + std::to_string(i) + " is not 0");
        }
    }
}

// Unitaries of two circuits agree (the reference compares compiled matrices).
void run_equalities(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        const uint32_t n = c["n"].u();
        QCircuit lhs = c.has("lhs_builder") ? build_builder(c["lhs_builder"]) : circuit_of(n, c["lhs"]);
        QCircuit rhs = circuit_of(n, c["rhs"]);
        lhs.compile();
        rhs.compile();
        const float tol = c.has("tol") && c["tol"].d() > 0 ? c["tol"].f() : 1e-6f;
        const float diff = max_abs_diff(lhs.circuit(), rhs.circuit());
        require(diff <= tol, "circuit matrices differ by " + std::to_string(diff));
    }
}

void run_ch_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        QSimulator qs = simulator_of(c["init"]);
        QSimulator want = simulator_of(c["expect_product"]);
        qs.simulate(circuit_of(c["n"].u(), c["gates"]));
        require(max_abs_diff(qs.statevector().host(), want.statevector().host()) <= 1e-6f, "state differs from product");
    }
}

void run_prob_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        QSimulator qs = simulator_of(c["init"]);
        for (uint32_t q = 0; q < qs.qubit_count(); ++q) {
            const float want = c["p1"][q].f();
            const float tol = want == 0.f ? 0.f : c["tol"].f();
            require(rel_close(qs.qubit_probability_true(q), want, tol), "P(qubit " + std::to_string(q) + " = 1)");
            require(rel_close(qs.qubit_probability_false(q), 1.f - want, tol == 0.f ? 0.f : tol),
                    "P(qubit " + std::to_string(q) + " = 0)");
        }
    }
}

void run_measure_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        QSimulator qs = simulator_of(c["init"]);
        const uint32_t want = c["peek_all"].u();
        require(qs.peek_measure_all() == want, "peek_measure_all");
        for (uint32_t q = 0; q < qs.qubit_count(); ++q)
            require(qs.peek_measure(q) == (((want >> (qs.qubit_count() - 1 - q)) & 1u) != 0),
                    "peek_measure(" + std::to_string(q) + ")");
    }
}

void run_qstate_cases(const J& cases, const J& prob_cases) {
    for (const auto& c : cases.arr) {
        g_where = "QState " + c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        QState in = c["in"].is_null() ? QState{} : qstate_of(c["in"]);
        const std::string& op = c["op"].s();
        QState out = in;
        if (op == "copy") {
            QState copy(in);
            QState assigned;
            assigned = copy;
            out = assigned;
        } else if (op == "X") out = X_op(in);
        else if (op == "Y") out = Y_op(in);
        else if (op == "Z") out = Z_op(in);
        else if (op == "H") out = Hadamard_op(in);
        const float tol = c["tol"].f();
        if (c.has("expect")) {
            const cfloat w0{c["expect"][0].f(), c["expect"][1].f()}, w1{c["expect"][2].f(), c["expect"][3].f()};
            require(crel_close(out[0], w0, tol) && crel_close(out[1], w1, tol), "amplitudes");
        } else {
            const QState want = qstate_of(c["expect_normalized"]);
            require(crel_close(out[0], want[0], tol) && crel_close(out[1], want[1], tol), "amplitudes (normalised)");
        }
    }
    {
        // normalisation of a complex pair: each amplitude divided by the FP32
        // magnitude (test/tests.cpp:92-103)
        g_where = "QState normalisation test/tests.cpp:92-103";
        const QState q({1.f, 2.f}, {3.f, 4.f});
        const float mag = std::sqrt(1.f * 1.f + 2.f * 2.f + 3.f * 3.f + 4.f * 4.f);
        require(q[0] == cfloat(1.f, 2.f) / mag && q[1] == cfloat(3.f, 4.f) / mag, "complex pair");
    }
    for (const auto& c : prob_cases.arr) {
        g_where = "QState probability test/tests.cpp:102-131";
        const QState q = qstate_of(c["in"]);
        const float tol = c["tol"].f();
        require(rel_close(q.probability_false(), c["p0"].f(), tol) && rel_close(q.probability_true(), c["p1"].f(), tol),
                "probabilities");
    }
}

void run_qsim_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = "QSimulator " + c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        const QState q = qstate_of(c["qubit"]);
        QSimulator qs(c["n"].u(), q);
        require(qs.qubit_count() == c["n"].u(), "qubit count");
        for (uint32_t i = 0; i < qs.qubit_count(); ++i) require(qs.qubit(i) == q, "per-qubit state");
        const cfloat want{c["expect"][0].f(), c["expect"][1].f()};
        require(crel_close(qs.state(c["index"].u()), want, c["tol"].f()), "amplitude");
    }
}

void run_collapse_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        const std::string& op = c["op"].s();
        // several draws, so both branches of a measurement are reached
        for (int draw = 0; draw < 16; ++draw) {
            QSimulator qs = simulator_of(c["init"]);
            if (op == "measure_all") {
                const uint32_t m = qs.measure_all();
                bool in_support = false;
                for (const auto& s : c["support"].arr) in_support |= (s.u() == m);
                require(in_support, "outcome outside the support");
                require(qs.state(m) == cfloat(1.f, 0.f), "collapsed amplitude is 1");
                require(qs.peek_measure_all() == m, "peek after collapse");
                const auto sv = qs.statevector().host();
                for (uint32_t i = 0; i < sv.size(); ++i)
                    if (i != m) require(sv[i] == cfloat(0.f, 0.f), "other amplitudes are 0");
            } else {
                const uint32_t q = c["qubit"].u();
                const bool one = qs.measure(q);
                const uint32_t seen = qs.peek_measure_all();
                require(one == (((seen >> (qs.qubit_count() - 1 - q)) & 1u) != 0), "measured qubit stays");
                for (const auto& e : c[one ? "after_one" : "after_zero"].arr)
                    require(crel_close(qs.state(e[0].u()), cfloat(e[1].f(), e[2].f()), c["tol"].f()),
                            "post-measurement amplitude " + std::to_string(e[0].u()));
                require(std::fabs(qs.qubit_probability_true(q) - (one ? 1.f : 0.f)) < 1e-6f, "P after collapse");
            }
        }
    }
}

// Counts in the 99.9% band around reps * p (z = 3.291).
bool in_band(uint32_t count, float p, uint32_t reps) {
    const float sd = std::sqrt(p * (1.f - p));
    const int expected = (int)(reps * p);
    const int half = (int)(std::sqrt((float)reps) * 3.291f * sd);
    return expected - half <= (int)count && (int)count <= expected + half;
}

void run_profile_cases(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        QSimulator qs = simulator_of(c["init"]);
        const uint32_t reps = c["reps"].u();
        const auto counts = qs.profile_measure_all(reps);
        uint64_t total = 0;
        for (uint32_t i = 0; i < counts.size(); ++i) {
            total += counts[i];
            require(in_band(counts[i], qs.state_probability(i), reps), "profile_measure_all count " + std::to_string(i));
        }
        require(total == reps, "profile_measure_all total");
        for (uint32_t q = 0; q < qs.qubit_count(); ++q) {
            const auto pr = qs.profile_measure(q, reps);
            require(pr[0] + pr[1] == reps, "profile_measure total");
            require(in_band(pr[1], qs.qubit(q).probability_true(), reps), "profile_measure(" + std::to_string(q) + ")");
        }
    }
    g_where = "QState::profile_measure test/tests.cpp:136-190";
    for (const QState& q : {QState({1.f, 0.f}, {0.f, 0.f}), QState({0.f, 1.f}, {0.f, 1.f}),
                            QState({0.6f, 0.f}, {0.f, 0.8f}), QState({0.f, 0.f}, {0.f, 1.f})}) {
        const auto pr = q.profile_measure(10000);
        require(pr[0] + pr[1] == 10000 && in_band(pr[1], q.probability_true(), 10000), "QState profile");
    }
}

void run_special_peek(const J& cases) {
    for (const auto& c : cases.arr) {
        g_where = c["name"].s() + " test/tests.cpp:" + c["lines"].s();
        QCircuit qc = circuit_of(c["n"].u(), c["circuit"]);
        qc.compile();
        for (const auto& r : c["runs"].arr) {
            QSimulator qs = simulator_of(r["init"]);
            qs.simulate(qc);
            require(qs.peek_measure_all() == r["peek_all"].u(), "peek_measure_all after the builder circuit");
        }
    }
}

// ---- checks of this build's own ------------------------------------------

// Builders give the reference's circuits (src/quantum_gates.cpp:36-178,
// 232-338), including where its index arithmetic packs or moves targets.
void builder_behaviour() {
    g_where = "Control_Group_Gate packs a group's targets";
    {
        // targets {0, 2} below control 3: the reference's group circuit has two
        // qubits placed at qubit 0, so X acts on qubits 0 and 1
        QCircuit got(4);
        got << Gate(Control_Group_Gate(4, 3, {0, 2}, X::gate()), 0);
        QCircuit want(4);
        want << CX{3, 0} << CX{3, 1};
        got.compile();
        want.compile();
        require(max_abs_diff(got.circuit(), want.circuit()) == 0.f, "packed placement");
        require(throws<std::invalid_argument>([] { Control_Group_Gate(4, 2, {0, 2}, X::gate()); },
                                              "Cannot add control gate at the target qubit positions"),
                "control among the targets");
        require(throws<std::invalid_argument>([] { Control_Group_Gate(3, 0, {0, 1, 2}, X::gate()); },
                                              "Cannot add control gate at the given position"),
                "too many targets");
        require(throws<std::invalid_argument>([] { Control_Group_Gate(3, 3, {0}, X::gate()); },
                                              "Invalid control qubit position"),
                "control out of range");
    }
    g_where = "NControl_Gate widths and placements";
    {
        // controls only below the targets: the circuit spans qubits - target_begin
        QCircuit below = NControl_Gate(5, std::vector<uint32_t>{3}, 1, X::gate());
        require(below.qubit_count() == 4, "width qubits - target_begin");
        QCircuit got(5), want(5);
        got << Gate(below, 1);
        want << CX{3, 1};
        got.compile();
        want.compile();
        require(max_abs_diff(got.circuit(), want.circuit()) == 0.f, "below-only, placed at target_begin");
        // one control above, none below: the gate lands on the last qubit
        QCircuit one_above = NControl_Gate(5, std::vector<uint32_t>{0}, 1, X::gate());
        require(one_above.qubit_count() == 5, "width with a control above");
        QCircuit moved(5);
        moved << CX{0, 4};
        one_above.compile();
        moved.compile();
        require(max_abs_diff(one_above.circuit(), moved.circuit()) == 0.f, "single control above moves the gate");
        // two controls above and one below: every qubit where it belongs
        QCircuit mixed = NControl_Gate(5, std::vector<uint32_t>{0, 1, 4}, 2, X::gate());
        // qubits 0, 1 and 4 set flip qubit 2; nothing else changes a basis state
        for (uint32_t idx = 0; idx < 32; ++idx) {
            std::vector<QState> init;
            for (uint32_t q = 0; q < 5; ++q) init.push_back((idx >> (4 - q)) & 1u ? QState::one() : QState::zero());
            QSimulator qs(5, init);
            qs.simulate(mixed);
            const bool fire = (idx & 0b11001u) == 0b11001u;
            require(qs.peek_measure_all() == (fire ? (idx ^ 0b00100u) : idx), "3-controlled X on basis state");
        }
        // the range form: consecutive controls right above the target
        QCircuit range = NControl_Gate(5, 1, 2, 3, X::gate());
        QCircuit ccx(5);
        ccx << CCNot{1, 2, 3};
        range.compile();
        ccx.compile();
        require(max_abs_diff(range.circuit(), ccx.circuit()) == 0.f, "range form = CCNot");
        require(throws<std::invalid_argument>([] { NControl_Gate(5, 0, 3, 2, X::gate()); },
                                              "Invalid control_qubit position"),
                "range overlapping the target");
        require(throws<std::invalid_argument>([] { NControl_Gate(4, std::vector<uint32_t>{1, 2}, 2, X::gate()); },
                                              "Cannot add control gate at the target qubit positions"),
                "control on the target");
    }
    g_where = "Adjoint_Gate inverts the shared gate objects";
    {
        QCircuit c(3);
        c << RotX{0, 0.5f} << CPhase{0, 1, 0.25f} << H{2} << RotZ{1, -1.5f};
        QCircuit adj = Adjoint_Gate(c);
        // the input shares the gate objects, so its angles are negated too
        require(static_cast<const RotX&>(*c.gate_list()[0]).angle == -0.5f, "input RotX negated");
        require(static_cast<const CPhase&>(*c.gate_list()[1]).angle == -0.25f, "input CPhase negated");
        require(static_cast<const RotZ&>(*c.gate_list()[3]).angle == 1.5f, "input RotZ negated");
        require(adj.representation() == "RotZ,0,1:1;H,0,1:2;Phase,1,1:0,1;RotX,0,1:0;", adj.representation());
        // adjoint times the original (as it was) is the identity
        QCircuit orig(3);
        orig << RotX{0, 0.5f} << CPhase{0, 1, 0.25f} << H{2} << RotZ{1, -1.5f};
        QCircuit both(3);
        both << Gate(orig, 0) << Gate(adj, 0);
        both.compile();
        const auto& m = both.circuit();
        float err = 0.f;
        for (uint32_t r = 0; r < m.dim; ++r)
            for (uint32_t col = 0; col < m.dim; ++col)
                err = std::fmax(err, std::abs(m(r, col) - cfloat(r == col ? 1.f : 0.f, 0.f)));
        require(err < 1e-5f, "adjoint * circuit = identity");
        // CH has no case in the reference's switch
        QCircuit with_ch(2);
        with_ch << RotY{0, 0.3f} << CH{0, 1} << RotX{1, 0.7f};
        require(throws<std::runtime_error>([&] { Adjoint_Gate(with_ch); }, "Unknown unsupported gate cannot be adjoint"),
                "CH throws");
        // the gate after CH (program order) was already inverted when it threw
        require(static_cast<const RotX&>(*with_ch.gate_list()[2]).angle == -0.7f, "partial inversion before the throw");
        require(static_cast<const RotY&>(*with_ch.gate_list()[0]).angle == 0.3f, "gates before CH untouched");
    }
}

// Representation strings (src/quantum.cpp:569-571, 885-896, 1331-1335,
// 1662-1677) and the cache keyed by them (:1705-1724).
void representations_and_cache() {
    g_where = "representation strings";
    QCircuit r(3);
    r << H{0} << CX{0, 1} << Phase{2, pi / 4} << CPhase{1, 2, -pi / 2} << CSwap{0, 1, 2} << Barrier{} << Barrier{false};
    require(r.representation() == "H,0,1:0;X,1,1:0,1;T,0,1:2;S†,1,1:1,2;Swap,1,2:0,12;B;P;", r.representation());
    QCircuit sub(2);
    sub << H{0} << CX{0, 1};
    require(Gate(sub, 1).to_string() == "H,0,1:1;X,1,1:1,2;", "Gate shifts its sub-circuit");
    require(ControlGate(sub, 0, 1).to_string() == "H,1,1:0,1;X,2,1:0,1,2;", "ControlGate adds its control");
    require(Gate(sub, 1, "Bell").to_string() == "Bell,0,2:1,2;", "named Gate");
    require(ControlGate(sub, 0, 1, "Bell").to_string() == "Bell,1,2:0,1,2;", "named ControlGate");
    require(throws<std::invalid_argument>([&] { Gate(sub, 0, "a;b"); }, "Name cannot contain commas, colons, nor semicolons"),
            "name with a separator");

    g_where = "circuit cache";
    clear_circuit_cache();
    QCircuit a = NControl_Gate(6, 0, 5, 5, Z::gate());
    QCircuit b = NControl_Gate(6, 0, 5, 5, Z::gate());
    require(Gate(a, 0).internal_circuit == Gate(b, 0).internal_circuit, "rebuilt circuit shares the entry");
    QCircuit o1(3), o2(3);
    o1 << X{0};
    o2 << H{1};
    QCircuit c1(3), c2(3);
    c1 << Gate{o1, 0, "Orc"};
    c2 << Gate{o2, 0, "Orc"};
    require(c1.representation() == c2.representation(), "named gates share a key");
    QCircuit w1(3), w2(3);
    w1 << Gate{c1, 0};
    w2 << Gate{c2, 0};
    QSimulator s1(3), s2(3);
    s1.simulate(w1);
    s2.simulate(w2);
    require(s1.state(4) == cfloat(1.f, 0.f), "X on qubit 0");
    require(std::fabs(s2.state(0).real() - 0.70710678f) < 1e-6f && std::fabs(s2.state(2).real() - 0.70710678f) < 1e-6f,
            "H on qubit 1");
    clear_circuit_cache();
}

// Exception types and messages of the reference (src/quantum.cpp).
void errors() {
    g_where = "errors";
    QCircuit qc(3);
    require(throws<std::out_of_range>([&] { qc << X{3}; }, "Cannot add gate at the given qubit position"), "X out of range");
    require(throws<std::invalid_argument>([&] { qc << CX{1, 1}; }), "CX on one qubit");
    require(throws<std::domain_error>([] {
                QCircuit one(1);
                one << CX{0, 1};
            }),
            "CX on a 1-qubit circuit");
    require(throws<std::invalid_argument>([] { QCircuit bad(31); }), "31 qubits");
    require(throws<std::invalid_argument>([] { QCircuit bad(0); }), "0 qubits");
    QCircuit two(2);
    two << H{0};
    require(throws<std::out_of_range>([&] { qc << Gate(two, 2); }, "Cannot add gate at the given qubit position"),
            "Gate past the end");
    require(throws<std::out_of_range>([&] { qc << ControlGate(two, 1, 0); },
                                      "Control qubit cannot be one of the target qubits of the gate"),
            "ControlGate control inside");
    require(throws<std::invalid_argument>([&] {
                QSimulator qs(2);
                qs.simulate(qc);
            }),
            "simulate with mismatched qubits");
    require(qc.gate_list().empty(), "failed insertions add nothing");
    QSimulator qs(3);
    require(throws<std::out_of_range>([&] { qs.qubit_probability_true(3); }), "probability of a missing qubit");
    require(throws<std::out_of_range>([&] { qs.measure(3); }), "measure a missing qubit");
}

// compile() (fused passes, or LDS tile passes from 20 qubits) against the
// per-gate path, through the API.
void compile_equivalence() {
    g_where = "compile equivalence";
    for (uint32_t n : {6u, 12u}) {
        QCircuit a = fourier_transform(n, false);
        QCircuit b = fourier_transform(n, true);
        require(a.representation() == b.representation(), "same representation");
        QSimulator sa(n, QState::plus()), sb(n, QState::plus());
        sa.simulate(a);
        sb.simulate(b);
        require(max_abs_diff(sa.statevector().host(), sb.statevector().host()) < 2e-6f, "QFT compiled = uncompiled");
    }
    const uint32_t n = 22;
    QCircuit a(n), b(n);
    for (QCircuit* qc : {&a, &b})
        for (uint32_t layer = 0; layer < 6; ++layer) {
            for (uint32_t q = 0; q < n; ++q) *qc << U3{q, 0.3f + 0.1f * q, 0.2f * layer, -0.7f + 0.05f * q};
            for (uint32_t q = layer % 2; q + 1 < n; q += 2) *qc << CZ{q, q + 1};
        }
    b.compile();
    require(b.compiled_passes() * 4 < b.gate_list().size(), "fused pass count");
    QSimulator sa(n), sb(n);
    sa.simulate(a);
    sb.simulate(b);
    require(max_abs_diff(sa.statevector().host(), sb.statevector().host()) < 1e-5f, "22-qubit brickwork");
}

// Grover: after k iterations from |0..0> the marked index has probability
// sin^2((2k + 1) theta), sin theta = 2^(-n/2).
void grover() {
    g_where = "grover";
    const uint32_t n = 4, marked = 5, marked_index = 0b1010;
    QCircuit oracle = grover_oracle(n, marked);
    for (uint32_t i = 0; i < (1u << n); ++i) {
        std::vector<QState> init;
        for (uint32_t q = 0; q < n; ++q) init.push_back((i >> (n - 1 - q)) & 1u ? QState::one() : QState::zero());
        QSimulator qs(n, init);
        qs.simulate(oracle);
        require(std::fabs(qs.state(i).real() - (i == marked_index ? -1.f : 1.f)) < 1e-6f, "oracle sign");
    }
    const double theta = std::asin(std::pow(2.0, -0.5 * n));
    for (uint32_t k = 1; k <= 4; ++k) {
        const double want = std::pow(std::sin((2 * k + 1) * theta), 2);
        for (bool fused : {false, true}) {
            QSimulator qs(n);
            qs.simulate(grover_search(n, oracle, k, "Oracle", fused));
            require(std::fabs(qs.state_probability(marked_index) - want) < 1e-5, "success probability");
        }
    }
    const uint32_t n10 = 10, idx10 = (1u << 9) | (1u << 7);
    const uint32_t iters = 25;
    QCircuit s10 = grover_search(n10, grover_oracle(n10, marked), iters);
    const double want10 = std::pow(std::sin((2 * iters + 1) * std::asin(std::pow(2.0, -5.0))), 2);
    QSimulator q10(n10);
    q10.simulate(s10);
    require(std::fabs(q10.state_probability(idx10) - want10) < 1e-4, "10-qubit search");
    require(throws<std::invalid_argument>([] { grover_oracle(4, 16); }), "marked state out of range");
}

// ---- dump mode ------------------------------------------------------------

struct Writer {
    std::ostringstream os;
    bool first = true;
    void key(const std::string& k) {
        os << (first ? "" : ",\n") << mini_json::quote(k) << ": ";
        first = false;
    }
    void states(const std::string& k, const std::vector<cfloat>& v) {
        key(k);
        os << "[";
        for (size_t i = 0; i < v.size(); ++i)
            os << (i ? "," : "") << mini_json::num(v[i].real()) << "," << mini_json::num(v[i].imag());
        os << "]";
    }
    void matrix(const std::string& k, const QCircuit& qc) {
        // row-major re, im
        const CircuitMatrix& m = qc.circuit();
        std::vector<cfloat> rows(m.data.size());
        for (uint32_t r = 0; r < m.dim; ++r)
            for (uint32_t c = 0; c < m.dim; ++c) rows[(size_t)r * m.dim + c] = m(r, c);
        states(k, rows);
    }
    void integer(const std::string& k, uint64_t v) {
        key(k);
        os << v;
    }
    void text(const std::string& k, const std::string& v) {
        key(k);
        os << mini_json::quote(v);
    }
};

int dump(const std::string& path) {
    Writer w;
    w.os << "{\n";
    // builders, as circuit matrices of the returned circuits
    auto builder = [&](const std::string& name, QCircuit qc) {
        w.integer(name + ".qubits", qc.qubit_count());
        qc.compile();
        w.matrix(name, qc);
    };
    builder("control_group_4_3_02", Control_Group_Gate(4, 3, {0, 2}, X::gate()));
    builder("control_group_5_2_0134", Control_Group_Gate(5, 2, {0, 1, 3, 4}, Y::gate()));
    builder("control_group_5_4_03", Control_Group_Gate(5, 4, {0, 3}, H::gate()));
    builder("control_group_5_0_24", Control_Group_Gate(5, 0, {2, 4}, X::gate()));
    builder("ncontrol_list_5_0_1", NControl_Gate(5, std::vector<uint32_t>{0}, 1, X::gate()));
    builder("ncontrol_list_5_3_1", NControl_Gate(5, std::vector<uint32_t>{3}, 1, X::gate()));
    builder("ncontrol_list_5_34_1", NControl_Gate(5, std::vector<uint32_t>{3, 4}, 1, Z::gate()));
    builder("ncontrol_list_5_014_2", NControl_Gate(5, std::vector<uint32_t>{0, 1, 4}, 2, X::gate()));
    builder("ncontrol_list_6_0245_3", NControl_Gate(6, std::vector<uint32_t>{0, 2, 4, 5}, 3, X::gate()));
    builder("ncontrol_list_5_02_4", NControl_Gate(5, std::vector<uint32_t>{0, 2}, 4, Y::gate()));
    builder("ncontrol_range_5_0_2_3", NControl_Gate(5, 0, 2, 3, X::gate()));
    builder("ncontrol_range_6_1_3_5", NControl_Gate(6, 1, 3, 5, H::gate()));
    builder("ncontrol_range_4_0_1_2", NControl_Gate(4, 0, 1, 2, Y::gate()));

    // set_basis: every transition from a fixed product state
    const std::vector<QState> init5 = {QState({0.6f, 0.f}, {0.f, 0.8f}), QState({0.3f, -0.1f}, {0.5f, 0.7f}),
                                       QState({1.f, 0.f}, {0.f, 0.f}), QState({0.2f, 0.4f}, {-0.6f, 0.1f}),
                                       QState({0.f, 0.f}, {1.f, 0.f})};
    {
        QSimulator qs(5, init5);
        w.states("set_basis.initial", qs.statevector().host());
    }
    using B = QSimulator::Basis;
    const std::pair<const char*, B> bases[] = {{"Z", B::Z}, {"Y", B::Y}, {"X", B::X}};
    for (const auto& from : bases)
        for (const auto& to : bases) {
            if (from.second == to.second) continue;
            QSimulator qs(5, init5);
            qs.set_basis(from.second);  // Z -> from (identity when from is Z)
            qs.set_basis(to.second);
            w.states(std::string("set_basis.") + from.first + "_to_" + to.first, qs.statevector().host());
            require(qs.get_basis() == to.second, "get_basis after set_basis");
        }

    // linear_entanglement_varstate: 6 qubits, depth 3, values 0.1 * (k + 1)
    {
        const uint32_t n = 6, depth = 3;
        std::vector<float> values(n * depth * 2);
        for (size_t k = 0; k < values.size(); ++k) values[k] = 0.1f * (float)(k + 1) - 1.3f;
        QCircuit qc = linear_entanglement_varstate(n, depth, values);
        w.text("varstate.representation", qc.representation());
        w.integer("varstate.gates", qc.gate_list().size());
        QSimulator qs(n);
        qs.simulate(qc);
        w.states("varstate.state", qs.statevector().host());
        QCircuit compiled = linear_entanglement_varstate(n, depth, values, true);
        QSimulator qc2(n);
        qc2.simulate(compiled);
        w.states("varstate.state_compiled", qc2.statevector().host());
        require(throws<std::invalid_argument>([&] { linear_entanglement_varstate(n, depth, std::vector<float>(5)); },
                                              "Parameter count passed is not the expected received"),
                "varstate parameter count");
    }

    // measure_all on a 4-qubit entangled state: the returned index and the state after
    {
        QCircuit qc(4);
        qc << H{0} << CX{0, 1} << RotY{2, 1.1f} << CRotX{2, 3, 0.8f};
        for (int k = 0; k < 4; ++k) {
            QSimulator qs(4);
            qs.simulate(qc);
            w.states("measure_all." + std::to_string(k) + ".before", qs.statevector().host());
            const uint32_t m = qs.measure_all();
            w.integer("measure_all." + std::to_string(k) + ".index", m);
            w.states("measure_all." + std::to_string(k) + ".after", qs.statevector().host());
        }
    }
    w.os << "\n}\n";
    std::ofstream out(path);
    out << w.os.str();
    std::cout << "dumped " << path << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s run <kat.json> | dump <out.json> [device]\n", argv[0]);
        return 2;
    }
    // device index as the reference's initialize reads it (argv[1])
    char dev0[] = "0";
    char* init_argv[] = {argv[0], argc > 3 ? argv[3] : dev0, nullptr};
    initialize(2, init_argv);
    seed(20240229u);
    const std::string mode = argv[1];
    try {
        if (mode == "dump") return dump(argv[2]);
        const J kat = mini_json::load(argv[2]);
        run_gate_cases(kat["gate_cases"]);
        run_equalities(kat["equalities"]);
        run_ch_cases(kat["ch_cases"]);
        run_prob_cases(kat["prob_cases"]);
        run_measure_cases(kat["measure_cases"]);
        run_qstate_cases(kat["qstate_cases"], kat["qstate_prob_cases"]);
        run_qsim_cases(kat["qsim_cases"]);
        run_collapse_cases(kat["collapse_cases"]);
        run_profile_cases(kat["profile_cases"]);
        run_special_peek(kat["special_peek"]);
        run_equalities(kat["special_equalities"]);
        builder_behaviour();
        representations_and_cache();
        errors();
        compile_equivalence();
        grover();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAILED [%s] unexpected exception: %s\n", g_where.c_str(), e.what());
        return 1;
    }
    std::cout << g_checks << " checks, ALL PASSED" << std::endl;
    return 0;
}
