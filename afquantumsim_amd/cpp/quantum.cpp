// aqs:: C++ API over the MI355X C ABI (include/aqs_hip.h).
//
// Mirrors the reference's src/quantum.cpp class by class.  Where the
// reference multiplies a 2^n x 2^n operator into QCircuit::circuit_, a gate
// here emits one aqs_op into the circuit; the circuit runs it on the state
// being simulated (simulate), records it for fusion (compile) or, when the
// gate is applied to a circuit directly, appends it to the compiled part.
#include "aqs/quantum.h"

#include <cmath>
#include <functional>
#include <mutex>
#include <random>
#include <sstream>
#include <unordered_map>

namespace aqs {

namespace {

// Reference RNG statics (src/quantum.cpp:57-62): mt19937 seeded from
// random_device, float uniforms on [0, 1).
std::random_device g_dv;
std::mt19937 g_rd_gen{g_dv()};
std::uniform_real_distribution<float> g_lin_dist{0.0f, 1.0f};

std::unordered_map<std::string, std::shared_ptr<QCircuit>> g_cached_circuits;
// States up to this size run op sequences in LDS-resident batches (aqs_apply
// on a complex64 state); simulate() and compile() hand them whole runs.
constexpr uint32_t kSmallStateQubits = 12;

void check(aqs_status st) {
    if (st == AQS_OK) return;
    std::string msg = aqs_last_error();
    switch (st) {
        case AQS_ERR_INVALID_ARGUMENT: throw std::invalid_argument{msg};
        case AQS_ERR_OUT_OF_RANGE: throw std::out_of_range{msg};
        case AQS_ERR_DOMAIN: throw std::domain_error{msg};
        default: throw std::runtime_error{msg};
    }
}

const float kHf = 0.70710678118f;  // src/quantum.cpp:46-49

aqs_op make_op(int kind, std::initializer_list<int> targets, std::initializer_list<int> controls = {}) {
    aqs_op op{};
    op.kind = kind;
    op.ntargets = 0;
    for (int t : targets) op.targets[op.ntargets++] = t;
    op.ncontrols = 0;
    for (int c : controls) op.controls[op.ncontrols++] = c;
    return op;
}

void set_m(aqs_op& op, std::complex<double> a, std::complex<double> b, std::complex<double> c,
           std::complex<double> d) {
    const std::complex<double> v[4] = {a, b, c, d};
    for (int i = 0; i < 4; ++i) {
        op.m[2 * i] = v[i].real();
        op.m[2 * i + 1] = v[i].imag();
    }
}

// Matrices as the reference builds them, in float (src/quantum.cpp:683-688,
// 730-735, 777-782, 861-866).
aqs_op op_rotx(int t, float angle) {
    const float c = std::cos(angle / 2.0f), s = std::sin(angle / 2.0f);
    aqs_op op = make_op(AQS_OP_MAT1, {t});
    set_m(op, c, {0.0, -s}, {0.0, -s}, c);
    return op;
}
aqs_op op_roty(int t, float angle) {
    const float c = std::cos(angle / 2.0f), s = std::sin(angle / 2.0f);
    aqs_op op = make_op(AQS_OP_MAT1, {t});
    set_m(op, c, -s, s, c);
    return op;
}
aqs_op op_rotz(int t, float angle) {
    const float c = std::cos(angle / 2.0f), s = std::sin(angle / 2.0f);
    aqs_op op = make_op(AQS_OP_DIAG1, {t});
    set_m(op, {c, -s}, 0.0, 0.0, {c, s});
    return op;
}
aqs_op op_h(int t) {
    aqs_op op = make_op(AQS_OP_MAT1, {t});
    set_m(op, kHf, kHf, kHf, -kHf);
    return op;
}
aqs_op op_y(int t) {
    aqs_op op = make_op(AQS_OP_MAT1, {t});
    set_m(op, 0.0, {0.0, -1.0}, {0.0, 1.0}, 0.0);
    return op;
}
aqs_op op_phase(int t, std::complex<double> ph) {
    aqs_op op = make_op(AQS_OP_PHASE, {t});
    set_m(op, ph, 0.0, 0.0, 0.0);
    return op;
}
std::complex<double> eiphi(float angle) { return {std::cos(angle), std::sin(angle)}; }

void add_control(aqs_op& op, int c) { op.controls[op.ncontrols++] = c; }

}  // namespace

namespace detail {

struct OpRecord {
    aqs_op op{};
    std::shared_ptr<std::vector<double>> mat;
};

OpRecord own(const aqs_op& op) {
    OpRecord r;
    r.op = op;
    if (op.kind == AQS_OP_MATK && op.u) {
        const size_t K = (size_t)1 << op.ntargets;
        r.mat = std::make_shared<std::vector<double>>(op.u, op.u + 2 * K * K);
        r.op.u = r.mat->data();
    }
    return r;
}

struct Segment {
    aqs_schedule sch = nullptr;
    std::vector<OpRecord> ops;
    Segment() = default;
    Segment(const Segment&) = delete;
    Segment& operator=(const Segment&) = delete;
    ~Segment() {
        if (sch) aqs_schedule_destroy(sch);
    }
    void run(aqs_state s) const {
        if (sch) {
            check(aqs_apply_schedule(s, sch));
            return;
        }
        std::vector<aqs_op> v;
        v.reserve(ops.size());
        for (const auto& r : ops) v.push_back(r.op);
        check(aqs_apply(s, v.data(), v.size()));
    }
};

struct Sink {
    aqs_state state = nullptr;
    std::vector<OpRecord>* record = nullptr;
    std::function<void(const aqs_op&)> forward;
};

}  // namespace detail

// ---------------------------------------------------------------------------

void initialize(int argc, char** argv) {
    const int device = (argc > 1) ? std::stoi(argv[1]) : 0;
    check(aqs_init(device));
}

void clear_circuit_cache() { g_cached_circuits.clear(); }

void seed(uint32_t value) { g_rd_gen.seed(value); }

// ---- QState (src/quantum.cpp:90-157) ---------------------------------------

QState::QState(const std::complex<float>& zeroState, const std::complex<float>& oneState)
    : state_{zeroState, oneState} {
    force_normalize();
}

QState::QState(const std::array<std::complex<float>, 2>& states) : state_{states[0], states[1]} { force_normalize(); }

QState& QState::set(const std::complex<float>& zero_state, const std::complex<float>& one_state) {
    state_[0] = zero_state;
    state_[1] = one_state;
    force_normalize();
    return *this;
}

bool QState::peek_measure() const {
    const float prob = probability_true();
    const float val = g_lin_dist(g_rd_gen);
    return val < prob;
}

bool QState::measure() {
    const bool m = peek_measure();
    state_[m] = 1.f;
    state_[!m] = 0.f;
    return m;
}

std::array<uint32_t, 2> QState::profile_measure(uint32_t rep_count) const {
    uint32_t t = 0;
    for (uint32_t i = 0; i < rep_count; ++i) t += static_cast<uint32_t>(peek_measure());
    return {rep_count - t, t};
}

void QState::force_normalize() {
    const float mag2 = state_[0].real() * state_[0].real() + state_[0].imag() * state_[0].imag() +
                       state_[1].real() * state_[1].real() + state_[1].imag() * state_[1].imag();
    if (mag2 == 0.f) throw std::invalid_argument{"Cannot normalize a null state"};
    const float mag = std::sqrt(mag2);
    state_[0] = state_[0] / mag;
    state_[1] = state_[1] / mag;
}

// ---- StateVector ---------------------------------------------------------------

StateVector::StateVector(uint32_t qubits) : n_{qubits} { check(aqs_state_create((int)qubits, AQS_C64, &h_)); }

StateVector::StateVector(const StateVector& o) : n_{o.n_} {
    if (o.h_) {
        check(aqs_state_create((int)n_, AQS_C64, &h_));
        check(aqs_state_copy(h_, o.h_));
    }
}

StateVector& StateVector::operator=(const StateVector& o) {
    if (this == &o) return *this;
    StateVector tmp(o);
    *this = std::move(tmp);
    return *this;
}

StateVector::StateVector(StateVector&& o) noexcept : h_{o.h_}, n_{o.n_} { o.h_ = nullptr; }

StateVector& StateVector::operator=(StateVector&& o) noexcept {
    if (this != &o) {
        if (h_) aqs_state_destroy(h_);
        h_ = o.h_;
        n_ = o.n_;
        o.h_ = nullptr;
    }
    return *this;
}

StateVector::~StateVector() {
    if (h_) aqs_state_destroy(h_);
}

std::vector<cfloat> StateVector::host() const {
    std::vector<cfloat> out(size());
    host(out.data());
    return out;
}

void StateVector::host(cfloat* out) const { check(aqs_state_download(h_, out)); }

void StateVector::upload(const cfloat* in, bool normalize) { check(aqs_state_upload(h_, in, normalize ? 1 : 0)); }

// ---- QCircuit (src/quantum.cpp:159-210) -------------------------------------

QCircuit::QCircuit(uint32_t qubit_count) : qubits_{qubit_count} {
    if (qubit_count < 1) throw std::invalid_argument{"Circuit must contain at least 1 qubit"};
    if (qubit_count > max_qubit_count)
        throw std::invalid_argument{"Maximum qubit count supported is " + std::to_string(max_qubit_count)};
}

QCircuit::QCircuit(const QCircuit& o)
    : gate_list_{o.gate_list_},
      representation_{o.representation_},
      qubits_{o.qubits_},
      cached_index_{o.cached_index_},
      fusion_width_{o.fusion_width_},
      compiled_{o.compiled_},
      matrix_{o.matrix_},
      unitary64_{o.unitary64_} {}

QCircuit::QCircuit(QCircuit&& o) noexcept
    : gate_list_{std::move(o.gate_list_)},
      representation_{std::move(o.representation_)},
      qubits_{o.qubits_},
      cached_index_{o.cached_index_},
      fusion_width_{o.fusion_width_},
      compiled_{std::move(o.compiled_)},
      matrix_{std::move(o.matrix_)},
      unitary64_{std::move(o.unitary64_)} {}

QCircuit& QCircuit::operator=(const QCircuit& o) {
    if (this != &o) {
        gate_list_ = o.gate_list_;
        representation_ = o.representation_;
        qubits_ = o.qubits_;
        cached_index_ = o.cached_index_;
        fusion_width_ = o.fusion_width_;
        compiled_ = o.compiled_;
        matrix_ = o.matrix_;
        unitary64_ = o.unitary64_;
        sink_ = nullptr;
    }
    return *this;
}

QCircuit& QCircuit::operator=(QCircuit&& o) noexcept {
    if (this != &o) {
        gate_list_ = std::move(o.gate_list_);
        representation_ = std::move(o.representation_);
        qubits_ = o.qubits_;
        cached_index_ = o.cached_index_;
        fusion_width_ = o.fusion_width_;
        compiled_ = std::move(o.compiled_);
        matrix_ = std::move(o.matrix_);
        unitary64_ = std::move(o.unitary64_);
        sink_ = nullptr;
    }
    return *this;
}

QCircuit::~QCircuit() = default;

void QCircuit::clear() {
    clear_cache();
    gate_list_.clear();
}

void QCircuit::clear_cache() {
    cached_index_ = 0;
    compiled_.clear();
    matrix_.reset();
    unitary64_.reset();
}

bool operator==(const QCircuit& lhs, const QCircuit& rhs) {
    if (lhs.qubit_count() != rhs.qubit_count()) return false;
    if (lhs.gate_list().size() != rhs.gate_list().size()) return false;
    if (lhs.representation() != rhs.representation()) return false;
    for (uint32_t i = 0; i < lhs.gate_list().size(); ++i)
        if (lhs.gate_list()[i] != rhs.gate_list()[i]) return false;
    return true;
}

void QCircuit::emit(const aqs_op& op) {
    if (sink_) {
        if (sink_->state) {
            check(aqs_apply(sink_->state, &op, 1));
        } else if (sink_->record) {
            sink_->record->push_back(detail::own(op));
        } else if (sink_->forward) {
            sink_->forward(op);
        }
        return;
    }
    // Applied to the circuit itself: becomes part of its compiled operator.
    if (compiled_.empty() || compiled_.back()->sch) compiled_.push_back(std::make_shared<detail::Segment>());
    compiled_.back()->ops.push_back(detail::own(op));
    matrix_.reset();
    unitary64_.reset();
}

void QCircuit::compile() {
    if (cached_index_ == gate_list_.size()) return;
    std::vector<detail::OpRecord> rec;
    detail::Sink sink;
    sink.record = &rec;
    detail::Sink* saved = sink_;
    sink_ = &sink;
    try {
        for (size_t i = cached_index_; i < gate_list_.size(); ++i) (*gate_list_[i])(*this);
    } catch (...) {
        sink_ = saved;
        throw;
    }
    sink_ = saved;
    std::vector<aqs_op> ops;
    ops.reserve(rec.size());
    for (auto& r : rec) ops.push_back(r.op);
    auto seg = std::make_shared<detail::Segment>();
    if (!ops.empty() && fusion_width_ == 0 && qubits_ <= kSmallStateQubits) {
        // A small state runs its ops in LDS-resident batches (aqs_apply),
        // faster than fused passes: QFT-10 0.054 ms against 0.142 ms.
        seg->ops = std::move(rec);
        compiled_.push_back(seg);
    } else if (!ops.empty()) {
        if (fusion_width_ == 0)
            check(aqs_fuse_auto(ops.data(), ops.size(), (int)qubits_, &seg->sch));
        else
            check(aqs_fuse(ops.data(), ops.size(), (int)qubits_, fusion_width_, &seg->sch));
    }
    if (seg->sch) compiled_.push_back(seg);
    cached_index_ = gate_list_.size();
    matrix_.reset();
    unitary64_.reset();
}

size_t QCircuit::compiled_passes() const {
    size_t p = 0;
    for (const auto& s : compiled_) {
        if (s->sch) {
            size_t passes = 0;
            check(aqs_schedule_info(s->sch, &passes, nullptr));
            p += passes;
        } else if (qubits_ <= kSmallStateQubits) {
            // launches of aqs_apply's small-state path: one per MATK op and
            // per single op, one per 64 ops of a longer run
            size_t run = 0;
            for (const auto& r : s->ops) {
                if (r.op.kind != AQS_OP_MATK) {
                    ++run;
                    continue;
                }
                p += (run >= 2 ? (run + 63) / 64 : run) + 1;
                run = 0;
            }
            p += run >= 2 ? (run + 63) / 64 : run;
        } else {
            p += s->ops.size();
        }
    }
    return p;
}

void QCircuit::run_compiled(aqs_state s) const {
    for (const auto& seg : compiled_) seg->run(s);
}

const CircuitMatrix& QCircuit::circuit() const {
    if (matrix_) return *matrix_;
    if (qubits_ > 12)
        throw std::invalid_argument{"circuit() materialises the unitary only for circuits of up to 12 qubits"};
    const uint32_t N = state_count();
    auto m = std::make_shared<CircuitMatrix>();
    m->dim = N;
    m->data.resize((size_t)N * N);
    StateVector col(qubits_);
    for (uint32_t j = 0; j < N; ++j) {
        check(aqs_state_set_basis(col.handle(), j));
        run_compiled(col.handle());
        col.host(m->data.data() + (size_t)j * N);
    }
    matrix_ = m;
    return *matrix_;
}

// ---- QSimulator (src/quantum.cpp:212-531) -----------------------------------

QSimulator::QSimulator(uint32_t qubit_count, const QState& initial_state, const QNoise& noise_generator)
    : states_(qubit_count, initial_state), noise_{noise_generator}, qubits_{qubit_count}, basis_{Basis::Z} {
    if (qubit_count < 1) throw std::invalid_argument{"Circuit must contain at least 1 qubit"};
    statevector_ = StateVector(qubit_count);
    if (initial_state == QState::zero())
        check(aqs_state_set_basis(statevector_.handle(), 0));
    else if (initial_state == QState::one())
        check(aqs_state_set_basis(statevector_.handle(), state_count() - 1));
    else
        generate_statevector();
}

QSimulator::QSimulator(uint32_t qubit_count, std::vector<QState> initial_states, const QNoise& noise_generator)
    : states_(std::move(initial_states)), noise_{noise_generator}, qubits_{qubit_count}, basis_{Basis::Z} {
    if (qubit_count != states_.size())
        throw std::invalid_argument{
            "The number of initial states must match the number of qubits in the circuit"};
    statevector_ = StateVector(qubit_count);
    generate_statevector();
}

QSimulator::QSimulator(uint32_t qubit_count, const std::vector<cfloat>& statevector, const QNoise& noise_generator)
    : states_(qubit_count), noise_{noise_generator}, qubits_{qubit_count}, basis_{Basis::Z} {
    if (statevector.size() != fast_pow2(qubit_count)) throw std::invalid_argument{"Invalid initial statevector shape"};
    statevector_ = StateVector(qubit_count);
    aqs_status st = aqs_state_upload(statevector_.handle(), statevector.data(), 1);
    if (st == AQS_ERR_INVALID_ARGUMENT) throw std::invalid_argument{"Cannot have a null statevector"};
    check(st);
}

void QSimulator::generate_statevector() {
    std::vector<double> qa(4 * (size_t)qubit_count());
    for (uint32_t i = 0; i < qubit_count(); ++i) {
        qa[4 * i + 0] = states_[i][0].real();
        qa[4 * i + 1] = states_[i][0].imag();
        qa[4 * i + 2] = states_[i][1].real();
        qa[4 * i + 3] = states_[i][1].imag();
    }
    check(aqs_state_product(statevector_.handle(), qa.data()));
}

void QSimulator::simulate(const QCircuit& circuit) {
    if (circuit.qubit_count() != qubits_)
        throw std::invalid_argument{"Number of qubit states and circuit input qubit states do not match"};
    circuit.run_compiled(statevector_.handle());
    QCircuit temp{qubit_count()};
    detail::Sink sink;
    if (qubit_count() > kSmallStateQubits) {
        // one aqs_apply per op, as the gates emit them
        sink.state = statevector_.handle();
        temp.sink_ = &sink;
        for (size_t i = circuit.cached_index_; i < circuit.gate_list().size(); ++i)
            circuit.gate_list()[i]->operator()(temp);
        temp.sink_ = nullptr;
        return;
    }
    // Small states: the ops are collected and applied by one aqs_apply, which
    // runs them in LDS-resident batches instead of one launch each.
    std::vector<detail::OpRecord> rec;
    sink.record = &rec;
    temp.sink_ = &sink;
    auto flush = [&] {
        std::vector<aqs_op> v;
        v.reserve(rec.size());
        for (const auto& r : rec) v.push_back(r.op);
        check(aqs_apply(statevector_.handle(), v.data(), v.size()));
    };
    try {
        for (size_t i = circuit.cached_index_; i < circuit.gate_list().size(); ++i)
            circuit.gate_list()[i]->operator()(temp);
    } catch (...) {
        // the ops emitted before the failing gate still apply
        temp.sink_ = nullptr;
        flush();
        throw;
    }
    temp.sink_ = nullptr;
    flush();
}

bool QSimulator::peek_measure(uint32_t qubit) const {
    if (qubit >= qubit_count()) throw std::out_of_range{"Cannot measure the state of the given qubit"};
    const float prob = qubit_probability_true(qubit);
    const float val = g_lin_dist(g_rd_gen);
    return val < prob;
}

bool QSimulator::measure(uint32_t qubit) {
    if (qubit >= qubit_count()) throw std::out_of_range{"Cannot measure the state of the given qubit"};
    const float val = g_lin_dist(g_rd_gen);
    int outcome = 0;
    check(aqs_measure_qubit(statevector_.handle(), (int)qubit, (double)val, &outcome, nullptr));
    return outcome != 0;
}

uint32_t QSimulator::peek_measure_all() const {
    const double u = g_lin_dist(g_rd_gen);
    uint64_t idx = 0;
    check(aqs_sample(statevector_.handle(), &u, 1, &idx));
    return (uint32_t)idx;
}

uint32_t QSimulator::measure_all() {
    const uint32_t m = peek_measure_all();
    check(aqs_collapse_to(statevector_.handle(), m));
    return m;
}

std::vector<uint64_t> QSimulator::sample(const std::vector<double>& uniforms) const {
    std::vector<uint64_t> out(uniforms.size());
    if (!uniforms.empty()) check(aqs_sample(statevector_.handle(), uniforms.data(), (uint32_t)uniforms.size(), out.data()));
    return out;
}

std::vector<uint32_t> QSimulator::profile_measure_all(uint32_t rep_count) const {
    std::vector<uint32_t> count(state_count());
    std::vector<double> u(rep_count);
    for (auto& x : u) x = g_lin_dist(g_rd_gen);
    for (uint64_t i : sample(u)) count[i] += 1;
    return count;
}

std::array<uint32_t, 2> QSimulator::profile_measure(uint32_t qubit, uint32_t rep_count) const {
    if (qubit >= qubit_count()) throw std::out_of_range{"Cannot profile measurement of the given qubit"};
    double p1 = 0.0;
    check(aqs_prob1(statevector_.handle(), (int)qubit, &p1));
    uint32_t t = 0;
    for (uint32_t i = 0; i < rep_count; ++i) t += (double)g_lin_dist(g_rd_gen) < p1 ? 1u : 0u;
    return {rep_count - t, t};
}

cfloat QSimulator::state(uint32_t st) const {
    assert(st < state_count());
    double v[2];
    check(aqs_amplitude(statevector_.handle(), st, v));
    return {(float)v[0], (float)v[1]};
}

float QSimulator::qubit_probability_true(uint32_t qubit) const {
    if (qubit >= qubit_count()) throw std::out_of_range{"Cannot obtain probability of the given qubit"};
    double p = 0.0;
    check(aqs_prob1(statevector_.handle(), (int)qubit, &p));
    return (float)p;
}

float QSimulator::state_probability(uint32_t st) const {
    if (st >= state_count()) throw std::out_of_range{"Cannot obtain probability of the given state"};
    const cfloat v = state(st);
    return v.real() * v.real() + v.imag() * v.imag();
}

std::vector<float> QSimulator::probabilities() const {
    std::vector<float> out(state_count());
    check(aqs_probabilities(statevector_.handle(), out.data()));
    return out;
}

// set_basis (src/quantum.cpp:416-465).  The reference's switch on the current
// basis has no break after the Y case, so leaving the Y basis applies the
// X-to-Z matrix; this is kept so results match the reference.
void QSimulator::set_basis(Basis basis) {
    if (basis == basis_) return;
    using cd = std::complex<double>;
    const double r = kHf;
    const cd z_to_x[4] = {r, r, r, -r};
    const cd z_to_y[4] = {r, r, cd(0, -r), cd(0, r)};
    const cd x_to_z[4] = {r, r, r, -r};
    cd m[4] = {1, 0, 0, 1};
    switch (basis_) {
        case Basis::Z: break;
        case Basis::Y:  // falls through as in the reference
        case Basis::X: std::copy(x_to_z, x_to_z + 4, m); break;
    }
    auto mul = [](const cd* a, const cd* b, cd* o) {
        cd t[4] = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                   a[2] * b[1] + a[3] * b[3]};
        std::copy(t, t + 4, o);
    };
    switch (basis) {
        case Basis::Z: break;
        case Basis::Y: mul(z_to_y, m, m); break;
        case Basis::X: mul(z_to_x, m, m); break;
    }
    std::vector<aqs_op> ops;
    for (uint32_t q = 0; q < qubits_; ++q) {
        aqs_op op = make_op(AQS_OP_MAT1, {(int)q});
        set_m(op, m[0], m[1], m[2], m[3]);
        ops.push_back(op);
    }
    check(aqs_apply(statevector_.handle(), ops.data(), ops.size()));
    basis_ = basis;
}

// ---- gates -----------------------------------------------------------------

namespace {

void check_target(const QCircuit& qc, uint32_t q) {
    if (q >= qc.qubit_count()) throw std::out_of_range{"Cannot add gate at the given qubit position"};
}

void check_ctrl_target(const QCircuit& qc, uint32_t c, uint32_t t) {
    if (qc.qubit_count() < 2) throw std::domain_error{"Gate not supported for given simulation"};
    check_target(qc, c);
    check_target(qc, t);
    if (c == t) throw std::invalid_argument{"Control qubit cannot be the same as the target qubit"};
}

template <typename G>
const G* as(const QGate& rhs) {
    return static_cast<const G*>(&rhs);
}

std::string name_1q(const char* name, uint32_t t) { return std::string(name) + ",0,1:" + std::to_string(t) + ";"; }
std::string name_c1(const char* name, uint32_t c, uint32_t t) {
    return std::string(name) + ",1,1:" + std::to_string(c) + "," + std::to_string(t) + ";";
}

// Phase naming (src/quantum.cpp:885-896, 1227-1244).
const char* phase_name(float angle) {
    if (angle == pi / 2) return "S";
    if (angle == -pi / 2) return "S†";
    if (angle == pi / 4) return "T";
    if (angle == -pi / 4) return "T†";
    return "Phase";
}

}  // namespace

// X ---------------------------------------------------------------------------
bool X::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& X::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(make_op(AQS_OP_X, {(int)target_qubit}));
    return qc;
}
std::string X::to_string() const { return name_1q("X", target_qubit); }
bool X::operator==(const QGate& rhs) const noexcept { return type() == rhs.type() && target_qubit == as<X>(rhs)->target_qubit; }
const QCircuit& X::gate() {
    static QCircuit qc = [] { QCircuit c(1); c << X{0}; c.compile(); return c; }();
    return qc;
}

// Y ---------------------------------------------------------------------------
bool Y::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& Y::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_y((int)target_qubit));
    return qc;
}
std::string Y::to_string() const { return name_1q("Y", target_qubit); }
bool Y::operator==(const QGate& rhs) const noexcept { return type() == rhs.type() && target_qubit == as<Y>(rhs)->target_qubit; }
const QCircuit& Y::gate() {
    static QCircuit qc = [] { QCircuit c(1); c << Y{0}; c.compile(); return c; }();
    return qc;
}

// Z ---------------------------------------------------------------------------
bool Z::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& Z::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_phase((int)target_qubit, -1.0));
    return qc;
}
std::string Z::to_string() const { return name_1q("Z", target_qubit); }
bool Z::operator==(const QGate& rhs) const noexcept { return type() == rhs.type() && target_qubit == as<Z>(rhs)->target_qubit; }
const QCircuit& Z::gate() {
    static QCircuit qc = [] { QCircuit c(1); c << Z{0}; c.compile(); return c; }();
    return qc;
}

// RotX / RotY / RotZ ------------------------------------------------------------
bool RotX::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& RotX::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_rotx((int)target_qubit, angle));
    return qc;
}
std::string RotX::to_string() const { return name_1q("RotX", target_qubit); }
bool RotX::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<RotX>(rhs)->target_qubit && angle == as<RotX>(rhs)->angle;
}
QCircuit RotX::gate(float angle) {
    QCircuit c(1);
    c << RotX{0, angle};
    c.compile();
    return c;
}

bool RotY::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& RotY::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_roty((int)target_qubit, angle));
    return qc;
}
std::string RotY::to_string() const { return name_1q("RotY", target_qubit); }
bool RotY::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<RotY>(rhs)->target_qubit && angle == as<RotY>(rhs)->angle;
}
QCircuit RotY::gate(float angle) {
    QCircuit c(1);
    c << RotY{0, angle};
    c.compile();
    return c;
}

bool RotZ::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& RotZ::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_rotz((int)target_qubit, angle));
    return qc;
}
std::string RotZ::to_string() const { return name_1q("RotZ", target_qubit); }
bool RotZ::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<RotZ>(rhs)->target_qubit && angle == as<RotZ>(rhs)->angle;
}
QCircuit RotZ::gate(float angle) {
    QCircuit c(1);
    c << RotZ{0, angle};
    c.compile();
    return c;
}

// H ---------------------------------------------------------------------------
bool H::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& H::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_h((int)target_qubit));
    return qc;
}
std::string H::to_string() const { return name_1q("H", target_qubit); }
bool H::operator==(const QGate& rhs) const noexcept { return type() == rhs.type() && target_qubit == as<H>(rhs)->target_qubit; }
const QCircuit& H::gate() {
    static QCircuit qc = [] { QCircuit c(1); c << H{0}; c.compile(); return c; }();
    return qc;
}

// Phase -------------------------------------------------------------------------
bool Phase::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& Phase::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(op_phase((int)target_qubit, eiphi(angle)));
    return qc;
}
std::string Phase::to_string() const { return name_1q(phase_name(angle), target_qubit); }
bool Phase::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<Phase>(rhs)->target_qubit && angle == as<Phase>(rhs)->angle;
}
QCircuit Phase::gate(float angle) {
    QCircuit c(1);
    c << Phase{0, angle};
    c.compile();
    return c;
}

// U3 ----------------------------------------------------------------------------
bool U3::check(const QCircuit& qc) const {
    check_target(qc, target_qubit);
    return true;
}
QCircuit& U3::operator()(QCircuit& qc) const {
    check(qc);
    const double t = theta, p = phi, l = lambda;
    const double c = std::cos(t / 2), s = std::sin(t / 2);
    aqs_op op = make_op(AQS_OP_MAT1, {(int)target_qubit});
    set_m(op, c, -std::polar(1.0, l) * s, std::polar(1.0, p) * s, std::polar(1.0, p + l) * c);
    qc.emit(op);
    return qc;
}
std::string U3::to_string() const { return name_1q("U3", target_qubit); }
bool U3::operator==(const QGate& rhs) const noexcept {
    if (type() != rhs.type()) return false;
    const U3* o = as<U3>(rhs);
    return target_qubit == o->target_qubit && theta == o->theta && phi == o->phi && lambda == o->lambda;
}

// Swap --------------------------------------------------------------------------
bool Swap::check(const QCircuit& qc) const {
    if (qc.qubit_count() < 2) throw std::domain_error{"Gate not supported for given simulation"};
    check_target(qc, target_qubit_A);
    check_target(qc, target_qubit_B);
    if (target_qubit_A == target_qubit_B)
        throw std::invalid_argument{"Cannot use the swap gate on the same target qubits"};
    return true;
}
QCircuit& Swap::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(make_op(AQS_OP_SWAP, {(int)target_qubit_A, (int)target_qubit_B}));
    return qc;
}
std::string Swap::to_string() const {
    return "Swap,0,2:" + std::to_string(target_qubit_A) + "," + std::to_string(target_qubit_B) + ";";
}
bool Swap::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit_A == as<Swap>(rhs)->target_qubit_A &&
           target_qubit_B == as<Swap>(rhs)->target_qubit_B;
}

// CX / CY / CZ / CH / CPhase ------------------------------------------------------
bool CX::check(const QCircuit& qc) const {
    check_ctrl_target(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CX::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(make_op(AQS_OP_X, {(int)target_qubit}, {(int)control_qubit}));
    return qc;
}
std::string CX::to_string() const { return name_c1("X", control_qubit, target_qubit); }
bool CX::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CX>(rhs)->target_qubit && control_qubit == as<CX>(rhs)->control_qubit;
}

bool CY::check(const QCircuit& qc) const {
    check_ctrl_target(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CY::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_y((int)target_qubit);
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CY::to_string() const { return name_c1("Y", control_qubit, target_qubit); }
bool CY::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CY>(rhs)->target_qubit && control_qubit == as<CY>(rhs)->control_qubit;
}

bool CZ::check(const QCircuit& qc) const {
    check_ctrl_target(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CZ::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_phase((int)target_qubit, -1.0);
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CZ::to_string() const { return name_c1("Z", control_qubit, target_qubit); }
bool CZ::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CZ>(rhs)->target_qubit && control_qubit == as<CZ>(rhs)->control_qubit;
}

bool CH::check(const QCircuit& qc) const {
    check_ctrl_target(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CH::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_h((int)target_qubit);
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CH::to_string() const { return name_c1("H", control_qubit, target_qubit); }
bool CH::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CH>(rhs)->target_qubit && control_qubit == as<CH>(rhs)->control_qubit;
}

bool CPhase::check(const QCircuit& qc) const {
    check_ctrl_target(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CPhase::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_phase((int)target_qubit, eiphi(angle));
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CPhase::to_string() const { return name_c1(phase_name(angle), control_qubit, target_qubit); }
bool CPhase::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CPhase>(rhs)->target_qubit &&
           control_qubit == as<CPhase>(rhs)->control_qubit && angle == as<CPhase>(rhs)->angle;
}

// CSwap -------------------------------------------------------------------------
bool CSwap::check(const QCircuit& qc) const {
    if (qc.qubit_count() < 3) throw std::domain_error{"Gate not supported for given circuit"};
    check_target(qc, target_qubit_A);
    check_target(qc, target_qubit_B);
    check_target(qc, control_qubit);
    if (control_qubit == target_qubit_A || control_qubit == target_qubit_B)
        throw std::invalid_argument{"Control qubit cannot be the same as the target qubit"};
    if (target_qubit_A == target_qubit_B) throw std::invalid_argument{"Cannot use the swap gate on the same target qubits"};
    return true;
}
QCircuit& CSwap::operator()(QCircuit& qc) const {
    check(qc);
    qc.emit(make_op(AQS_OP_SWAP, {(int)target_qubit_A, (int)target_qubit_B}, {(int)control_qubit}));
    return qc;
}
// The reference writes the two targets without a separator (src/quantum.cpp:1331-1335).
std::string CSwap::to_string() const {
    return "Swap,1,2:" + std::to_string(control_qubit) + "," + std::to_string(target_qubit_A) +
           std::to_string(target_qubit_B) + ";";
}
bool CSwap::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit_A == as<CSwap>(rhs)->target_qubit_A &&
           target_qubit_B == as<CSwap>(rhs)->target_qubit_B && control_qubit == as<CSwap>(rhs)->control_qubit;
}

// CRotX / CRotY / CRotZ: controlled rotations (src/quantum.cpp:1405-1510) --------
namespace {
void check_crot(const QCircuit& qc, uint32_t c, uint32_t t) {
    if (t >= qc.qubit_count() || c >= qc.qubit_count())
        throw std::out_of_range{"Cannot add gate at the given qubit position"};
    if (t == c) throw std::invalid_argument{"Control qubit cannot be the same as the target qubit"};
}
}  // namespace

bool CRotX::check(const QCircuit& qc) const {
    check_crot(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CRotX::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_rotx((int)target_qubit, angle);
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CRotX::to_string() const { return name_c1("RotX", control_qubit, target_qubit); }
bool CRotX::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CRotX>(rhs)->target_qubit &&
           control_qubit == as<CRotX>(rhs)->control_qubit && angle == as<CRotX>(rhs)->angle;
}

bool CRotY::check(const QCircuit& qc) const {
    check_crot(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CRotY::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_roty((int)target_qubit, angle);
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CRotY::to_string() const { return name_c1("RotY", control_qubit, target_qubit); }
bool CRotY::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CRotY>(rhs)->target_qubit &&
           control_qubit == as<CRotY>(rhs)->control_qubit && angle == as<CRotY>(rhs)->angle;
}

bool CRotZ::check(const QCircuit& qc) const {
    check_crot(qc, control_qubit, target_qubit);
    return true;
}
QCircuit& CRotZ::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = op_rotz((int)target_qubit, angle);
    add_control(op, (int)control_qubit);
    qc.emit(op);
    return qc;
}
std::string CRotZ::to_string() const { return name_c1("RotZ", control_qubit, target_qubit); }
bool CRotZ::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CRotZ>(rhs)->target_qubit &&
           control_qubit == as<CRotZ>(rhs)->control_qubit && angle == as<CRotZ>(rhs)->angle;
}

// CCNot / Or ------------------------------------------------------------------------
namespace {
void check_two_ctrl(const QCircuit& qc, uint32_t a, uint32_t b, uint32_t t) {
    if (qc.qubit_count() < 3) throw std::domain_error{"Gate not supported for given simulation"};
    check_target(qc, a);
    check_target(qc, b);
    check_target(qc, t);
    if (a == t || b == t) throw std::invalid_argument{"Control qubit cannot be the same as the target qubit"};
}
}  // namespace

bool CCNot::check(const QCircuit& qc) const {
    check_two_ctrl(qc, control_qubit_A, control_qubit_B, target_qubit);
    return true;
}
QCircuit& CCNot::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = make_op(AQS_OP_X, {(int)target_qubit}, {(int)control_qubit_A});
    // A == B is accepted by the reference and acts as a single control.
    if (control_qubit_B != control_qubit_A) add_control(op, (int)control_qubit_B);
    qc.emit(op);
    return qc;
}
std::string CCNot::to_string() const {
    return "X,2,1:" + std::to_string(control_qubit_A) + "," + std::to_string(control_qubit_B) + "," +
           std::to_string(target_qubit) + ";";
}
bool CCNot::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<CCNot>(rhs)->target_qubit &&
           control_qubit_A == as<CCNot>(rhs)->control_qubit_A && control_qubit_B == as<CCNot>(rhs)->control_qubit_B;
}

bool Or::check(const QCircuit& qc) const {
    check_two_ctrl(qc, control_qubit_A, control_qubit_B, target_qubit);
    return true;
}
QCircuit& Or::operator()(QCircuit& qc) const {
    check(qc);
    aqs_op op = make_op(AQS_OP_X, {(int)target_qubit}, {(int)control_qubit_A});
    if (control_qubit_B != control_qubit_A) add_control(op, (int)control_qubit_B);
    op.flags = AQS_CTRL_ANY;
    qc.emit(op);
    return qc;
}
// Equivalent-circuit representation (src/quantum.cpp:1662-1677).
std::string Or::to_string() const {
    std::stringstream b;
    b << "P;" << "X,0,1:" << control_qubit_A << ";" << "X,0,1:" << control_qubit_B << ";" << "X,0,1:" << target_qubit
      << ";" << "X,2,1:" << control_qubit_A << "," << control_qubit_B << "," << target_qubit << ";"
      << "X,0,1:" << control_qubit_A << ";" << "X,0,1:" << control_qubit_B << ";" << "P;";
    return b.str();
}
bool Or::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit == as<Or>(rhs)->target_qubit &&
           control_qubit_A == as<Or>(rhs)->control_qubit_A && control_qubit_B == as<Or>(rhs)->control_qubit_B;
}

// Gate / ControlGate (src/quantum.cpp:1679-1960) ------------------------------------
namespace {

// Representation rewriting: every "name,c,t:q1,...;" entry gets its qubits
// shifted (Gate) or an extra leading control (ControlGate); entries without
// qubit lists (barriers) are copied.
std::string shift_representation(const std::string& rep, uint32_t offset, bool add_control, uint32_t control) {
    std::stringstream out;
    size_t begin = 0;
    while (true) {
        size_t end = rep.find(';', begin);
        if (end == std::string::npos) break;
        std::string entry = rep.substr(begin, end - begin);
        begin = end + 1;
        size_t colon = entry.find(':');
        size_t comma = entry.find(',');
        if (comma == std::string::npos) {
            out << entry << ";";
            continue;
        }
        std::string name = entry.substr(0, comma);
        std::string counts = colon == std::string::npos ? "" : entry.substr(comma + 1, colon - comma - 1);
        std::string qubits = colon == std::string::npos ? entry.substr(comma + 1) : entry.substr(colon + 1);
        out << name << ",";
        if (!counts.empty()) {
            if (add_control) {
                size_t c2 = counts.find(',');
                out << std::stoi(counts.substr(0, c2)) + 1 << counts.substr(c2) << ":";
            } else {
                out << counts << ":";
            }
        }
        if (add_control) out << control << ",";
        std::stringstream qs(qubits);
        std::string q;
        bool first = true;
        while (std::getline(qs, q, ',')) {
            if (!first) out << ",";
            out << std::stoi(q) + offset;
            first = false;
        }
        out << ";";
    }
    return out.str();
}

std::string named_representation(const std::string& name, uint32_t k, uint32_t begin, bool control, uint32_t c) {
    if (name.find_first_of(",;:") != std::string::npos)
        throw std::invalid_argument{"Name cannot contain commas, colons, nor semicolons"};
    std::stringstream b;
    b << name << "," << (control ? 1 : 0) << "," << k << ":";
    if (control) b << c << ",";
    b << begin;
    for (uint32_t i = 1; i < k; ++i) b << "," << i + begin;
    b << ";";
    return b.str();
}

// Same gates by value (each gate's operator==), not only the same gate
// objects as QCircuit's operator== requires.
bool same_content(const QCircuit& a, const QCircuit& b) {
    if (a.qubit_count() != b.qubit_count() || a.gate_list().size() != b.gate_list().size() ||
        a.representation() != b.representation())
        return false;
    for (size_t i = 0; i < a.gate_list().size(); ++i)
        if (!(*a.gate_list()[i] == *b.gate_list()[i])) return false;
    return true;
}

// The circuit cache keyed by representation (src/quantum.cpp:1710-1724).  A
// circuit rebuilt with the same content (NControl_Gate called again, as in
// every Grover diffusion step) keeps the cached entry, and with it the
// entry's compiled passes and unitary; other circuits replace the entry.
std::shared_ptr<QCircuit> cache_circuit(const QCircuit& circuit) {
    auto it = g_cached_circuits.find(circuit.representation());
    if (it == g_cached_circuits.end()) {
        auto res = g_cached_circuits.insert({circuit.representation(), std::make_shared<QCircuit>(circuit)});
        if (!res.second) throw std::runtime_error{"Could not store circuit"};
        return res.first->second;
    }
    if (!(circuit == *(it->second)) && !same_content(circuit, *(it->second)))
        g_cached_circuits[circuit.representation()] = std::make_shared<QCircuit>(circuit);
    return g_cached_circuits.find(circuit.representation())->second;
}

}  // namespace

namespace detail {
struct Access {
    static void set_sink(QCircuit& c, Sink* s) { c.sink_ = s; }
    static void run_compiled(const QCircuit& c, aqs_state s) { c.run_compiled(s); }
    static std::shared_ptr<const std::vector<double>>& unitary_cache(const QCircuit& c) { return c.unitary64_; }
};
}  // namespace detail

namespace {

// The internal circuit's ops, qubits shifted by `offset` (and one extra
// control when control >= 0), emitted into qc one by one.
void emit_flattened(QCircuit& qc, const QCircuit& internal, uint32_t offset, int control) {
    QCircuit proxy(internal.qubit_count());
    detail::Sink sink;
    sink.forward = [&](const aqs_op& op) {
        aqs_op o = op;
        for (int j = 0; j < o.ntargets; ++j) o.targets[j] += (int)offset;
        for (int j = 0; j < o.ncontrols; ++j) o.controls[j] += (int)offset;
        o.more_controls <<= offset;
        if (control >= 0) {
            if (o.flags & AQS_CTRL_ANY)
                throw std::invalid_argument{"an Or gate cannot take an extra control in a wide ControlGate"};
            if (o.ncontrols < 3)
                o.controls[o.ncontrols++] = control;
            else
                o.more_controls |= 1ull << control;  // the fourth control and beyond
        }
        qc.emit(o);
    };
    detail::Access::set_sink(proxy, &sink);
    try {
        for (const auto& g : internal.gate_list()) (*g)(proxy);
    } catch (...) {
        detail::Access::set_sink(proxy, nullptr);
        throw;
    }
    detail::Access::set_sink(proxy, nullptr);
}

// The internal circuit's unitary in FP64 (row-major, interleaved), computed
// on the device with a complex128 state, one basis column at a time.
std::vector<double> unitary64(const QCircuit& internal) {
    const uint32_t k = internal.qubit_count();
    const uint64_t K = 1ull << k;
    aqs_state s = nullptr;
    check(aqs_state_create((int)k, AQS_C128, &s));
    std::vector<double> u(2 * K * K), col(2 * K);
    try {
        for (uint64_t j = 0; j < K; ++j) {
            check(aqs_state_set_basis(s, j));
            detail::Access::run_compiled(internal, s);
            check(aqs_state_download(s, col.data()));
            for (uint64_t r = 0; r < K; ++r) {
                u[2 * (r * K + j)] = col[2 * r];
                u[2 * (r * K + j) + 1] = col[2 * r + 1];
            }
        }
    } catch (...) {
        aqs_state_destroy(s);
        throw;
    }
    aqs_state_destroy(s);
    return u;
}

// True when the K x K unitary (row-major, interleaved) is exactly the
// identity except for its last diagonal entry.
bool phase_on_all_ones(const std::vector<double>& u, size_t K) {
    for (size_t r = 0; r < K; ++r)
        for (size_t c = 0; c < K; ++c) {
            if (r == K - 1 && c == K - 1) continue;
            const double re = u[2 * (r * K + c)], im = u[2 * (r * K + c) + 1];
            if (re != (r == c ? 1.0 : 0.0) || im != 0.0) return false;
        }
    return true;
}

void emit_subcircuit(QCircuit& qc, QCircuit& internal, uint32_t begin, int control) {
    internal.compile();
    const uint32_t k = internal.qubit_count();
    if (k <= 5) {
        auto& cached = detail::Access::unitary_cache(internal);
        if (!cached) cached = std::make_shared<const std::vector<double>>(unitary64(internal));
        const std::vector<double>& u = *cached;
        aqs_op op{};
        op.ntargets = (int)k;
        for (uint32_t j = 0; j < k; ++j) op.targets[j] = (int)(begin + j);
        if (control >= 0) op.controls[op.ncontrols++] = control;
        if (phase_on_all_ones(u, (size_t)1 << k)) {
            // diag(1, ..., 1, z) exactly (a multi-controlled Z or phase, as in
            // NControl_Gate): one PHASE op on the all-ones index of its
            // qubits, a plain phase pass instead of a dense 2^k block
            op.kind = AQS_OP_PHASE;
            const size_t last = ((size_t)1 << k) - 1;
            op.m[0] = u[2 * (last * ((size_t)1 << k) + last)];
            op.m[1] = u[2 * (last * ((size_t)1 << k) + last) + 1];
        } else {
            op.kind = AQS_OP_MATK;
            op.u = u.data();
        }
        qc.emit(op);
    } else {
        emit_flattened(qc, internal, begin, control);
    }
}

}  // namespace

Gate::Gate(const QCircuit& circuit_, uint32_t target_qubit_begin_, std::string name)
    : representation{}, qubit_count{circuit_.qubit_count()}, target_qubit_begin{target_qubit_begin_} {
    internal_circuit = cache_circuit(circuit_);
    if (name.empty())
        representation = shift_representation(circuit_.representation(), target_qubit_begin_, false, 0);
    else
        representation = named_representation(name, circuit_.qubit_count(), target_qubit_begin_, false, 0);
}

bool Gate::check(const QCircuit& qc) const {
    const auto qubits = qc.qubit_count();
    if (target_qubit_begin >= qubits) throw std::out_of_range{"Cannot add gate at the given qubit position"};
    if (target_qubit_begin + qubit_count > qubits) throw std::out_of_range{"Cannot add gate at the given qubit position"};
    return true;
}

QCircuit& Gate::operator()(QCircuit& qc) const {
    check(qc);
    emit_subcircuit(qc, *internal_circuit, target_qubit_begin, -1);
    return qc;
}

bool Gate::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit_begin == as<Gate>(rhs)->target_qubit_begin &&
           internal_circuit == as<Gate>(rhs)->internal_circuit;
}

ControlGate::ControlGate(const QCircuit& circuit_, uint32_t control_qubit_, uint32_t target_qubit_begin_,
                         std::string name)
    : representation{},
      qubit_count{circuit_.qubit_count()},
      control_qubit{control_qubit_},
      target_qubit_begin{target_qubit_begin_} {
    internal_circuit = cache_circuit(circuit_);
    if (name.empty())
        representation = shift_representation(circuit_.representation(), target_qubit_begin_, true, control_qubit_);
    else
        representation = named_representation(name, circuit_.qubit_count(), target_qubit_begin_, true, control_qubit_);
}

bool ControlGate::check(const QCircuit& qc) const {
    const auto qubits = qc.qubit_count();
    if (target_qubit_begin + qubit_count > qubits) throw std::out_of_range{"Gate must fit inside the circuit qubit count"};
    if (target_qubit_begin <= control_qubit && control_qubit < target_qubit_begin + qubit_count)
        throw std::out_of_range{"Control qubit cannot be one of the target qubits of the gate"};
    if (qubit_count >= qubits) throw std::invalid_argument{"Cannot add a bigger gate to the circuit"};
    if (control_qubit >= qubits) throw std::out_of_range{"Control qubit must be inside the circuit qubit range"};
    return true;
}

QCircuit& ControlGate::operator()(QCircuit& qc) const {
    check(qc);
    emit_subcircuit(qc, *internal_circuit, target_qubit_begin, (int)control_qubit);
    return qc;
}

bool ControlGate::operator==(const QGate& rhs) const noexcept {
    return type() == rhs.type() && target_qubit_begin == as<ControlGate>(rhs)->target_qubit_begin &&
           control_qubit == as<ControlGate>(rhs)->control_qubit &&
           internal_circuit == as<ControlGate>(rhs)->internal_circuit;
}

// ---- single-qubit host ops (src/quantum.cpp:1962-2032) --------------------------

namespace {
QState apply2(const QState& s, cfloat a, cfloat b, cfloat c, cfloat d) {
    const cfloat t0 = a * s[0] + b * s[1];
    const cfloat t1 = c * s[0] + d * s[1];
    return {t0, t1};
}
}  // namespace

QState X_op(const QState& s) { return apply2(s, 0.f, 1.f, 1.f, 0.f); }
QState Y_op(const QState& s) { return apply2(s, 0.f, cfloat{0.f, -1.f}, cfloat{0.f, 1.f}, 0.f); }
QState Z_op(const QState& s) { return apply2(s, 1.f, 0.f, 0.f, -1.f); }
QState RotateX_op(const QState& s, float angle) {
    const float c = std::cos(angle / 2.f), sn = std::sin(angle / 2.f);
    return apply2(s, c, cfloat{0.f, -sn}, cfloat{0.f, -sn}, c);
}
QState RotateY_op(const QState& s, float angle) {
    const float c = std::cos(angle / 2.f), sn = std::sin(angle / 2.f);
    return apply2(s, c, -sn, sn, c);
}
QState RotateZ_op(const QState& s, float angle) {
    const float c = std::cos(angle / 2.f), sn = std::sin(angle / 2.f);
    return apply2(s, cfloat{c, -sn}, 0.f, 0.f, cfloat{c, sn});
}
QState Hadamard_op(const QState& s) { return apply2(s, kHf, kHf, kHf, -kHf); }
QState Phase_op(const QState& s, float angle) {
    return apply2(s, 1.f, 0.f, 0.f, cfloat{std::cos(angle), std::sin(angle)});
}

}  // namespace aqs
