/*
 * aqs/quantum.h -- the aqs:: C++ API of afQuantumSim on the MI355X engine.
 *
 * Same classes, constructors, public fields, representation strings and
 * exception types as the reference's include/quantum.h, so user code that
 * builds QCircuits and runs QSimulators compiles against this header.  The
 * one deliberate change: ArrayFire types leave the interface.
 *   af::cfloat       -> aqs::cfloat (std::complex<float>)
 *   af::array state  -> aqs::StateVector (an aqs_state on the device,
 *                       include/aqs_hip.h), host copies via host()
 *   af::array matrix -> aqs::CircuitMatrix (host, column-major 2^n x 2^n),
 *                       materialised on demand for small circuits
 * Gates do not build 2^n x 2^n operators: QGate::operator()(QCircuit&)
 * emits one device pass (aqs_op) into the circuit, which either runs it on
 * the state being simulated or records it for compile()'s fusion.
 */
#ifndef AQS_QUANTUM_H
#define AQS_QUANTUM_H

#include <array>
#include <cassert>
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "aqs_hip.h"

namespace aqs {

using cfloat = std::complex<float>;

static constexpr float pi = 3.14159265358979323846f;
static constexpr uint32_t max_qubit_count = 30;

inline uint32_t fast_pow2(uint32_t p) { return 1u << p; }

class QGate;
class QState;
class QCircuit;
class QSimulator;

/* initialize (include/quantum.h:138-139): device = argv[1] or 0. */
void initialize(int argc, char** argv);
/* clear_circuit_cache (include/quantum.h:145). */
void clear_circuit_cache();
/* Seeds the host generator behind peek_measure / measure / profile_measure*
 * (the reference seeds it from std::random_device only). */
void seed(uint32_t value);

/* A single qubit on the host (include/quantum.h:151-342). */
class QState {
   public:
    QState() = default;
    QState(const std::complex<float>& zeroState, const std::complex<float>& oneState);
    QState(const std::array<std::complex<float>, 2>& states);

    static std::array<std::complex<float>, 2> create_state(float polar_angle, float azimuthal_angle) {
        return {std::cos(polar_angle / 2.f),
                std::complex<float>{std::cos(azimuthal_angle), std::sin(azimuthal_angle)} *
                    std::sin(polar_angle / 2.f)};
    }

    QState& set(const std::complex<float>& zero_state, const std::complex<float>& one_state);
    bool peek_measure() const;
    bool measure();
    std::array<uint32_t, 2> profile_measure(uint32_t rep_count) const;

    float probability_true() const noexcept {
        return state_[1].real() * state_[1].real() + state_[1].imag() * state_[1].imag();
    }
    float probability_false() const noexcept { return 1.f - probability_true(); }

    const cfloat& operator[](bool index) const noexcept { return state_[static_cast<int>(index)]; }
    bool operator!=(const QState& other) const noexcept {
        return state_[0] != other.state_[0] || state_[1] != other.state_[1];
    }
    bool operator==(const QState& other) const noexcept { return !(*this != other); }
    cfloat* data() noexcept { return state_; }
    const cfloat* data() const noexcept { return state_; }

    static const QState& zero() {
        static const QState s{1.0f, 0.0f};
        return s;
    }
    static const QState& one() {
        static const QState s{0.0f, 1.0f};
        return s;
    }
    static const QState& plus() {
        static const QState s{0.70710678118f, 0.70710678118f};
        return s;
    }
    static const QState& minus() {
        static const QState s{0.70710678118f, -0.70710678118f};
        return s;
    }

   private:
    cfloat state_[2]{{1.0f, 0.0f}, {0.0f, 0.0f}};
    void force_normalize();
};

/* Host copy of a circuit's unitary, column-major (element (r, c) at c*N + r). */
struct CircuitMatrix {
    uint32_t dim = 0;
    std::vector<cfloat> data;
    cfloat operator()(uint32_t r, uint32_t c) const { return data[(size_t)c * dim + r]; }
    bool operator==(const CircuitMatrix& o) const { return dim == o.dim && data == o.data; }
};

/* Device state vector of a QSimulator (replaces its af::array statevector_). */
class StateVector {
   public:
    StateVector() = default;
    explicit StateVector(uint32_t qubits);
    StateVector(const StateVector& o);
    StateVector& operator=(const StateVector& o);
    StateVector(StateVector&& o) noexcept;
    StateVector& operator=(StateVector&& o) noexcept;
    ~StateVector();

    aqs_state handle() const noexcept { return h_; }
    uint32_t qubits() const noexcept { return n_; }
    size_t size() const noexcept { return (size_t)1 << n_; }
    std::vector<cfloat> host() const;
    void host(cfloat* out) const;
    void upload(const cfloat* in, bool normalize);

   private:
    aqs_state h_ = nullptr;
    uint32_t n_ = 0;
};

namespace detail {
struct OpRecord;  // an aqs_op with owned matrix storage
struct Segment;   // one compiled piece of a circuit
struct Sink;      // where emitted ops go
struct Access;    // library-internal access to QCircuit's sink and compiled part
}  // namespace detail

/* Quantum circuit (include/quantum.h:350-469). */
class QCircuit {
   public:
    friend class QSimulator;
    friend struct detail::Access;

    QCircuit() = delete;
    QCircuit(const QCircuit&);
    QCircuit(QCircuit&&) noexcept;
    QCircuit& operator=(const QCircuit&);
    QCircuit& operator=(QCircuit&&) noexcept;
    ~QCircuit();

    QCircuit(uint32_t qubit_count);

    template <typename T>
    friend QCircuit& operator<<(QCircuit& qc, const T& gate);
    template <typename T>
    friend QCircuit& operator<<(QCircuit& qc, const std::vector<T>& gates);

    uint32_t qubit_count() const noexcept { return qubits_; }
    uint32_t state_count() const noexcept { return fast_pow2(qubits_); }

    /* The unitary of the compiled gates [0, cached index), materialised on
     * the device column by column (circuits up to 12 qubits). */
    const CircuitMatrix& circuit() const;

    auto& gate_list() noexcept { return gate_list_; }
    const auto& gate_list() const noexcept { return gate_list_; }
    std::string& representation() noexcept { return representation_; }
    const std::string& representation() const noexcept { return representation_; }

    /* Compiles the not-yet-compiled gates into device passes: <= 5-qubit
     * dense blocks, or from 20 qubits up LDS-resident tile passes when the
     * library's cost model rates them cheaper (aqs_fuse_auto). */
    void compile();
    void clear();
    void clear_cache();

    /* Where a gate's operator() sends its device pass. */
    void emit(const aqs_op& op);

    /* Number of fused device passes the compiled part runs. */
    size_t compiled_passes() const;
    /* compile() with dense blocks of at most k (1..5) qubits only; 0 (the
     * default) lets the library choose between blocks and tile passes. */
    void set_fusion_width(int k) { fusion_width_ = k; }

    friend bool operator==(const QCircuit& lhs, const QCircuit& rhs);

   private:
    std::vector<std::shared_ptr<QGate>> gate_list_;
    std::string representation_;
    uint32_t qubits_ = 0;
    std::size_t cached_index_ = 0;
    int fusion_width_ = 0;
    std::vector<std::shared_ptr<detail::Segment>> compiled_;
    detail::Sink* sink_ = nullptr;
    mutable std::shared_ptr<CircuitMatrix> matrix_;
    /* FP64 unitary of the compiled part (up to 5 qubits), computed once for
     * Gate / ControlGate emission; dropped with matrix_. */
    mutable std::shared_ptr<const std::vector<double>> unitary64_;

    void run_compiled(aqs_state s) const;
};

class QNoise {};

/* Simulator (include/quantum.h:482-742). */
class QSimulator {
   public:
    enum class Basis : int8_t { Z, Y, X };

    explicit QSimulator(uint32_t qubit_count, const QState& initial_state = aqs::QState::zero(),
                        const QNoise& noise_generator = QNoise{});
    explicit QSimulator(uint32_t qubit_count, std::vector<QState> initial_states,
                        const QNoise& noise_generator = QNoise{});
    /* From a host vector of 2^n amplitudes, normalised in FP64. */
    explicit QSimulator(uint32_t qubit_count, const std::vector<cfloat>& statevector,
                        const QNoise& noise_generator = QNoise{});

    void generate_statevector();
    void simulate(const QCircuit& circuit);

    bool peek_measure(uint32_t qubit) const;
    bool measure(uint32_t qubit);
    uint32_t measure_all();
    uint32_t peek_measure_all() const;
    std::array<uint32_t, 2> profile_measure(uint32_t qubit, uint32_t rep_count) const;
    std::vector<uint32_t> profile_measure_all(uint32_t rep_count) const;

    /* Sampling with caller-supplied uniforms in [0, 1): index k is the first
     * basis state whose cumulative probability exceeds u[k]. */
    std::vector<uint64_t> sample(const std::vector<double>& uniforms) const;

    QState& qubit(uint32_t index) noexcept {
        assert(index < qubit_count());
        return states_[index];
    }
    const QState& qubit(uint32_t index) const noexcept {
        assert(index < qubit_count());
        return states_[index];
    }

    cfloat state(uint32_t state) const;
    float qubit_probability_true(uint32_t qubit) const;
    float qubit_probability_false(uint32_t qubit) const { return 1.f - qubit_probability_true(qubit); }
    float state_probability(uint32_t state) const;
    std::vector<float> probabilities() const;
    uint32_t qubit_count() const noexcept { return qubits_; }
    uint32_t state_count() const noexcept { return fast_pow2(qubits_); }
    Basis get_basis() const noexcept { return basis_; }
    void set_basis(Basis basis);

    StateVector& statevector() noexcept { return statevector_; }
    const StateVector& statevector() const noexcept { return statevector_; }

   private:
    std::vector<QState> states_;
    StateVector statevector_;
    QNoise noise_;
    uint32_t qubits_;
    Basis basis_ = Basis::Z;
};

/* Gate base class (include/quantum.h:748-824). */
class QGate {
   public:
    virtual ~QGate() = default;
    virtual QCircuit& operator()(QCircuit& qc) const = 0;
    virtual std::string to_string() const = 0;
    virtual uint32_t type() const noexcept = 0;
    virtual bool check(const QCircuit& qc) const = 0;
    virtual bool operator==(const QGate& rhs) const noexcept = 0;

   protected:
    enum class GateTypes : uint32_t {
        Barrier, X, Y, Z, Hadamard, Phase, Swap, RotX, RotY, RotZ, CX, CY, CZ, CHadamard, CPhase,
        CSwap, CRotX, CRotY, CRotZ, CCX, Or, Circuit, ControlCircuit
    };
};

template <typename T>
QCircuit& operator<<(QCircuit& qc, const T& gate) {
    static_assert(std::is_base_of<QGate, T>::value, "Gate must inherit from QGate class");
    if (gate.check(qc)) {
        qc.representation_.append(gate.to_string());
        qc.gate_list().push_back(std::make_shared<T>(gate));
    }
    return qc;
}

template <typename T>
QCircuit& operator<<(QCircuit& qc, const std::vector<T>& gates) {
    static_assert(std::is_base_of<QGate, T>::value, "Gate must inherit from QGate class");
    for (const auto& gate : gates) {
        if (gate.check(qc)) {
            qc.representation_.append(gate.to_string());
            qc.gate_list().push_back(std::make_shared<T>(gate));
        }
    }
    return qc;
}

#define AQS_GATE_COMMON(NAME, TYPE)                                                   \
    QCircuit& operator()(QCircuit&) const override;                                   \
    std::string to_string() const override;                                          \
    bool check(const QCircuit&) const override;                                       \
    bool operator==(const QGate& rhs) const noexcept override;                        \
    uint32_t type() const noexcept override { return static_cast<uint32_t>(GateTypes::TYPE); } \
    static constexpr uint32_t static_type() noexcept { return static_cast<uint32_t>(GateTypes::TYPE); }

class Barrier : public QGate {
   public:
    Barrier(bool visible_ = true) noexcept : visible{visible_} {}
    bool check(const QCircuit&) const override { return true; }
    QCircuit& operator()(QCircuit& qc) const override { return qc; }
    std::string to_string() const override { return visible ? "B;" : "P;"; }
    uint32_t type() const noexcept override { return static_cast<uint32_t>(GateTypes::Barrier); }
    bool operator==(const QGate& rhs) const noexcept override { return type() == rhs.type(); }
    static constexpr uint32_t static_type() noexcept { return static_cast<uint32_t>(GateTypes::Barrier); }
    bool visible = true;
};

class X : public QGate {
   public:
    X(uint32_t target_qubit_) noexcept : target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(X, X)
    static const QCircuit& gate();
    uint32_t target_qubit;
};
using Not = X;

class Y : public QGate {
   public:
    Y(uint32_t target_qubit_) noexcept : target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(Y, Y)
    static const QCircuit& gate();
    uint32_t target_qubit;
};

class Z : public QGate {
   public:
    Z(uint32_t target_qubit_) noexcept : target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(Z, Z)
    static const QCircuit& gate();
    uint32_t target_qubit;
};

class RotX : public QGate {
   public:
    RotX(uint32_t target_qubit_, float angle_) noexcept : target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(RotX, RotX)
    static QCircuit gate(float angle);
    uint32_t target_qubit;
    float angle;
};

class RotY : public QGate {
   public:
    RotY(uint32_t target_qubit_, float angle_) noexcept : target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(RotY, RotY)
    static QCircuit gate(float angle);
    uint32_t target_qubit;
    float angle;
};

class RotZ : public QGate {
   public:
    RotZ(uint32_t target_qubit_, float angle_) noexcept : target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(RotZ, RotZ)
    static QCircuit gate(float angle);
    uint32_t target_qubit;
    float angle;
};

class H : public QGate {
   public:
    H(uint32_t target_qubit_) noexcept : target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(H, Hadamard)
    static const QCircuit& gate();
    uint32_t target_qubit;
};

class Phase : public QGate {
   public:
    Phase(uint32_t target_qubit_, float angle_) noexcept : target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(Phase, Phase)
    static QCircuit gate(float angle);
    uint32_t target_qubit;
    float angle;
};

/* U3(theta, phi, lambda) = Phase(phi) RotY(theta) Phase(lambda): no
 * counterpart in the reference (its TODO list asks for a unitary gate). */
class U3 : public QGate {
   public:
    U3(uint32_t target_qubit_, float theta_, float phi_, float lambda_) noexcept
        : target_qubit{target_qubit_}, theta{theta_}, phi{phi_}, lambda{lambda_} {}
    QCircuit& operator()(QCircuit&) const override;
    std::string to_string() const override;
    bool check(const QCircuit&) const override;
    bool operator==(const QGate& rhs) const noexcept override;
    /* shares the Circuit type id space past the reference's enum */
    uint32_t type() const noexcept override { return static_type(); }
    static constexpr uint32_t static_type() noexcept { return 64; }
    uint32_t target_qubit;
    float theta, phi, lambda;
};

class Swap : public QGate {
   public:
    Swap(uint32_t target_qubit_A_, uint32_t target_qubit_B_) noexcept
        : target_qubit_A{target_qubit_A_}, target_qubit_B{target_qubit_B_} {}
    AQS_GATE_COMMON(Swap, Swap)
    uint32_t target_qubit_A;
    uint32_t target_qubit_B;
};

class CX : public QGate {
   public:
    CX(uint32_t control_qubit_, uint32_t target_qubit_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(CX, CX)
    uint32_t control_qubit;
    uint32_t target_qubit;
};
using CNot = CX;
using Xor = CX;

class CY : public QGate {
   public:
    CY(uint32_t control_qubit_, uint32_t target_qubit_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(CY, CY)
    uint32_t control_qubit;
    uint32_t target_qubit;
};

class CZ : public QGate {
   public:
    CZ(uint32_t control_qubit_, uint32_t target_qubit_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(CZ, CZ)
    uint32_t control_qubit;
    uint32_t target_qubit;
};

class CH : public QGate {
   public:
    CH(uint32_t control_qubit_, uint32_t target_qubit_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(CH, CHadamard)
    uint32_t control_qubit;
    uint32_t target_qubit;
};

class CPhase : public QGate {
   public:
    CPhase(uint32_t control_qubit_, uint32_t target_qubit_, float angle_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(CPhase, CPhase)
    uint32_t control_qubit;
    uint32_t target_qubit;
    float angle;
};

class CSwap : public QGate {
   public:
    CSwap(uint32_t control_qubit_, uint32_t target_qubit_A_, uint32_t target_qubit_B_) noexcept
        : control_qubit{control_qubit_}, target_qubit_A{target_qubit_A_}, target_qubit_B{target_qubit_B_} {}
    AQS_GATE_COMMON(CSwap, CSwap)
    uint32_t control_qubit;
    uint32_t target_qubit_A;
    uint32_t target_qubit_B;
};

class CRotX : public QGate {
   public:
    CRotX(uint32_t control_qubit_, uint32_t target_qubit_, float angle_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(CRotX, CRotX)
    uint32_t control_qubit;
    uint32_t target_qubit;
    float angle;
};

class CRotY : public QGate {
   public:
    CRotY(uint32_t control_qubit_, uint32_t target_qubit_, float angle_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(CRotY, CRotY)
    uint32_t control_qubit;
    uint32_t target_qubit;
    float angle;
};

class CRotZ : public QGate {
   public:
    CRotZ(uint32_t control_qubit_, uint32_t target_qubit_, float angle_) noexcept
        : control_qubit{control_qubit_}, target_qubit{target_qubit_}, angle{angle_} {}
    AQS_GATE_COMMON(CRotZ, CRotZ)
    uint32_t control_qubit;
    uint32_t target_qubit;
    float angle;
};

class CCNot : public QGate {
   public:
    CCNot(uint32_t control_qubit_A_, uint32_t control_qubit_B_, uint32_t target_qubit_) noexcept
        : control_qubit_A{control_qubit_A_}, control_qubit_B{control_qubit_B_}, target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(CCNot, CCX)
    uint32_t control_qubit_A;
    uint32_t control_qubit_B;
    uint32_t target_qubit;
};
using And = CCNot;

class Or : public QGate {
   public:
    Or(uint32_t control_qubit_A_, uint32_t control_qubit_B_, uint32_t target_qubit_) noexcept
        : control_qubit_A{control_qubit_A_}, control_qubit_B{control_qubit_B_}, target_qubit{target_qubit_} {}
    AQS_GATE_COMMON(Or, Or)
    uint32_t control_qubit_A;
    uint32_t control_qubit_B;
    uint32_t target_qubit;
};

class Gate : public QGate {
   public:
    Gate(const QCircuit& circuit_, uint32_t target_qubit_begin_, std::string name = "");
    bool check(const QCircuit&) const override;
    QCircuit& operator()(QCircuit&) const override;
    std::string to_string() const override { return representation; }
    uint32_t type() const noexcept override { return static_cast<uint32_t>(GateTypes::Circuit); }
    bool operator==(const QGate& rhs) const noexcept override;
    static constexpr uint32_t static_type() noexcept { return static_cast<uint32_t>(GateTypes::Circuit); }

    std::shared_ptr<QCircuit> internal_circuit;
    std::string representation;
    uint32_t qubit_count;
    uint32_t target_qubit_begin;
};

class ControlGate : public QGate {
   public:
    ControlGate(const QCircuit& circuit_, uint32_t control_qubit_, uint32_t target_qubit_begin_,
                std::string name = "");
    bool check(const QCircuit&) const override;
    QCircuit& operator()(QCircuit&) const override;
    std::string to_string() const override { return representation; }
    uint32_t type() const noexcept override { return static_cast<uint32_t>(GateTypes::ControlCircuit); }
    bool operator==(const QGate& rhs) const noexcept override;
    static constexpr uint32_t static_type() noexcept { return static_cast<uint32_t>(GateTypes::ControlCircuit); }

    std::shared_ptr<QCircuit> internal_circuit;
    std::string representation;
    uint32_t qubit_count;
    uint32_t control_qubit;
    uint32_t target_qubit_begin;
};

#undef AQS_GATE_COMMON

/* Single-qubit helpers on host QStates (include/quantum.h:1590-1653). */
QState X_op(const QState& state);
QState Y_op(const QState& state);
QState Z_op(const QState& state);
QState RotateX_op(const QState& state, float angle);
QState RotateY_op(const QState& state, float angle);
QState RotateZ_op(const QState& state, float angle);
QState Hadamard_op(const QState& state);
QState Phase_op(const QState& state, float angle);

}  // namespace aqs

#endif
