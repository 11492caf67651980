// The reference's published algorithm benchmarks (benchmark/benchmark.cpp,
// numbers in benchmark/results.md:33-35) run through this library's aqs::
// C++ API: QFT on 10 qubits, Grover search on 10 qubits (25 iterations) and
// Shor's period finding for N = 15, a = 13 on 12 qubits.  As there, each
// timed repetition builds the circuit and the simulator, runs simulate()
// on the uncompiled circuit and (Shor) samples 10^4 shots; here every
// repetition also waits for the device (the reference's timer stopped
// without af::sync, benchmark/benchmark.cpp:191-197).  A second pass times
// the same work with compile() before simulate().
//
// Differences kept on purpose: Shor's state is regenerated after the last
// qubit is set to |1> (the reference sets it without generate_statevector,
// so its run starts from |0...0>), and the factors found are checked.
//
// Prints one JSON line: per benchmark the mean and sample standard
// deviation in ms over `reps` repetitions (default 100, after one untimed
// warm-up that loads the device code), plus correctness checks.
#include <aqs/quantum.h>
#include <aqs/quantum_algo.h>
#include <aqs_hip.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <set>
#include <string>
#include <vector>

namespace {

void wait_for(const aqs::QSimulator& qs) { aqs_state_sync(qs.statevector().handle()); }

double qft(uint32_t n, bool compile) {
    aqs::QCircuit qc = aqs::fourier_transform(n);
    if (compile) qc.compile();
    aqs::QSimulator qs{n};
    qs.simulate(qc);
    wait_for(qs);
    return 0.0;
}

// Largest deviation of QFT|0...0> from the uniform superposition.
double qft_check(uint32_t n) {
    aqs::QCircuit qc = aqs::fourier_transform(n);
    aqs::QSimulator qs{n};
    qs.simulate(qc);
    const double u = 1.0 / std::sqrt((double)(1u << n));
    double err = 0.0;
    for (const auto& a : qs.statevector().host()) err = std::max(err, std::abs(std::complex<double>(a.real() - u, a.imag())));
    return err;
}

double grover(uint32_t n, bool compile) {
    const uint32_t marked = 5;
    const auto iterations = static_cast<uint32_t>(aqs::pi * std::sqrt((float)(1u << n)) / 4.f);
    aqs::QCircuit qc = aqs::grover_search(n, aqs::grover_oracle(n, marked), iterations);
    if (compile) qc.compile();
    aqs::QSimulator qs{n};
    qs.simulate(qc);
    wait_for(qs);
    // qubit i carries bit i of the marked value; qubit 0 is the top index bit
    uint32_t index = 0;
    for (uint32_t i = 0; i < n; ++i)
        if ((marked >> i) & 1u) index |= 1u << (n - 1 - i);
    return qs.state_probability(index);
}

// The 4-qubit permutation circuit of x -> a^(2^j) x mod 15 for a = 13 (a
// rotation of the 4 bits, then every bit flipped), composed with itself j
// times as the reference does.
aqs::QCircuit mult_mod15(int a, uint32_t j) {
    aqs::QCircuit c{4};
    if (a == 13 || a == 2) c << aqs::Swap{0, 1} << aqs::Swap{1, 2} << aqs::Swap{2, 3};
    if (a == 7 || a == 8) c << aqs::Swap{2, 3} << aqs::Swap{1, 2} << aqs::Swap{0, 1};
    if (a == 11 || a == 4) c << aqs::Swap{1, 3} << aqs::Swap{0, 2};
    if (a == 7 || a == 11 || a == 13)
        for (uint32_t q = 0; q < 4; ++q) c << aqs::X{q};
    for (uint32_t i = 0; i < j; ++i) c << aqs::Gate{c, 0};
    return c;
}

// Best fraction p/q with q <= max_den from the continued-fraction
// convergents of x.
std::pair<long, long> convergent(double x, long max_den) {
    long p0 = 0, q0 = 1, p1 = 1, q1 = 0;
    double v = x;
    for (int it = 0; it < 64; ++it) {
        const long a = (long)std::floor(v);
        const long p2 = a * p1 + p0, q2 = a * q1 + q0;
        if (q2 > max_den) break;
        p0 = p1;
        q0 = q1;
        p1 = p2;
        q1 = q2;
        const double f = v - (double)a;
        if (f < 1e-12) break;
        v = 1.0 / f;
    }
    return {p1, q1};
}

long ipow(long b, long e) {
    long r = 1;
    while (e-- > 0) r *= b;
    return r;
}

// The benchmark's Shor instance; with `factors`, collects the nontrivial
// factors of 15 that the most likely measured periods give.
double shor(bool compile, std::set<long>* factors) {
    const uint32_t out = 8;
    const int N = 15, a = 13;
    aqs::QCircuit qc{out + 4};
    aqs::QSimulator qs{out + 4};
    qs.qubit(qs.qubit_count() - 1) = aqs::QState::one();
    qs.generate_statevector();
    aqs::QCircuit h1{out / 2}, h2{out - out / 2};
    for (uint32_t i = 0; i < h1.qubit_count(); ++i) h1 << aqs::H{i};
    for (uint32_t i = 0; i < h2.qubit_count(); ++i) h2 << aqs::H{i};
    qc << aqs::Gate{h1, 0} << aqs::Gate{h2, h1.qubit_count()};
    for (uint32_t i = 0; i < out; ++i)
        qc << aqs::ControlGate{mult_mod15(a, i), out - 1 - i, out, "a^(2^" + std::to_string(i) + ") mod 15"};
    qc << aqs::Gate{aqs::inverse_fourier_transform(out), 0, "QFT+"};
    if (compile) qc.compile();
    qs.simulate(qc);
    const auto hist = qs.profile_measure_all(10000);
    (void)hist;
    // most likely outputs from the per-qubit marginals, as the reference does
    std::vector<float> p1(out);
    for (uint32_t q = 0; q < out; ++q) p1[q] = qs.qubit_probability_true(q);
    std::vector<std::pair<float, uint32_t>> ranked;
    for (uint32_t s = 0; s < (1u << out); ++s) {
        float p = 1.f;
        for (uint32_t q = 0; q < out; ++q) p *= (s & (1u << q)) ? p1[q] : 1.f - p1[q];
        ranked.push_back({p, s});
    }
    std::partial_sort(ranked.begin(), ranked.begin() + 4, ranked.end(),
                      [](const auto& x, const auto& y) { return x.first > y.first; });
    for (int i = 0; i < 4; ++i) {
        const long r = convergent((double)ranked[i].second / (double)(1u << out), N).second;
        if (r % 2) continue;
        for (long g : {std::gcd(ipow(a, r / 2) - 1, (long)N), std::gcd(ipow(a, r / 2) + 1, (long)N)})
            if (factors && g != 1 && g != N) factors->insert(g);
    }
    return 0.0;
}

struct Stat {
    double mean_ms, std_ms;
};

Stat time_it(int reps, const std::function<void()>& f) {
    f();  // warm-up: device code loading, first allocations
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        f();
        const auto t1 = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    const double m = std::accumulate(t.begin(), t.end(), 0.0) / reps;
    double v = 0.0;
    for (double x : t) v += (x - m) * (x - m);
    return {m, reps > 1 ? std::sqrt(v / (reps - 1)) : 0.0};
}

}  // namespace

int main(int argc, char** argv) {
    const int reps = argc > 1 ? std::max(1, std::atoi(argv[1])) : 100;
    char* dev_argv[] = {argv[0], nullptr};
    aqs::initialize(1, dev_argv);
    std::string js = "{\"reps\": " + std::to_string(reps);
    for (bool compile : {false, true}) {
        aqs::clear_circuit_cache();
        const Stat g = time_it(reps, [&] { grover(10, compile); });
        aqs::clear_circuit_cache();
        const Stat q = time_it(reps, [&] { qft(10, compile); });
        aqs::clear_circuit_cache();
        const Stat s = time_it(reps, [&] { shor(compile, nullptr); });
        char buf[512];
        std::snprintf(buf, sizeof buf,
                      ", \"%s\": {\"qft10_ms\": [%.4f, %.4f], \"grover10_ms\": [%.4f, %.4f], \"shor12_ms\": [%.4f, %.4f]}",
                      compile ? "compiled" : "uncompiled", q.mean_ms, q.std_ms, g.mean_ms, g.std_ms, s.mean_ms,
                      s.std_ms);
        js += buf;
    }
    std::set<long> factors;
    shor(false, &factors);
    std::string fs;
    for (long f : factors) fs += (fs.empty() ? "" : ", ") + std::to_string(f);
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  ", \"checks\": {\"qft10_max_dev_from_uniform\": %.3e, \"grover10_marked_probability\": %.6f, "
                  "\"shor15_factors\": [%s]}}",
                  qft_check(10), grover(10, false), fs.c_str());
    js += buf;
    std::puts(js.c_str());
    return 0;
}
