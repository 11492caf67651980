// Host tests of the request server's scheduling core (serve_core.cpp) with a
// stand-in backend: no GPU.  Built twice by afquantumsim_amd/build.py, plain
// and with -fsanitize=thread, and run by tests/test_serve_cpu.py.
//
// The stand-in's "state" of n qubits holds 2^n bytes of an injected budget;
// "running" a request sleeps for the time its first uniform asks for (in
// milliseconds), polling the cancel flag, and writes indices derived from the
// uniforms.  It checks the core's invariants from the device side: resident
// bytes never exceed the budget, and running requests never exceed the
// workers or the budget.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "serve_core.h"

using namespace aqs_serve;

namespace {

int g_checks = 0;
const char* g_case = "";

void require(bool ok, const char* what) {
    ++g_checks;
    if (!ok) {
        std::fprintf(stderr, "FAILED [%s] %s\n", g_case, what);
        std::exit(1);
    }
}

class FakeBackend final : public Backend {
   public:
    explicit FakeBackend(size_t budget) : budget_(budget) {}
    ~FakeBackend() override {
        if (on_destroy_) on_destroy_(created_.load(), destroyed_.load(), violation_.load());
    }
    size_t footprint(int n) const override { return n == kGroupKey ? arena_ : (size_t)1 << n; }
    void* create_state(int n) override {
        const size_t b = footprint(n);
        const size_t now = live_bytes_.fetch_add(b) + b;
        if (now > budget_) violation_ = true;
        created_.fetch_add(1);
        return new int(n);
    }
    void destroy_state(void* st) override {
        int* p = (int*)st;
        live_bytes_.fetch_sub(footprint(*p));
        destroyed_.fetch_add(1);
        delete p;
    }
    void run(Request& r, void* st, int) override {
        const int n = *(int*)st;
        if (n != r.n) violation_ = true;
        const int now = running_.fetch_add(1) + 1;
        const size_t bytes = running_bytes_.fetch_add(footprint(r.n)) + footprint(r.n);
        int prev = max_running_.load();
        while (now > prev && !max_running_.compare_exchange_weak(prev, now)) {
        }
        if (bytes > budget_) violation_ = true;
        const double ms = r.u.empty() ? 0.0 : r.u[0];
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds((long)(ms * 1000));
        bool cancelled = false;
        while (std::chrono::steady_clock::now() < until) {
            if (r.cancel.load()) {
                cancelled = true;
                break;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        running_bytes_.fetch_sub(footprint(r.n));
        running_.fetch_sub(1);
        if (cancelled) throw ServeError{AQS_ERR_RUNTIME, "request cancelled"};
        if (r.n == fail_qubits_) throw std::runtime_error("backend failure");
        if (r.n == oom_qubits_) throw std::bad_alloc();
        r.indices.resize(r.shots);
        for (uint32_t k = 0; k < r.shots; ++k) r.indices[k] = (uint64_t)(r.u[k] * 1000.0) + r.id;
        r.norm2 = 1.0;
        r.passes = r.ops.size();
    }

    // Groups: members of at most group_qubits_ qubits, each taking item_bytes_
    // of an arena of arena_ bytes; "running" a group sleeps for the longest
    // member's time.  Every group's member ids are recorded in launch order.
    int group_max_qubits() const override { return group_qubits_; }
    size_t group_bytes(const Request& r) const override { return r.n == unfit_qubits_ ? arena_ + 1 : item_bytes_; }
    void run_group(std::vector<Request*>& rs, void* arena, int) override {
        if (*(int*)arena != kGroupKey) violation_ = true;
        size_t bytes = 0;
        double ms = 0.0;
        for (Request* r : rs) {
            if (r->n > group_qubits_) violation_ = true;
            bytes += item_bytes_;
            ms = std::max(ms, r->u.empty() ? 0.0 : r->u[0]);
        }
        if (bytes > arena_) violation_ = true;
        {
            std::lock_guard<std::mutex> lock(groups_mu_);
            std::vector<uint64_t> ids;
            for (Request* r : rs) ids.push_back(r->id);
            groups_.push_back(ids);
        }
        const int now = running_.fetch_add((int)rs.size()) + (int)rs.size();
        const size_t held = running_bytes_.fetch_add(arena_) + arena_;
        int prev = max_running_.load();
        while (now > prev && !max_running_.compare_exchange_weak(prev, now)) {
        }
        if (held > budget_) violation_ = true;
        std::this_thread::sleep_for(std::chrono::microseconds((long)(ms * 1000)));
        running_bytes_.fetch_sub(arena_);
        running_.fetch_sub((int)rs.size());
        for (Request* r : rs)
            if (r->n == fail_qubits_) throw std::runtime_error("group failure");
        for (Request* r : rs) {
            r->indices.resize(r->shots);
            for (uint32_t k = 0; k < r->shots; ++k) r->indices[k] = (uint64_t)(r->u[k] * 1000.0) + r->id;
            r->norm2 = 1.0;
            r->passes = 1;
        }
    }
    std::vector<std::vector<uint64_t>> groups() {
        std::lock_guard<std::mutex> lock(groups_mu_);
        return groups_;
    }

    int group_qubits_ = 0, unfit_qubits_ = -1;
    size_t arena_ = 0, item_bytes_ = 0;
    std::mutex groups_mu_;
    std::vector<std::vector<uint64_t>> groups_;

    size_t budget_;
    std::atomic<size_t> live_bytes_{0}, running_bytes_{0};
    std::atomic<int> running_{0}, max_running_{0}, created_{0}, destroyed_{0};
    std::atomic<bool> violation_{false};
    int fail_qubits_ = -1, oom_qubits_ = -1;
    std::function<void(int, int, bool)> on_destroy_;  // final record, reported as the backend goes
};

std::shared_ptr<Request> make(uint64_t id, int n, double ms, uint32_t shots = 4) {
    auto r = std::make_shared<Request>();
    r->id = id;
    r->n = n;
    r->shots = shots;
    r->u.assign(shots ? shots : 1, 0.0);
    r->u[0] = ms;
    for (uint32_t k = 1; k < shots; ++k) r->u[k] = 0.001 * k;
    r->ops.resize(3);
    return r;
}

template <typename F>
bool raises(aqs_status code, F&& f) {
    try {
        f();
    } catch (const ServeError& e) {
        return e.code == code;
    }
    return false;
}

// Every request completes with its own result; concurrency stays within the
// workers.
void basic() {
    g_case = "basic";
    auto* be = new FakeBackend(1 << 20);
    int created = 0, destroyed = 0;
    bool violation = false;
    {
        auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 1 << 20, 4);
    Server& s = *s_owner;
        for (uint64_t id = 0; id < 24; ++id) s.submit(make(id, 8 + (int)(id % 5), 2.0));
        for (uint64_t id = 0; id < 24; ++id) {
            auto r = s.wait(id, -1);
            require(r->status == AQS_OK, "status");
            require(r->indices.size() == 4 && r->indices[1] == 1 + id, "indices belong to the request");
            require(r->t_submit <= r->t_admit && r->t_admit <= r->t_done, "timestamps ordered");
        }
        const Stats st = s.stats();
        require(st.completed == 24 && st.admitted == 24 && st.pending == 0, "counts");
        require(st.max_concurrent <= 4 && be->max_running_.load() <= 4, "at most 4 running");
        require(st.max_concurrent >= 2, "requests overlapped");
        // the backend is owned by the server: read its record before the
        // server (and with it the backend) goes; destruction frees the idle
        // states through a second backend pointer taken here
        be->on_destroy_ = [&](int c, int d, bool v) {
            created = c;
            destroyed = d;
            violation = v;
        };
    }
    require(created > 0 && created == destroyed, "every state freed at destruction");
    require(!violation, "device-side invariants");
}

// Admission by footprint: with room for three 10-qubit states, eight workers
// never run more than three of them; a request larger than the budget and a
// duplicate id are rejected at submit.
void budget_admission() {
    g_case = "budget admission";
    const size_t budget = 3 * 1024;
    auto* be = new FakeBackend(budget);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), budget, 8);
    Server& s = *s_owner;
    for (uint64_t id = 0; id < 12; ++id) s.submit(make(id, 10, 5.0));
    require(raises(AQS_ERR_INVALID_ARGUMENT, [&] { s.submit(make(100, 12, 1.0)); }), "larger than the budget");
    require(raises(AQS_ERR_INVALID_ARGUMENT, [&] { s.submit(make(5, 10, 1.0)); }), "duplicate id");
    for (uint64_t id = 0; id < 12; ++id) require(s.wait(id, -1)->status == AQS_OK, "status");
    const Stats st = s.stats();
    require(st.max_concurrent <= 3 && be->max_running_.load() <= 3, "at most three fit");
    require(st.rejected == 2, "two rejections counted");
    require(st.peak_resident_bytes <= budget && !be->violation_, "resident bytes within the budget");
    require(raises(AQS_ERR_OUT_OF_RANGE, [&] { s.wait(100, 0); }), "rejected request is unknown");
}

// Starvation limit: a large request that does not fit while small ones run
// may be overtaken only until it has waited the limit; after that nothing
// submitted later starts before it.
void starvation() {
    g_case = "starvation";
    const size_t budget = 4096;  // one 12-qubit state, or four 10-qubit ones
    auto* be = new FakeBackend(budget);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), budget, 4);
    Server& s = *s_owner;
    s.set_starvation_limit(0.02);
    s.submit(make(1, 10, 60.0));  // occupies a quarter of the budget for 60 ms
    s.submit(make(2, 12, 5.0));   // needs everything
    std::this_thread::sleep_for(std::chrono::milliseconds(40));  // request 2 is now starved
    for (uint64_t id = 3; id < 9; ++id) s.submit(make(id, 10, 5.0));
    auto big = s.wait(2, -1);
    require(big->status == AQS_OK, "big request ran");
    for (uint64_t id = 3; id < 9; ++id) {
        auto r = s.wait(id, -1);
        require(r->t_admit >= big->t_admit, "nothing overtakes a starved request");
    }
    require(s.wait(1, -1)->status == AQS_OK, "first request");
    require(big->t_admit >= big->t_submit + 0.02, "big request waited for room");
    require(!be->violation_, "device-side invariants");

    // without starvation (a long limit) small requests overtake
    g_case = "overtaking";
    auto* be2 = new FakeBackend(budget);
    auto s2_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be2), budget, 4);
    Server& s2 = *s2_owner;
    s2.set_starvation_limit(10.0);
    s2.submit(make(1, 10, 40.0));
    s2.submit(make(2, 12, 1.0));
    s2.submit(make(3, 10, 1.0));
    auto r3 = s2.wait(3, -1);
    auto r2 = s2.wait(2, -1);
    require(r3->t_admit < r2->t_admit, "a small request overtook the waiting large one");
    s2.wait(1, -1);
}

// Idle states stay resident and are reused; a new size evicts idle states
// (largest first) when the budget requires it.
void eviction() {
    g_case = "eviction";
    const size_t budget = 4096 + 1024;
    auto* be = new FakeBackend(budget);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), budget, 2);
    Server& s = *s_owner;
    s.submit(make(1, 12, 1.0));
    s.wait(1, -1);
    s.submit(make(2, 12, 1.0));  // reuses the idle 12-qubit state
    s.wait(2, -1);
    require(be->created_.load() == 1, "idle state reused");
    s.submit(make(3, 10, 1.0));  // fits beside it
    s.wait(3, -1);
    require(s.stats().evictions == 0, "no eviction needed yet");
    s.submit(make(4, 11, 1.0));  // 2048 more: the idle 12-qubit state goes
    s.wait(4, -1);
    const Stats st = s.stats();
    require(st.evictions == 1, "one eviction");
    require(st.resident_bytes <= budget && st.peak_resident_bytes <= budget && !be->violation_, "within the budget");
}

// Cancelling a queued request drops it; cancelling a running one stops it;
// a finished one is reported as not cancelled.
void cancellation() {
    g_case = "cancellation";
    auto* be = new FakeBackend(1 << 20);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 1 << 20, 1);
    Server& s = *s_owner;
    s.submit(make(1, 8, 200.0));  // runs for 200 ms unless cancelled
    s.submit(make(2, 8, 1.0));    // queued behind it
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    require(s.cancel(2), "queued request cancelled");
    auto r2 = s.wait(2, 1000);
    require(r2->status == AQS_ERR_RUNTIME && r2->error == "request cancelled", "queued: cancelled status");
    const double t0 = now_s();
    require(s.cancel(1), "running request cancelled");
    auto r1 = s.wait(1, 2000);
    require(r1->status == AQS_ERR_RUNTIME && r1->error == "request cancelled", "running: cancelled status");
    require(now_s() - t0 < 0.15, "running request stopped early");
    s.submit(make(3, 8, 1.0));
    auto r3 = s.wait(3, -1);
    require(r3->status == AQS_OK, "server keeps working");
    require(raises(AQS_ERR_OUT_OF_RANGE, [&] { s.cancel(3); }), "collected request is unknown");
    s.submit(make(4, 8, 1.0));
    std::this_thread::sleep_for(std::chrono::milliseconds(30));
    require(!s.cancel(4), "finished request is not cancelled");
    require(s.wait(4, -1)->status == AQS_OK, "finished result kept");
    require(s.stats().cancelled == 2, "cancel count");
}

// Failures inside the backend become the request's status (the worker thread
// survives), and a wait with a timeout reports an unfinished request.
void failures() {
    g_case = "failures";
    auto* be = new FakeBackend(1 << 20);
    be->fail_qubits_ = 9;
    be->oom_qubits_ = 7;
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 1 << 20, 2);
    Server& s = *s_owner;
    s.submit(make(1, 9, 1.0));
    s.submit(make(2, 7, 1.0));
    s.submit(make(3, 8, 100.0));
    auto r1 = s.wait(1, -1);
    require(r1->status == AQS_ERR_RUNTIME && r1->error == "backend failure", "exception becomes a status");
    auto r2 = s.wait(2, -1);
    require(r2->status == AQS_ERR_NO_MEMORY, "bad_alloc becomes a status");
    require(raises(AQS_ERR_RUNTIME, [&] { s.wait(3, 1); }), "timeout");
    require(s.wait(3, -1)->status == AQS_OK, "still collectable after a timeout");
}

// Many threads submit, wait, wait twice and cancel at once (the race the
// server's wait must survive: the map grows while a waiter sleeps).
void concurrent() {
    g_case = "concurrent";
    auto* be = new FakeBackend(64 * 1024);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 64 * 1024, 4);
    Server& s = *s_owner;
    s.set_starvation_limit(0.005);
    std::atomic<int> ok{0}, cancelled{0}, collected_twice{0}, errors{0};
    std::vector<std::thread> threads;
    for (int t = 0; t < 6; ++t)
        threads.emplace_back([&, t] {
            std::mt19937 rng(1234 + t);
            for (int k = 0; k < 40; ++k) {
                const uint64_t id = (uint64_t)t * 1000 + k;
                s.submit(make(id, 6 + (int)(rng() % 10), 0.2 + 0.1 * (rng() % 5)));
                if (k % 7 == 3) s.cancel(id);
                if (k % 5 == 0) {
                    // two waiters on one id: one result, one "already collected"
                    std::atomic<int> got{0}, lost{0};
                    auto waiter = [&] {
                        try {
                            s.wait(id, -1);
                            got.fetch_add(1);
                        } catch (const ServeError& e) {
                            if (e.code == AQS_ERR_OUT_OF_RANGE) lost.fetch_add(1);
                        }
                    };
                    std::thread a(waiter), b(waiter);
                    a.join();
                    b.join();
                    if (got.load() == 1 && lost.load() == 1) collected_twice.fetch_add(1);
                    else errors.fetch_add(1);
                    continue;
                }
                auto r = s.wait(id, -1);
                if (r->status == AQS_OK) ok.fetch_add(1);
                else if (r->error == "request cancelled") cancelled.fetch_add(1);
                else errors.fetch_add(1);
            }
        });
    for (auto& th : threads) th.join();
    require(errors.load() == 0, "no unexpected outcome");
    require(collected_twice.load() == 6 * 8, "each double wait hands the result over once");
    require(ok.load() + cancelled.load() == 6 * 32, "every other request accounted for");
    const Stats st = s.stats();
    require(st.pending == 0 && st.running == 0, "drained");
    require(!be->violation_, "device-side invariants");
}

FakeBackend* group_backend(size_t budget, int qubits, size_t arena, size_t item) {
    auto* be = new FakeBackend(budget);
    be->group_qubits_ = qubits;
    be->arena_ = arena;
    be->item_bytes_ = item;
    return be;
}

size_t group_of(const std::vector<std::vector<uint64_t>>& groups, uint64_t id) {
    for (size_t g = 0; g < groups.size(); ++g)
        if (std::find(groups[g].begin(), groups[g].end(), id) != groups[g].end()) return g;
    return (size_t)-1;
}

// Small requests queued behind a busy worker run as groups: in arrival
// order, at most the request cap and the arena's bytes per group; larger
// requests, and small ones the backend cannot group, run alone.
void grouping() {
    g_case = "grouping";
    auto* be = group_backend(1 << 20, 9, 4096, 1000);  // four members fill the arena
    be->unfit_qubits_ = 7;
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 1 << 20, 1);
    Server& s = *s_owner;
    s.submit(make(1, 12, 30.0));  // occupies the one worker
    for (uint64_t id = 2; id < 12; ++id) s.submit(make(id, id == 5 ? 11 : id == 8 ? 7 : 8, 1.0));
    for (uint64_t id = 1; id < 12; ++id) {
        auto r = s.wait(id, -1);
        require(r->status == AQS_OK, "status");
        require(r->indices.size() == 4 && r->indices[1] == 1 + id, "indices belong to the request");
    }
    const auto groups = be->groups();
    // 8 groupable requests (2,3,4,6,7,9,10,11) in groups of at most 4
    require(groups.size() == 2, "two groups");
    require(groups[0] == std::vector<uint64_t>({2, 3, 4, 6}), "first group: arrival order, bytes cap");
    require(groups[1] == std::vector<uint64_t>({7, 9, 10, 11}), "second group");
    require(group_of(groups, 5) == (size_t)-1 && group_of(groups, 8) == (size_t)-1, "large and unfit run alone");
    const Stats st = s.stats();
    require(st.groups == 2 && st.grouped == 8, "group counts");
    require(st.completed == 11 && st.admitted == 11, "counts");
    require(!be->violation_, "device-side invariants");

    g_case = "group request cap";
    auto* be2 = group_backend(1 << 20, 9, 1 << 16, 1000);
    auto s2_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be2), 1 << 20, 1);
    Server& s2 = *s2_owner;
    s2.set_grouping(9, 3);
    s2.submit(make(1, 12, 30.0));
    for (uint64_t id = 2; id < 9; ++id) s2.submit(make(id, 8, 1.0));
    for (uint64_t id = 1; id < 9; ++id) require(s2.wait(id, -1)->status == AQS_OK, "status");
    const auto g2 = be2->groups();
    require(g2.size() == 3 && g2[0].size() == 3 && g2[1].size() == 3 && g2[2].size() == 1, "at most 3 per group");

    g_case = "grouping off";
    auto* be3 = group_backend(1 << 20, 9, 1 << 16, 1000);
    auto s3_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be3), 1 << 20, 1);
    Server& s3 = *s3_owner;
    s3.set_grouping(0, 64);
    s3.submit(make(1, 12, 10.0));
    for (uint64_t id = 2; id < 6; ++id) s3.submit(make(id, 8, 1.0));
    for (uint64_t id = 1; id < 6; ++id) require(s3.wait(id, -1)->status == AQS_OK, "status");
    require(be3->groups().empty() && s3.stats().groups == 0, "no groups when off");

    g_case = "arena over budget";
    auto* be4 = group_backend(2048, 9, 4096, 100);  // the arena never fits: coalescing stays off
    auto s4_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be4), 2048, 1);
    Server& s4 = *s4_owner;
    s4.submit(make(1, 10, 10.0));
    for (uint64_t id = 2; id < 5; ++id) s4.submit(make(id, 8, 1.0));
    for (uint64_t id = 1; id < 5; ++id) require(s4.wait(id, -1)->status == AQS_OK, "status");
    require(be4->groups().empty(), "no groups without room for the arena");
    require(!be4->violation_, "device-side invariants");
}

// Collecting a group stops at a starved request: a later small request does
// not start before it.
void group_starvation() {
    g_case = "group starvation";
    const size_t arena = 1024;
    const size_t budget = 4096 + 1024;  // the 12-qubit state plus one arena (or one 10-qubit state)
    auto* be = group_backend(budget, 9, arena, 100);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), budget, 2);
    Server& s = *s_owner;
    s.set_starvation_limit(0.02);
    s.submit(make(1, 12, 100.0));  // worker 0, 4096 bytes
    s.submit(make(2, 10, 60.0));   // worker 1, 1024 bytes: the budget is full
    s.submit(make(3, 8, 1.0));     // small
    s.submit(make(4, 12, 1.0));    // large: needs request 1's room
    s.submit(make(5, 8, 1.0));     // small, behind the large one
    auto r4 = s.wait(4, -1);
    auto r5 = s.wait(5, -1);
    auto r3 = s.wait(3, -1);
    require(r3->status == AQS_OK && r4->status == AQS_OK && r5->status == AQS_OK, "status");
    require(r5->t_admit >= r4->t_admit, "a group does not overtake a starved request");
    require(group_of(be->groups(), 3) != group_of(be->groups(), 5), "not coalesced past it");
    s.wait(1, -1);
    s.wait(2, -1);
    require(!be->violation_, "device-side invariants");
}

// A member cancelled while its group runs completes as cancelled; the other
// members keep their results.  A failing group fails every member.
void group_cancel_and_failure() {
    g_case = "group cancel";
    auto* be = group_backend(1 << 20, 9, 1 << 16, 100);
    be->fail_qubits_ = 6;
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 1 << 20, 1);
    Server& s = *s_owner;
    s.submit(make(1, 12, 20.0));
    for (uint64_t id = 2; id < 5; ++id) s.submit(make(id, 8, 60.0));  // one group of three, 60 ms
    std::this_thread::sleep_for(std::chrono::milliseconds(40));       // the group is running
    require(s.cancel(3), "running member cancelled");
    auto r3 = s.wait(3, -1);
    require(r3->status == AQS_ERR_RUNTIME && r3->error == "request cancelled", "cancelled member");
    require(s.wait(2, -1)->status == AQS_OK && s.wait(4, -1)->status == AQS_OK, "other members finish");
    require(s.wait(1, -1)->status == AQS_OK, "first request");

    g_case = "group failure";
    s.submit(make(10, 12, 20.0));
    s.submit(make(11, 8, 1.0));
    s.submit(make(12, 6, 1.0));  // the backend fails any group holding a 6-qubit member
    s.submit(make(13, 8, 1.0));
    require(s.wait(10, -1)->status == AQS_OK, "alone");
    for (uint64_t id = 11; id < 14; ++id) {
        auto r = s.wait(id, -1);
        require(r->status == AQS_ERR_RUNTIME && r->error == "group failure", "every member fails");
    }
    s.submit(make(14, 8, 1.0));
    require(s.wait(14, -1)->status == AQS_OK, "server keeps working");
    require(s.stats().cancelled == 1, "cancel count");
    require(!be->violation_, "device-side invariants");
}

// The concurrent case with coalescing on: submits, double waits and cancels
// from many threads while groups form.
void concurrent_grouped() {
    g_case = "concurrent grouped";
    auto* be = group_backend(64 * 1024, 10, 8 * 1024, 1000);
    auto s_owner = std::make_unique<Server>(std::unique_ptr<Backend>(be), 64 * 1024, 3);
    Server& s = *s_owner;
    s.set_starvation_limit(0.005);
    std::atomic<int> ok{0}, cancelled{0}, errors{0};
    std::vector<std::thread> threads;
    for (int t = 0; t < 6; ++t)
        threads.emplace_back([&, t] {
            std::mt19937 rng(777 + t);
            for (int k = 0; k < 40; ++k) {
                const uint64_t id = (uint64_t)t * 1000 + k;
                s.submit(make(id, 6 + (int)(rng() % 10), 0.2 + 0.1 * (rng() % 5)));
                if (k % 7 == 3) s.cancel(id);
                auto r = s.wait(id, -1);
                if (r->status == AQS_OK) ok.fetch_add(1);
                else if (r->error == "request cancelled") cancelled.fetch_add(1);
                else errors.fetch_add(1);
            }
        });
    for (auto& th : threads) th.join();
    require(errors.load() == 0, "no unexpected outcome");
    require(ok.load() + cancelled.load() == 6 * 40, "every request accounted for");
    const Stats st = s.stats();
    require(st.pending == 0 && st.running == 0, "drained");
    require(st.groups > 0, "groups formed");
    require(!be->violation_, "device-side invariants");
}

}  // namespace

int main() {
    basic();
    budget_admission();
    starvation();
    eviction();
    cancellation();
    failures();
    concurrent();
    grouping();
    group_starvation();
    group_cancel_and_failure();
    concurrent_grouped();
    std::printf("%d checks, ALL PASSED\n", g_checks);
    return 0;
}
