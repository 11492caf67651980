// Scheduling core of the request server: queue, admission by device-memory
// footprint, resident-state pool, starvation limit, cancellation, waiting.
//
// Host-only (no HIP): the device work goes through a Backend, so the policy
// and the thread handling can be tested on a CPU with a stand-in backend
// (tests/cpp/serve_core_test.cpp, also built under ThreadSanitizer).
//
// The reference has no serving layer (a caller runs one QSimulator to
// completion, src/quantum.cpp:277-291, 467-501); this is new.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "aqs_hip.h"

namespace aqs_serve {

double now_s();

// Raised inside the core and the backends; the C boundary turns it into an
// aqs_status plus message.
struct ServeError {
    aqs_status code;
    std::string msg;
};

struct Request {
    uint64_t id = 0;
    int n = 0;
    std::vector<aqs_op> ops;               // MATK ops point into mats
    std::vector<std::vector<double>> mats;
    std::vector<double> u;                 // uniforms, one per shot
    uint32_t shots = 0;
    int keep_state = 0;
    size_t footprint = 0;                  // device bytes of its state
    size_t group_cost = 0;                 // its arena bytes in a group (Backend::group_bytes)
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::atomic<bool> cancel{false};       // set by cancel() while it runs

    // results (written by the worker before `done`)
    int status = AQS_OK;
    std::string error;
    std::vector<uint64_t> indices;
    std::vector<float> state;              // complex64 interleaved, when keep_state
    double norm2 = 0;
    size_t passes = 0;
    bool done = false;                     // guarded by the server mutex
};

// Pool key of the group arena (no request has 0 qubits).
constexpr int kGroupKey = 0;

// Device side of the server.  create_state / destroy_state / run /
// run_group are called from worker threads; run must check `r.cancel`
// between its stages and raise ServeError{AQS_ERR_RUNTIME, "request
// cancelled"} when it is set.
//
// Coalescing: a backend whose group_max_qubits() is above 0 can run several
// requests of at most that many qubits together (one batched launch) inside
// a group arena, a device buffer of footprint(kGroupKey) bytes that lives in
// the resident-state pool under the key kGroupKey.  group_bytes(r) (asked
// once, at submit) is an upper bound of the arena bytes r takes in a group,
// above footprint(kGroupKey) for a request that cannot run in one; a group's
// members sum to at most footprint(kGroupKey).  run_group fills every member's results or
// raises (then every member fails with that error).
class Backend {
   public:
    virtual ~Backend() = default;
    virtual void worker_start(int worker) { (void)worker; }
    // device bytes a state of n qubits (n = kGroupKey: the group arena)
    // holds while resident
    virtual size_t footprint(int n) const = 0;
    virtual void* create_state(int n) = 0;
    virtual void destroy_state(void* st) = 0;
    virtual void run(Request& r, void* st, int worker) = 0;
    virtual int group_max_qubits() const { return 0; }
    virtual size_t group_bytes(const Request& r) const {
        (void)r;
        return 0;
    }
    virtual void run_group(std::vector<Request*>& rs, void* arena, int worker) {
        (void)rs;
        (void)arena;
        (void)worker;
        throw ServeError{AQS_ERR_RUNTIME, "backend runs no request groups"};
    }
};

struct Stats {
    uint64_t admitted = 0, completed = 0, rejected = 0, evictions = 0, pending = 0, cancelled = 0;
    int32_t running = 0, max_concurrent = 0;
    uint64_t resident_bytes = 0, peak_resident_bytes = 0, budget_bytes = 0;
    uint64_t groups = 0, grouped = 0;  // group launches, requests run in them
};

// W worker threads each run one admitted request (or group) at a time.
//   - Admission: a request starts when a worker is free and the footprints of
//     the running requests plus its own fit the budget; idle resident states
//     are evicted (largest first) when a new state needs their memory.
//   - Order: arrival order, except that a later request may start ahead of
//     an earlier one that does not fit yet, unless the earlier one has waited
//     longer than the starvation limit (then nothing overtakes it).
//   - A request that can never fit the budget is rejected at submit.
//   - Idle states stay resident per qubit count and are reused.
//   - Coalescing (when the backend offers it and the group arena fits the
//     budget): a worker that takes a request of at most the group limit of
//     qubits also takes the later pending requests within that limit, in
//     arrival order, up to the request cap and the arena's bytes, and runs
//     them as one group.  Collection stops at a request that has waited
//     longer than the starvation limit, so a group never overtakes it.
class Server {
   public:
    Server(std::unique_ptr<Backend> backend, size_t budget_bytes, int workers);
    ~Server();  // finishes pending and running work, then frees every state
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    void set_starvation_limit(double seconds);
    // Requests of at most max_qubits qubits (0: none) are coalesced, at most
    // max_requests per group.  Limited by the backend's group_max_qubits().
    void set_grouping(int max_qubits, int max_requests);
    // Queues r (its footprint is set here).  ServeError on a duplicate id or a
    // request larger than the budget.
    void submit(std::shared_ptr<Request> r);
    // Waits for request `id` (timeout_ms < 0: no limit) and hands it over;
    // each result is collected once (a second waiter gets AQS_ERR_OUT_OF_RANGE).
    std::shared_ptr<Request> wait(uint64_t id, int timeout_ms);
    // A pending request is dropped and completes as cancelled; a running one
    // is told to stop at its next stage.  Returns false when it had already
    // finished.  ServeError for an unknown id.
    bool cancel(uint64_t id);
    Stats stats();

   private:
    std::unique_ptr<Backend> backend_;
    size_t budget_ = 0;
    double starvation_s_ = 0.25;

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    bool stopping_ = false;

    std::deque<std::shared_ptr<Request>> pending_;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all_;  // submitted and not yet collected
    // per worker: the assigned request, or a group; its admission footprint
    std::vector<std::vector<std::shared_ptr<Request>>> slot_;
    std::vector<bool> slot_group_;
    std::vector<size_t> slot_fp_;
    std::vector<std::thread> workers_;

    int group_qubits_ = 0;       // coalescing limit (0: off)
    int group_requests_ = 64;    // members per group
    size_t arena_bytes_ = 0;     // footprint(kGroupKey)

    std::map<int, std::vector<void*>> idle_;  // resident idle states by qubit count
    size_t resident_ = 0;                     // bytes of every resident state
    size_t in_use_ = 0;                       // footprints of the running requests
    Stats st_;

    bool groupable_locked(const Request& r) const;
    void dispatch_locked();
    void* acquire_state(int n);
    void release_state(int n, void* st);
    void finish_locked(Request& r);
    void run_single(Request& r, int w);
    void run_grouped(std::vector<std::shared_ptr<Request>>& job, int w);
    void worker_main(int w);
};

}  // namespace aqs_serve
