// Continuous-batching request server (host side).
//
// The reference has no serving layer: a caller builds a QSimulator per
// circuit and runs it to completion (src/quantum.cpp:277-291, 467-501).
// Here a server takes circuits of mixed size as they arrive and keeps the
// device busy with several of them at once:
//
//   - each request is a gate list, a qubit count, caller-supplied uniforms
//     and a shot count; the result is the sampled basis indices;
//   - W worker threads each own a HIP stream (default priority); a worker
//     runs one request at a time: compile() fusion on the host, then
//     |0..0>, the fused passes and sampling on its stream;
//   - admission is by HBM footprint: a request starts only when a worker is
//     free and the bytes of the resident states plus its own fit the
//     budget; idle states stay resident in a per-size pool and are reused,
//     and freed only when an admission needs their memory;
//   - requests are admitted in arrival order; a later request may start
//     ahead of an earlier one that does not fit yet, unless the earlier one
//     has waited longer than the starvation limit;
//   - a request that can never fit the budget is rejected at submit.
// Gate kernels are deterministic, so a request's result does not depend on
// what else runs beside it.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

using namespace aqs_hip;

namespace {

using Clock = std::chrono::steady_clock;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

struct Request {
    uint64_t id = 0;
    int n = 0;
    std::vector<aqs_op> ops;
    std::vector<std::vector<double>> mats;
    std::vector<double> u;
    uint32_t shots = 0;
    int keep_state = 0;
    size_t footprint = 0;
    // results
    int status = AQS_OK;
    std::string error;
    std::vector<uint64_t> indices;
    std::vector<float> state;  // complex64 interleaved when keep_state
    double t_submit = 0, t_admit = 0, t_done = 0;
    double norm2 = 0;
    size_t passes = 0;
    bool done = false;
};

size_t state_footprint(int n) {
    const uint64_t N = 1ull << n;
    const size_t blocks = (size_t)((N + 1023) / 1024);
    return (size_t)(N * 8) + std::max<size_t>(65536, (blocks + 1024) * 32) + 65536;
}

}  // namespace

struct aqs_server_s {
    size_t budget = 0;
    int fuse_qubits = 0;  // 0: aqs_fuse_auto
    double starvation_s = 0.25;
    int device = 0;

    std::mutex mu;
    std::condition_variable cv_work;   // workers wait for assignments
    std::condition_variable cv_done;   // waiters for results
    bool stopping = false;

    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<std::shared_ptr<Request>> slot;  // per worker: assigned request
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;

    // resident states: idle pool per size, bytes of everything resident
    std::map<int, std::vector<aqs_state>> idle;
    size_t resident = 0;
    size_t in_use = 0;

    // statistics
    size_t peak_resident = 0;
    int max_concurrent = 0;
    int running = 0;
    uint64_t admitted = 0, rejected = 0, completed = 0, evictions = 0;
};

namespace {

// Called with s->mu held: assign pending requests to idle workers.
void dispatch_locked(aqs_server s) {
    const double t = now_s();
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        int w = -1;
        for (size_t i = 0; i < s->slot.size(); ++i)
            if (!s->slot[i]) {
                w = (int)i;
                break;
            }
        if (w < 0) return;
        auto& r = *it;
        // memory: reuse an idle state of this size, else make room
        bool reuse = !s->idle[r->n].empty();
        size_t need = reuse ? 0 : r->footprint;
        size_t idle_bytes = s->resident - s->in_use;
        bool fits = s->in_use + r->footprint <= s->budget && (s->resident + need <= s->budget || idle_bytes >= need);
        if (!fits) {
            // head-of-line: an old request blocks later ones (no starvation)
            if (t - r->t_submit > s->starvation_s) return;
            ++it;
            continue;
        }
        s->slot[w] = r;
        s->in_use += r->footprint;
        r->t_admit = t;
        ++s->admitted;
        ++s->running;
        s->max_concurrent = std::max(s->max_concurrent, s->running);
        it = s->pending.erase(it);
        s->cv_work.notify_all();
    }
}

// Take a state of n qubits for a worker: idle pool first, else evict idle
// states until the new one fits, then allocate.
aqs_state acquire_state(aqs_server s, int n) {
    {
        std::lock_guard<std::mutex> lock(s->mu);
        auto& v = s->idle[n];
        if (!v.empty()) {
            aqs_state st = v.back();
            v.pop_back();
            return st;
        }
        const size_t need = state_footprint(n);
        while (s->resident + need > s->budget) {
            aqs_state victim = nullptr;
            int vn = -1;
            for (auto& kv : s->idle)
                if (!kv.second.empty()) {  // evict the largest idle states first
                    vn = kv.first;
                }
            if (vn < 0) break;
            victim = s->idle[vn].back();
            s->idle[vn].pop_back();
            s->resident -= state_footprint(vn);
            ++s->evictions;
            aqs_state_destroy(victim);
        }
        s->resident += need;
        s->peak_resident = std::max(s->peak_resident, s->resident);
    }
    aqs_state st = nullptr;
    aqs_status rc = aqs_state_create(n, AQS_C64, &st);
    if (rc != AQS_OK) {
        std::lock_guard<std::mutex> lock(s->mu);
        s->resident -= state_footprint(n);
        fail(rc, aqs_last_error());
    }
    return st;
}

void run_request(aqs_server s, Request& r, hipStream_t stream) {
    aqs_schedule sch = nullptr;
    aqs_state st = nullptr;
    auto ck = [&](aqs_status rc) {
        if (rc != AQS_OK) fail(rc, aqs_last_error());
    };
    try {
        if (s->fuse_qubits == 0)
            ck(aqs_fuse_auto(r.ops.data(), r.ops.size(), r.n, &sch));
        else
            ck(aqs_fuse(r.ops.data(), r.ops.size(), r.n, s->fuse_qubits, &sch));
        ck(aqs_schedule_info(sch, &r.passes, nullptr));
        st = acquire_state(s, r.n);
        ck(aqs_state_set_stream(st, stream));
        ck(aqs_state_set_basis(st, 0));
        ck(aqs_apply_schedule(st, sch));
        r.indices.resize(r.shots);
        if (r.shots) ck(aqs_sample(st, r.u.data(), r.shots, r.indices.data()));
        ck(aqs_norm2(st, &r.norm2));
        if (r.keep_state) {
            r.state.resize(2ull << r.n);
            ck(aqs_state_download(st, r.state.data()));
        }
    } catch (const Error& e) {
        r.status = e.code;
        r.error = e.msg;
    }
    if (sch) aqs_schedule_destroy(sch);
    if (st) {
        std::lock_guard<std::mutex> lock(s->mu);
        s->idle[r.n].push_back(st);
    }
}

void worker_main(aqs_server s, int w) {
    (void)hipSetDevice(s->device);
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            s->cv_work.wait(lock, [&] { return s->stopping || s->slot[w]; });
            if (!s->slot[w]) return;  // stopping and nothing assigned
            r = s->slot[w];
        }
        run_request(s, *r, s->streams[w]);
        {
            std::lock_guard<std::mutex> lock(s->mu);
            r->t_done = now_s();
            r->done = true;
            s->slot[w] = nullptr;
            s->in_use -= r->footprint;
            --s->running;
            ++s->completed;
            dispatch_locked(s);
        }
        s->cv_done.notify_all();
    }
}

aqs_server check_server(aqs_server s) {
    if (!s) fail(AQS_ERR_INVALID_ARGUMENT, "null server handle");
    return s;
}

}  // namespace

extern "C" {

aqs_status aqs_server_create(size_t hbm_budget_bytes, int workers, int fuse_qubits, aqs_server* out) {
    return guard([&] {
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output handle");
        *out = nullptr;
        if (workers < 1 || workers > 16) fail(AQS_ERR_INVALID_ARGUMENT, "workers must be 1..16");
        if (fuse_qubits < 0 || fuse_qubits > 5) fail(AQS_ERR_INVALID_ARGUMENT, "fuse_qubits must be 0..5");
        auto* s = new aqs_server_s();
        s->budget = hbm_budget_bytes;
        s->fuse_qubits = fuse_qubits;
        s->device = current_device();
        s->slot.assign(workers, nullptr);
        try {
            for (int w = 0; w < workers; ++w) {
                hipStream_t st;
                AQS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
                s->streams.push_back(st);
            }
            for (int w = 0; w < workers; ++w) s->workers.emplace_back(worker_main, s, w);
        } catch (...) {
            {
                std::lock_guard<std::mutex> lock(s->mu);
                s->stopping = true;
            }
            s->cv_work.notify_all();
            for (auto& t : s->workers) t.join();
            for (auto st : s->streams) (void)hipStreamDestroy(st);
            delete s;
            throw;
        }
        *out = s;
    });
}

aqs_status aqs_server_set_starvation_limit(aqs_server s, double seconds) {
    return guard([&] {
        check_server(s);
        std::lock_guard<std::mutex> lock(s->mu);
        s->starvation_s = seconds;
    });
}

aqs_status aqs_server_submit(aqs_server s, uint64_t id, int nqubits, const aqs_op* ops, size_t nops,
                             const double* uniforms, uint32_t shots, int keep_state) {
    return guard([&] {
        check_server(s);
        if (nqubits < 1 || nqubits > AQS_MAX_QUBITS) fail(AQS_ERR_INVALID_ARGUMENT, "bad qubit count");
        if (nops && !ops) fail(AQS_ERR_INVALID_ARGUMENT, "null op array");
        if (shots && !uniforms) fail(AQS_ERR_INVALID_ARGUMENT, "null uniforms");
        auto r = std::make_shared<Request>();
        r->id = id;
        r->n = nqubits;
        r->shots = shots;
        r->keep_state = keep_state;
        r->footprint = state_footprint(nqubits);
        if (r->footprint > s->budget)
            fail(AQS_ERR_INVALID_ARGUMENT, "request needs more device memory than the server budget");
        r->ops.assign(ops, ops + nops);
        for (auto& op : r->ops) {
            validate_op(op, nqubits);
            if (op.kind == AQS_OP_MATK) {
                const size_t K = (size_t)1 << op.ntargets;
                r->mats.emplace_back(op.u, op.u + 2 * K * K);
                op.u = r->mats.back().data();
            }
        }
        // matrices were moved into r->mats; re-point after the vector settled
        size_t mi = 0;
        for (auto& op : r->ops)
            if (op.kind == AQS_OP_MATK) op.u = r->mats[mi++].data();
        r->u.assign(uniforms, uniforms + shots);
        r->t_submit = now_s();
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->all.count(id)) {
            ++s->rejected;
            fail(AQS_ERR_INVALID_ARGUMENT, "duplicate request id");
        }
        s->all[id] = r;
        s->pending.push_back(r);
        dispatch_locked(s);
    });
}

aqs_status aqs_server_wait(aqs_server s, uint64_t id, int timeout_ms, aqs_request_result* res, uint64_t* indices,
                           void* state_host) {
    return guard([&] {
        check_server(s);
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            auto it = s->all.find(id);
            if (it == s->all.end()) fail(AQS_ERR_OUT_OF_RANGE, "unknown request id");
            r = it->second;
            auto pred = [&] { return r->done; };
            if (timeout_ms < 0)
                s->cv_done.wait(lock, pred);
            else if (!s->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), pred))
                fail(AQS_ERR_RUNTIME, "request not finished within the timeout");
            s->all.erase(it);
        }
        if (res) {
            res->id = r->id;
            res->status = r->status;
            res->nqubits = r->n;
            res->shots = r->shots;
            res->passes = r->passes;
            res->t_submit = r->t_submit;
            res->t_admit = r->t_admit;
            res->t_done = r->t_done;
            res->norm2 = r->norm2;
        }
        if (indices && r->status == AQS_OK) std::copy(r->indices.begin(), r->indices.end(), indices);
        if (state_host && r->keep_state && r->status == AQS_OK)
            std::copy(r->state.begin(), r->state.end(), (float*)state_host);
        if (r->status != AQS_OK) fail((aqs_status)r->status, r->error);
    });
}

aqs_status aqs_server_stats(aqs_server s, aqs_server_stats_t* out) {
    return guard([&] {
        check_server(s);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!out) return;
        out->admitted = s->admitted;
        out->completed = s->completed;
        out->rejected = s->rejected;
        out->evictions = s->evictions;
        out->pending = s->pending.size();
        out->running = s->running;
        out->max_concurrent = s->max_concurrent;
        out->resident_bytes = s->resident;
        out->peak_resident_bytes = s->peak_resident;
        out->budget_bytes = s->budget;
    });
}

aqs_status aqs_server_destroy(aqs_server s) {
    return guard([&] {
        if (!s) return;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            // finish admitted and pending work first
            s->cv_done.wait(lock, [&] { return s->pending.empty() && s->running == 0; });
            s->stopping = true;
        }
        s->cv_work.notify_all();
        for (auto& t : s->workers) t.join();
        for (auto& kv : s->idle)
            for (aqs_state st : kv.second) aqs_state_destroy(st);
        for (auto st : s->streams) (void)hipStreamDestroy(st);
        delete s;
    });
}

}  // extern "C"
