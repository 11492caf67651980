// Request server: the C ABI and the HIP backend of the scheduling core
// (serve_core.h).
//
// A request is a gate list, a qubit count, caller-supplied uniforms and a
// shot count; its result is the sampled basis indices.  Each worker thread
// owns a HIP stream (default priority) and runs one request at a time:
// compile() fusion on the host (aqs_fuse_auto, or aqs_fuse of a fixed
// width), then |0..0>, the fused passes and sampling on its stream.  Gate
// kernels are deterministic, so a request's result does not depend on what
// else runs beside it.
//
// Requests of at most 16 qubits with no MATK op are coalesced: a worker runs
// a group of them as one batched launch (run_group: each state in LDS tiles,
// no host fusion), then samples each state, inside a group arena of
// kArenaBytes that stays resident in the state pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <memory>

#include "batch.h"
#include "internal.h"
#include "serve_core.h"

using namespace aqs_hip;
using aqs_serve::Request;
using aqs_serve::ServeError;

namespace {

// Device bytes of a resident state: amplitudes, reduction / sampling scratch,
// staging.
size_t state_footprint(int n) {
    const uint64_t N = 1ull << n;
    const size_t blocks = (size_t)((N + 1023) / 1024);
    return (size_t)(N * 8) + std::max<size_t>(65536, (blocks + 1024) * 32) + 65536;
}

// Device bytes of a group arena: 64 states of 16 qubits with 1500-gate
// circuits and 1000 shots each fit.
constexpr size_t kArenaBytes = 96ull << 20;

// What the state pool holds for the backend: a request's state or a group arena.
struct Resident {
    aqs_state state = nullptr;
    void* arena = nullptr;
};

class HipBackend final : public aqs_serve::Backend {
   public:
    HipBackend(int device, int fuse_qubits, int workers) : device_(device), fuse_qubits_(fuse_qubits) {
        AQS_HIP(hipSetDevice(device_));
        for (int w = 0; w < workers; ++w) {
            hipStream_t st;
            hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (e != hipSuccess) {
                release_streams();
                hip_check(e, "hipStreamCreateWithFlags");
            }
            streams_.push_back(st);
        }
    }
    ~HipBackend() override { release_streams(); }

    void worker_start(int) override { (void)hipSetDevice(device_); }
    size_t footprint(int n) const override { return n == aqs_serve::kGroupKey ? kArenaBytes : state_footprint(n); }

    void* create_state(int n) override {
        auto res = std::make_unique<Resident>();
        if (n == aqs_serve::kGroupKey) {
            const hipError_t e = hipMalloc(&res->arena, kArenaBytes);
            if (e != hipSuccess) throw ServeError{AQS_ERR_NO_MEMORY, "group arena allocation failed"};
            return res.release();
        }
        const aqs_status rc = aqs_state_create(n, AQS_C64, &res->state);
        if (rc != AQS_OK) throw ServeError{rc, aqs_last_error()};
        return res.release();
    }
    void destroy_state(void* p) override {
        std::unique_ptr<Resident> res((Resident*)p);
        if (res->state) aqs_state_destroy(res->state);
        if (res->arena) (void)hipFree(res->arena);
    }

    int group_max_qubits() const override { return kGroupMaxQubits; }
    size_t group_bytes(const Request& r) const override {
        for (const auto& op : r.ops)
            if (op.kind == AQS_OP_MATK) return std::numeric_limits<size_t>::max();
        return group_item_bytes(r.n, r.ops.size(), r.shots);
    }

    void run_group(std::vector<Request*>& rs, void* arena, int worker) override {
        std::vector<GroupItem> items(rs.size());
        for (size_t k = 0; k < rs.size(); ++k) {
            Request& r = *rs[k];
            r.indices.resize(r.shots);
            if (r.keep_state) r.state.resize(2ull << r.n);
            GroupItem& it = items[k];
            it.n = r.n;
            it.ops = r.ops.data();
            it.nops = r.ops.size();
            it.u = r.u.data();
            it.shots = r.shots;
            it.indices = r.indices.data();
            it.norm2 = &r.norm2;
            it.state = r.keep_state ? r.state.data() : nullptr;
        }
        const aqs_status rc = guard([&] {
            aqs_hip::run_group(items.data(), (int)items.size(), ((Resident*)arena)->arena, kArenaBytes,
                               streams_[(size_t)worker]);
        });
        if (rc != AQS_OK) throw ServeError{rc, aqs_last_error()};
        for (size_t k = 0; k < rs.size(); ++k) rs[k]->passes = items[k].passes;
    }

    void run(Request& r, void* state, int worker) override {
        aqs_state st = ((Resident*)state)->state;
        auto ck = [](aqs_status rc) {
            if (rc != AQS_OK) throw ServeError{rc, aqs_last_error()};
        };
        auto stop_if_cancelled = [&] {
            if (r.cancel.load()) throw ServeError{AQS_ERR_RUNTIME, "request cancelled"};
        };
        aqs_schedule sch = nullptr;
        try {
            if (fuse_qubits_ == 0)
                ck(aqs_fuse_auto(r.ops.data(), r.ops.size(), r.n, &sch));
            else
                ck(aqs_fuse(r.ops.data(), r.ops.size(), r.n, fuse_qubits_, &sch));
            ck(aqs_schedule_info(sch, &r.passes, nullptr));
            stop_if_cancelled();
            ck(aqs_state_set_stream(st, streams_[(size_t)worker]));
            ck(aqs_state_set_basis(st, 0));
            ck(aqs_apply_schedule(st, sch));
            stop_if_cancelled();
            r.indices.resize(r.shots);
            if (r.shots) ck(aqs_sample(st, r.u.data(), r.shots, r.indices.data()));
            ck(aqs_norm2(st, &r.norm2));
            if (r.keep_state) {
                r.state.resize(2ull << r.n);
                ck(aqs_state_download(st, r.state.data()));
            }
        } catch (...) {
            if (sch) aqs_schedule_destroy(sch);
            throw;
        }
        aqs_schedule_destroy(sch);
    }

   private:
    int device_;
    int fuse_qubits_;
    std::vector<hipStream_t> streams_;

    void release_streams() {
        for (auto s : streams_) (void)hipStreamDestroy(s);
        streams_.clear();
    }
};

// ServeError -> the library's Error, so guard() reports it.
template <typename F>
auto serve_call(F&& f) -> decltype(f()) {
    try {
        return f();
    } catch (const ServeError& e) {
        fail(e.code, e.msg);
    }
}

}  // namespace

struct aqs_server_s {
    std::unique_ptr<aqs_serve::Server> core;
};

namespace {

aqs_server check_server(aqs_server s) {
    if (!s || !s->core) fail(AQS_ERR_INVALID_ARGUMENT, "null server handle");
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
        auto backend = std::make_unique<HipBackend>(current_device(), fuse_qubits, workers);
        auto s = std::make_unique<aqs_server_s>();
        serve_call([&] { s->core = std::make_unique<aqs_serve::Server>(std::move(backend), hbm_budget_bytes, workers); });
        *out = s.release();
    });
}

aqs_status aqs_server_set_starvation_limit(aqs_server s, double seconds) {
    return guard([&] { check_server(s)->core->set_starvation_limit(seconds); });
}

aqs_status aqs_server_set_grouping(aqs_server s, int max_qubits, int max_requests) {
    return guard([&] {
        check_server(s);
        if (max_qubits < 0 || max_qubits > kGroupMaxQubits)
            fail(AQS_ERR_INVALID_ARGUMENT, "group qubit limit must be 0..16");
        if (max_requests < 1 || max_requests > 256) fail(AQS_ERR_INVALID_ARGUMENT, "group size must be 1..256");
        serve_call([&] { s->core->set_grouping(max_qubits, max_requests); });
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
        r->ops.assign(ops, ops + nops);
        for (auto& op : r->ops) {
            validate_op(op, nqubits);
            if (op.kind == AQS_OP_MATK) {
                const size_t K = (size_t)1 << op.ntargets;
                r->mats.emplace_back(op.u, op.u + 2 * K * K);
            }
        }
        // point MATK ops at the request's own copies (after mats settled)
        size_t mi = 0;
        for (auto& op : r->ops)
            if (op.kind == AQS_OP_MATK) op.u = r->mats[mi++].data();
        r->u.assign(uniforms, uniforms + shots);
        serve_call([&] { s->core->submit(std::move(r)); });
    });
}

aqs_status aqs_server_wait(aqs_server s, uint64_t id, int timeout_ms, aqs_request_result* res, uint64_t* indices,
                           void* state_host) {
    return guard([&] {
        check_server(s);
        std::shared_ptr<Request> r = serve_call([&] { return s->core->wait(id, timeout_ms); });
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

aqs_status aqs_server_cancel(aqs_server s, uint64_t id, int* cancelled) {
    return guard([&] {
        check_server(s);
        const bool c = serve_call([&] { return s->core->cancel(id); });
        if (cancelled) *cancelled = c ? 1 : 0;
    });
}

aqs_status aqs_server_stats(aqs_server s, aqs_server_stats_t* out) {
    return guard([&] {
        check_server(s);
        if (!out) return;
        const aqs_serve::Stats st = s->core->stats();
        out->admitted = st.admitted;
        out->completed = st.completed;
        out->rejected = st.rejected;
        out->evictions = st.evictions;
        out->pending = st.pending;
        out->running = st.running;
        out->max_concurrent = st.max_concurrent;
        out->resident_bytes = st.resident_bytes;
        out->peak_resident_bytes = st.peak_resident_bytes;
        out->budget_bytes = st.budget_bytes;
        out->cancelled = st.cancelled;
        out->groups = st.groups;
        out->grouped = st.grouped;
    });
}

aqs_status aqs_server_destroy(aqs_server s) {
    return guard([&] {
        if (!s) return;
        std::unique_ptr<aqs_server_s> owned(s);
        owned->core.reset();  // finishes queued work, frees every resident state
    });
}

}  // extern "C"
