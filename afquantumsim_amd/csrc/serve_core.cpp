// Scheduling core of the request server (host-only; see serve_core.h).
#include "serve_core.h"

#include <algorithm>
#include <new>

namespace aqs_serve {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Server::Server(std::unique_ptr<Backend> backend, size_t budget_bytes, int workers)
    : backend_(std::move(backend)), budget_(budget_bytes) {
    if (!backend_) throw ServeError{AQS_ERR_INVALID_ARGUMENT, "null backend"};
    if (workers < 1 || workers > 16) throw ServeError{AQS_ERR_INVALID_ARGUMENT, "workers must be 1..16"};
    st_.budget_bytes = budget_bytes;
    slot_.assign((size_t)workers, {});
    slot_group_.assign((size_t)workers, false);
    slot_fp_.assign((size_t)workers, 0);
    if (backend_->group_max_qubits() > 0) {
        arena_bytes_ = backend_->footprint(kGroupKey);
        if (arena_bytes_ <= budget_) group_qubits_ = backend_->group_max_qubits();
    }
    try {
        for (int w = 0; w < workers; ++w) workers_.emplace_back(&Server::worker_main, this, w);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stopping_ = true;
        }
        cv_work_.notify_all();
        for (auto& t : workers_) t.join();
        throw;
    }
}

Server::~Server() {
    {
        std::unique_lock<std::mutex> lock(mu_);
        cv_done_.wait(lock, [&] { return pending_.empty() && st_.running == 0; });
        stopping_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : workers_) t.join();
    for (auto& kv : idle_)
        for (void* s : kv.second) backend_->destroy_state(s);
}

void Server::set_starvation_limit(double seconds) {
    std::lock_guard<std::mutex> lock(mu_);
    starvation_s_ = seconds;
}

void Server::set_grouping(int max_qubits, int max_requests) {
    if (max_qubits < 0 || max_requests < 1)
        throw ServeError{AQS_ERR_INVALID_ARGUMENT, "group limits: qubits >= 0, requests >= 1"};
    std::lock_guard<std::mutex> lock(mu_);
    const int cap = arena_bytes_ <= budget_ ? backend_->group_max_qubits() : 0;
    group_qubits_ = std::min(max_qubits, cap);
    group_requests_ = max_requests;
}

void Server::submit(std::shared_ptr<Request> r) {
    if (!r) throw ServeError{AQS_ERR_INVALID_ARGUMENT, "null request"};
    r->footprint = backend_->footprint(r->n);
    r->group_cost = backend_->group_max_qubits() > 0 ? backend_->group_bytes(*r) : 0;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lock(mu_);
    if (r->footprint > budget_) {
        ++st_.rejected;
        throw ServeError{AQS_ERR_INVALID_ARGUMENT, "request needs more device memory than the server budget"};
    }
    if (all_.count(r->id)) {
        ++st_.rejected;
        throw ServeError{AQS_ERR_INVALID_ARGUMENT, "duplicate request id"};
    }
    all_[r->id] = r;
    pending_.push_back(std::move(r));
    dispatch_locked();
}

// A request that may run in a group (called with mu_ held).
bool Server::groupable_locked(const Request& r) const {
    return group_qubits_ > 0 && r.n <= group_qubits_ && r.group_cost <= arena_bytes_;
}

// Assign pending requests to idle workers (called with mu_ held).
void Server::dispatch_locked() {
    const double t = now_s();
    size_t i = 0;
    while (i < pending_.size()) {
        int w = -1;
        for (size_t k = 0; k < slot_.size(); ++k)
            if (slot_[k].empty()) {
                w = (int)k;
                break;
            }
        if (w < 0) return;
        Request& r = *pending_[i];
        const bool group = groupable_locked(r);
        // Running requests hold at most their own footprints (a group: the
        // arena); every other resident state is idle and can be evicted
        // (acquire_state), so the running footprints alone decide whether
        // this one fits.
        const size_t fp = group ? arena_bytes_ : r.footprint;
        if (in_use_ + fp > budget_) {
            // an old request holds its place: nothing behind it may start
            if (t - r.t_submit > starvation_s_) return;
            ++i;
            continue;
        }
        std::vector<std::shared_ptr<Request>> job{pending_[i]};
        pending_.erase(pending_.begin() + (std::ptrdiff_t)i);
        if (group) {
            size_t bytes = r.group_cost;
            for (size_t j = i; j < pending_.size() && job.size() < (size_t)group_requests_;) {
                Request& q = *pending_[j];
                if (groupable_locked(q) && bytes + q.group_cost <= arena_bytes_) {
                    bytes += q.group_cost;
                    job.push_back(pending_[j]);
                    pending_.erase(pending_.begin() + (std::ptrdiff_t)j);
                    continue;
                }
                if (t - q.t_submit > starvation_s_) break;  // never overtake a starved request
                ++j;
            }
        }
        for (auto& q : job) {
            q->t_admit = t;
            ++st_.admitted;
            ++st_.running;
        }
        st_.max_concurrent = std::max(st_.max_concurrent, st_.running);
        slot_[(size_t)w] = std::move(job);
        slot_group_[(size_t)w] = group;
        slot_fp_[(size_t)w] = fp;
        in_use_ += fp;
        cv_work_.notify_all();
    }
}

// A state of n qubits for a worker: an idle one of that size, else a new one
// after evicting idle states (largest first) until it fits.
void* Server::acquire_state(int n) {
    const size_t need = backend_->footprint(n);
    {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = idle_.find(n);
        if (it != idle_.end() && !it->second.empty()) {
            void* st = it->second.back();
            it->second.pop_back();
            return st;
        }
        // Evicted states are freed before the lock is released, so no other
        // worker allocates against memory that is still held.
        while (resident_ + need > budget_) {
            auto largest = idle_.end();
            for (auto i = idle_.begin(); i != idle_.end(); ++i)
                if (!i->second.empty()) largest = i;
            if (largest == idle_.end()) break;
            void* victim = largest->second.back();
            largest->second.pop_back();
            backend_->destroy_state(victim);
            resident_ -= backend_->footprint(largest->first);
            ++st_.evictions;
        }
        resident_ += need;
        st_.peak_resident_bytes = std::max<uint64_t>(st_.peak_resident_bytes, resident_);
    }
    try {
        return backend_->create_state(n);
    } catch (...) {
        std::lock_guard<std::mutex> lock(mu_);
        resident_ -= need;
        throw;
    }
}

void Server::release_state(int n, void* st) {
    std::lock_guard<std::mutex> lock(mu_);
    idle_[n].push_back(st);
}

// Mark r finished (called with mu_ held).
void Server::finish_locked(Request& r) {
    r.t_done = now_s();
    r.done = true;
}

namespace {

// The failure of a request, from whatever the backend raised.
template <typename F>
void run_guarded(Request& r, F&& f) {
    try {
        f();
    } catch (const ServeError& e) {
        r.status = e.code;
        r.error = e.msg;
    } catch (const std::bad_alloc&) {
        r.status = AQS_ERR_NO_MEMORY;
        r.error = "host allocation failed";
    } catch (const std::exception& e) {
        r.status = AQS_ERR_RUNTIME;
        r.error = e.what();
    } catch (...) {
        r.status = AQS_ERR_RUNTIME;
        r.error = "unknown failure while running the request";
    }
}

void mark_cancelled(Request& r) {
    r.status = AQS_ERR_RUNTIME;
    r.error = "request cancelled";
}

}  // namespace

void Server::run_single(Request& r, int w) {
    void* st = nullptr;
    run_guarded(r, [&] {
        if (r.cancel.load()) throw ServeError{AQS_ERR_RUNTIME, "request cancelled"};
        st = acquire_state(r.n);
        backend_->run(r, st, w);
    });
    if (st) release_state(r.n, st);
}

// Members cancelled before the launch are left out; one cancelled during it
// completes as cancelled and its results are dropped.
void Server::run_grouped(std::vector<std::shared_ptr<Request>>& job, int w) {
    std::vector<Request*> live;
    for (auto& q : job) {
        if (q->cancel.load())
            mark_cancelled(*q);
        else
            live.push_back(q.get());
    }
    if (live.empty()) return;
    void* arena = nullptr;
    Request failed;
    run_guarded(failed, [&] {
        arena = acquire_state(kGroupKey);
        backend_->run_group(live, arena, w);
    });
    if (arena) release_state(kGroupKey, arena);
    for (Request* q : live) {
        if (failed.status != AQS_OK) {
            q->status = failed.status;
            q->error = failed.error;
        } else if (q->cancel.load()) {
            mark_cancelled(*q);
            q->indices.clear();
            q->state.clear();
        }
    }
}

void Server::worker_main(int w) {
    backend_->worker_start(w);
    for (;;) {
        std::vector<std::shared_ptr<Request>> job;
        bool group = false;
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_work_.wait(lock, [&] { return stopping_ || !slot_[(size_t)w].empty(); });
            if (slot_[(size_t)w].empty()) return;  // stopping with nothing assigned
            job = slot_[(size_t)w];
            group = slot_group_[(size_t)w];
        }
        if (group)
            run_grouped(job, w);
        else
            run_single(*job[0], w);
        {
            std::lock_guard<std::mutex> lock(mu_);
            for (auto& q : job) {
                finish_locked(*q);
                if (q->status == AQS_ERR_RUNTIME && q->cancel.load()) ++st_.cancelled;
                --st_.running;
                ++st_.completed;
            }
            if (group) {
                ++st_.groups;
                st_.grouped += job.size();
            }
            slot_[(size_t)w].clear();
            in_use_ -= slot_fp_[(size_t)w];
            slot_fp_[(size_t)w] = 0;
            dispatch_locked();
        }
        cv_done_.notify_all();
    }
}

std::shared_ptr<Request> Server::wait(uint64_t id, int timeout_ms) {
    std::unique_lock<std::mutex> lock(mu_);
    std::shared_ptr<Request> r;
    {
        auto it = all_.find(id);
        if (it == all_.end()) throw ServeError{AQS_ERR_OUT_OF_RANGE, "unknown request id"};
        r = it->second;
    }
    auto ready = [&] { return r->done; };
    // A system_clock deadline waits through pthread_cond_timedwait; a
    // steady_clock one (wait_for) goes through pthread_cond_clockwait, which
    // the ThreadSanitizer runtime of this toolchain does not intercept.
    if (timeout_ms < 0)
        cv_done_.wait(lock, ready);
    else if (!cv_done_.wait_until(lock, std::chrono::system_clock::now() + std::chrono::milliseconds(timeout_ms), ready))
        throw ServeError{AQS_ERR_RUNTIME, "request not finished within the timeout"};
    // the map may have changed while the lock was released: look the id up
    // again; another waiter may already have collected this result
    auto it = all_.find(id);
    if (it == all_.end() || it->second != r)
        throw ServeError{AQS_ERR_OUT_OF_RANGE, "request result was already collected"};
    all_.erase(it);
    return r;
}

bool Server::cancel(uint64_t id) {
    bool finished_here = false;
    {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = all_.find(id);
        if (it == all_.end()) throw ServeError{AQS_ERR_OUT_OF_RANGE, "unknown request id"};
        Request& r = *it->second;
        if (r.done) return false;
        r.cancel.store(true);
        auto p = std::find(pending_.begin(), pending_.end(), it->second);
        if (p != pending_.end()) {
            pending_.erase(p);
            r.status = AQS_ERR_RUNTIME;
            r.error = "request cancelled";
            finish_locked(r);
            ++st_.cancelled;
            finished_here = true;
            dispatch_locked();
        }
    }
    if (finished_here) cv_done_.notify_all();
    return true;
}

Stats Server::stats() {
    std::lock_guard<std::mutex> lock(mu_);
    Stats s = st_;
    s.pending = pending_.size();
    s.resident_bytes = resident_;
    return s;
}

}  // namespace aqs_serve
