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
    slot_.assign((size_t)workers, nullptr);
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

void Server::submit(std::shared_ptr<Request> r) {
    if (!r) throw ServeError{AQS_ERR_INVALID_ARGUMENT, "null request"};
    r->footprint = backend_->footprint(r->n);
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

// Assign pending requests to idle workers (called with mu_ held).
void Server::dispatch_locked() {
    const double t = now_s();
    for (auto it = pending_.begin(); it != pending_.end();) {
        int w = -1;
        for (size_t i = 0; i < slot_.size(); ++i)
            if (!slot_[i]) {
                w = (int)i;
                break;
            }
        if (w < 0) return;
        Request& r = **it;
        // Running requests hold at most their own footprints; every other
        // resident state is idle and can be evicted (acquire_state), so the
        // running footprints alone decide whether this one fits.
        const bool fits = in_use_ + r.footprint <= budget_;
        if (!fits) {
            // an old request holds its place: nothing behind it may start
            if (t - r.t_submit > starvation_s_) return;
            ++it;
            continue;
        }
        slot_[(size_t)w] = *it;
        in_use_ += r.footprint;
        r.t_admit = t;
        ++st_.admitted;
        ++st_.running;
        st_.max_concurrent = std::max(st_.max_concurrent, st_.running);
        it = pending_.erase(it);
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

void Server::worker_main(int w) {
    backend_->worker_start(w);
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_work_.wait(lock, [&] { return stopping_ || slot_[(size_t)w]; });
            if (!slot_[(size_t)w]) return;  // stopping with nothing assigned
            r = slot_[(size_t)w];
        }
        void* st = nullptr;
        try {
            if (r->cancel.load()) throw ServeError{AQS_ERR_RUNTIME, "request cancelled"};
            st = acquire_state(r->n);
            backend_->run(*r, st, w);
        } catch (const ServeError& e) {
            r->status = e.code;
            r->error = e.msg;
        } catch (const std::bad_alloc&) {
            r->status = AQS_ERR_NO_MEMORY;
            r->error = "host allocation failed";
        } catch (const std::exception& e) {
            r->status = AQS_ERR_RUNTIME;
            r->error = e.what();
        } catch (...) {
            r->status = AQS_ERR_RUNTIME;
            r->error = "unknown failure while running the request";
        }
        if (st) release_state(r->n, st);
        {
            std::lock_guard<std::mutex> lock(mu_);
            finish_locked(*r);
            if (r->status == AQS_ERR_RUNTIME && r->cancel.load()) ++st_.cancelled;
            slot_[(size_t)w] = nullptr;
            in_use_ -= r->footprint;
            --st_.running;
            ++st_.completed;
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
