// Request server (see lq_serve.hpp).
#include "lq_serve.hpp"

#include <algorithm>
#include <map>

namespace lqamd {

namespace {
uint64_t request_bytes(int n, int dtype) { return 3ull * (static_cast<uint64_t>(dtype ? 16 : 8) << n); }
} // namespace

Server::Server(const ServerConfig &cfg) : cfg_(cfg) {
    hip_check(hipGetDevice(&device_), "hipGetDevice");
    // by default only requests whose two states fit the packed kernel's LDS
    // are packed: from 13 qubits (complex128) the kernel works on its states
    // in global memory one gate at a time, and a launch lasts as long as its
    // largest request (~110 ms with 16-qubit depth-16 requests), while the
    // engine's fused passes on concurrent workers serve those requests in a
    // few ms each (measured: 209 against 177 requests/s at saturation, p50
    // 8 ms against 1.7 s)
    const int lim = batch_max_qubits(cfg_.dtype);
    cfg_.small_max_qubits = cfg_.small_max_qubits < 0 ? batch_lds_qubits(cfg_.dtype)
                                                      : std::min(cfg_.small_max_qubits, lim);
    cfg_.max_batch = std::max(1, cfg_.max_batch);
    cfg_.batch_streams = std::max(1, cfg_.batch_streams);
    cfg_.large_workers = std::max(1, cfg_.large_workers);
    cfg_.prep_threads = std::max(1, cfg_.prep_threads);
    cfg_.lane_workers = std::max(0, cfg_.lane_workers);
    if (cfg_.batched) {
        for (int i = 0; i < cfg_.prep_threads; ++i) threads_.emplace_back([this] { prep_loop(); });
        threads_.emplace_back([this] { batch_loop(); });
        for (int i = 0; i < cfg_.large_workers; ++i) threads_.emplace_back([this] { large_loop(false); });
        for (int i = 0; i < cfg_.lane_workers; ++i) threads_.emplace_back([this] { large_loop(true); });
    } else {
        threads_.emplace_back([this] { large_loop(false); });
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    for (auto &t : threads_) t.join();
}

uint64_t Server::submit(int n, const OpsList &ops, const std::shared_ptr<Obs> &ham, const std::vector<uint64_t> &tp) {
    LQ_ABORT_IF(n < 1 || n > 40, "invalid number of qubits");
    LQ_ABORT_IF(!ham, "null observable");
    auto r = std::make_shared<Req>();
    r->n = n;
    r->ops = ops;
    r->ham = ham;
    r->tp = tp;
    std::sort(r->tp.begin(), r->tp.end());
    r->bytes = request_bytes(n, cfg_.dtype);
    LQ_ABORT_IF(r->bytes > cfg_.hbm_budget, "request exceeds the HBM budget");
    {
        std::lock_guard<std::mutex> lk(mu_);
        LQ_ABORT_IF(stop_, "server stopped");
        r->id = next_id_++;
        r->t0 = Clock::now();
        live_[r->id] = r;
        ++stats_.submitted;
        if (cfg_.batched) incoming_.push_back(r);
        else large_ready_.push_back(r);
    }
    cv_work_.notify_all();
    return r->id;
}

bool Server::wait(uint64_t id, double timeout_s, ServeResult *out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = live_.find(id);
    LQ_ABORT_IF(it == live_.end(), "unknown request id");
    ReqP r = it->second;
    const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    if (!cv_done_.wait_until(lk, until, [&] { return r->done; })) return false;
    *out = r->res;
    live_.erase(id);
    return true;
}

Server::Stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    Stats s = stats_;
    s.hbm_in_use = hbm_in_use_;
    return s;
}

void Server::reset_timing() {
    std::lock_guard<std::mutex> lk(mu_);
    stats_.ready_wait_max_ms = stats_.launch_call_max_ms = stats_.launch_max_ms = 0.0;
    for (double &v : stats_.slow_launch_seg_ms) v = 0.0;
}

void Server::complete(const ReqP &r) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        r->res.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t0).count();
        r->done = true;
        ++stats_.completed;
        if (r->res.path == 0) ++stats_.packed;
        else ++stats_.engine;
        r->prog = BatchProgram{};
        r->ops = OpsList{};
    }
    cv_done_.notify_all();
}

void Server::prep_loop() {
    (void)hipSetDevice(device_);
    while (true) {
        ReqP r;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !incoming_.empty(); });
            if (stop_) return;
            r = incoming_.front();
            incoming_.pop_front();
        }
        bool small = false;
        if (r->n <= cfg_.small_max_qubits) {
            std::string why;
            try {
                small = build_batch_program(r->n, cfg_.dtype, r->ops, *r->ham, r->tp, &r->prog, &why);
            } catch (const std::exception &) {
                small = false;
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            r->t_ready = Clock::now();
            (small ? small_ready_ : large_ready_).push_back(r);
        }
        cv_work_.notify_all();
    }
}

void Server::batch_loop() {
    (void)hipSetDevice(device_);
    const int K = cfg_.batch_streams;
    std::vector<std::unique_ptr<BatchRunner>> runners;
    for (int i = 0; i < K; ++i) runners.emplace_back(std::make_unique<BatchRunner>());
    std::vector<std::vector<ReqP>> flight(K);
    std::vector<Clock::time_point> launched(K);
    auto ms_since = [](Clock::time_point t) {
        return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
    };
    auto retire = [&](int i, bool block) {
        if (flight[i].empty()) return false;
        std::vector<double> e;
        std::vector<std::vector<double>> g;
        std::string err;
        try {
            // a device error seen by the query fails the requests of this launch
            if (!block && !runners[i]->done()) return false;
            runners[i]->finish(&e, &g);
        } catch (const std::exception &ex) {
            err = ex.what();
        }
        {
            const double t = ms_since(launched[i]);
            double seg[3];
            runners[i]->segments(seg);
            std::lock_guard<std::mutex> lk(mu_);
            if (t > stats_.launch_max_ms) {
                stats_.launch_max_ms = t;
                for (int j = 0; j < 3; ++j) stats_.slow_launch_seg_ms[j] = seg[j];
            }
        }
        for (size_t j = 0; j < flight[i].size(); ++j) {
            Req &r = *flight[i][j];
            r.res.path = 0;
            if (err.empty()) {
                r.res.energy = e[j];
                r.res.grad = std::move(g[j]);
            } else {
                r.res.error = err;
            }
            complete(flight[i][j]);
        }
        flight[i].clear();
        return true;
    };
    while (true) {
        for (int i = 0; i < K; ++i) retire(i, false);
        int free_slot = -1;
        for (int i = 0; i < K && free_slot < 0; ++i)
            if (flight[i].empty()) free_slot = i;
        std::vector<ReqP> take;
        {
            std::unique_lock<std::mutex> lk(mu_);
            bool busy = false;
            for (int i = 0; i < K; ++i) busy = busy || !flight[i].empty();
            if (stop_ && !busy) return;
            if (free_slot >= 0 && !small_ready_.empty() && !stop_) {
                // at most max_batch requests, and no more program bytes than
                // the runner holds without reallocating (a reallocation waits
                // for every engine request in flight; one launch of ~2000
                // requests already fills the device several times over)
                const size_t cap = runners[free_slot]->capacity();
                size_t cnt = 0, bytes = 0;
                while (cnt < small_ready_.size() && cnt < static_cast<size_t>(cfg_.max_batch)) {
                    const size_t b = BatchRunner::program_bytes(small_ready_[cnt]->prog);
                    if (cnt > 0 && bytes + b > cap) break;
                    bytes += b;
                    ++cnt;
                }
                take.assign(small_ready_.begin(), small_ready_.begin() + static_cast<long>(cnt));
                small_ready_.erase(small_ready_.begin(), small_ready_.begin() + static_cast<long>(cnt));
                for (const auto &r : take) stats_.ready_wait_max_ms = std::max(stats_.ready_wait_max_ms, ms_since(r->t_ready));
            } else {
                // poll the launches in flight; wake early for new work
                cv_work_.wait_for(lk, std::chrono::microseconds(busy ? 20 : 2000), [&] {
                    return stop_ || (free_slot >= 0 && !small_ready_.empty());
                });
                continue;
            }
        }
        std::vector<const BatchProgram *> progs;
        progs.reserve(take.size());
        for (const auto &r : take) progs.push_back(&r->prog);
        try {
            launched[free_slot] = Clock::now();
            runners[free_slot]->launch(progs, cfg_.dtype);
            flight[free_slot] = std::move(take);
            const double t = ms_since(launched[free_slot]);
            std::lock_guard<std::mutex> lk(mu_);
            stats_.launch_call_max_ms = std::max(stats_.launch_call_max_ms, t);
            ++stats_.launches;
        } catch (const std::exception &ex) {
            for (const auto &r : take) {
                r->res.error = ex.what();
                complete(r);
            }
        }
    }
}

void Server::run_engine(StateVector &sv, Req &r) {
    sv.reset();
    for (const auto &op : r.ops.ops) sv.apply_oprec(op, false);
    std::vector<double> jac(r.tp.size());
    double ev = 0.0;
    if (!r.tp.empty()) {
        adjoint_jacobian(sv, r.ops, {r.ham}, r.tp, false, jac.data(), &ev);
    } else {
        ev = sv.expval_obs(*r.ham);
    }
    r.res.energy = ev;
    r.res.grad = std::move(jac);
    r.res.path = 1;
}

void Server::large_loop(bool lane) {
    (void)hipSetDevice(device_);
    // the worker's own stream: the pooled states it takes run on it, so the
    // server's streams are its workers and packed runners only, each with a
    // hardware queue of its own (states keep their own streams while idle)
    // (if it cannot be created, requests run on the states' own streams:
    // nothing may throw out of a worker thread)
    hipStream_t ws = nullptr;
    if (hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) != hipSuccess) ws = nullptr;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            if (!s) return;
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{ws};
    // Engine states come from a pool shared by the workers: a request takes
    // an idle state of its size if there is one, else allocates one once the
    // budget has room, evicting idle states (oldest first) to make it; the
    // state goes back to the pool afterwards. One pool instead of a cache per
    // worker: states of a size seen by several workers are not duplicated,
    // and memory is freed (hipFree waits for the device) only when needed.
    while (true) {
        ReqP r;
        std::unique_ptr<StateVector> sv;
        std::vector<std::unique_ptr<StateVector>> evicted;
        size_t k = 0; // position of the request taken in large_ready_
        {
            std::unique_lock<std::mutex> lk(mu_);
            while (true) {
                if (stop_) return;
                // a lane worker takes the oldest request of its sizes, the
                // others the oldest request
                k = 0;
                if (lane)
                    while (k < large_ready_.size() && large_ready_[k]->n > cfg_.lane_max_qubits) ++k;
                if (k < large_ready_.size()) {
                    const ReqP &h = large_ready_[k];
                    auto it = std::find_if(idle_.rbegin(), idle_.rend(), [&](const auto &e) { return e.first == h->n; });
                    if (it != idle_.rend()) {
                        sv = std::move(it->second);
                        idle_.erase(std::next(it).base());
                        break;
                    }
                    while (hbm_in_use_ + h->bytes > cfg_.hbm_budget && !idle_.empty()) {
                        hbm_in_use_ -= request_bytes(idle_.front().first, cfg_.dtype);
                        evicted.push_back(std::move(idle_.front().second));
                        idle_.erase(idle_.begin());
                        ++stats_.states_freed;
                    }
                    if (hbm_in_use_ + h->bytes <= cfg_.hbm_budget) {
                        hbm_in_use_ += h->bytes;
                        ++stats_.states_made;
                        break;
                    }
                }
                cv_work_.wait(lk);
            }
            r = large_ready_[k];
            large_ready_.erase(large_ready_.begin() + static_cast<long>(k));
        }
        evicted.clear(); // frees device memory outside the lock
        try {
            if (!sv) sv = std::make_unique<StateVector>(r->n, cfg_.dtype);
            sv->bind_stream(ws);
            run_engine(*sv, *r);
        } catch (const std::exception &ex) {
            r->res.error = ex.what();
            r->res.path = 1;
        }
        if (sv) {
            try {
                sv->bind_stream(nullptr); // idle states hold their own stream
            } catch (const std::exception &ex) {
                if (r->res.error.empty()) r->res.error = ex.what();
                sv.reset(); // a state whose stream failed is not reused
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (sv) idle_.emplace_back(r->n, std::move(sv));
            else hbm_in_use_ -= r->bytes; // not allocated, or dropped: no longer held
        }
        complete(r);
        cv_work_.notify_all();
    }
}

} // namespace lqamd
