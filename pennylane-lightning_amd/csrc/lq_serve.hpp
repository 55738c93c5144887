// Request server for energy + adjoint-gradient evaluations with continuous
// batching on one GPU.
//
// A request is (n qubits, ops list, Pauli-sum Hamiltonian, trainable
// parameters); its result is <H> and d<H>/d theta. Requests are admitted as
// they arrive:
//   * small ones (state fits the packed kernel) get their program built by a
//     pool of host threads and are launched many per launch, one workgroup
//     each; up to `batch_streams` launches are in flight, and every request
//     that became ready while a launch ran joins the next one;
//   * larger ones run through the engine (fused tile passes + adjoint sweep)
//     on `large_workers` threads with their own streams, admitted while the
//     HBM they need (3 states of 2^n amplitudes) fits the budget; their
//     engine states come from a pool of idle states the workers share;
//     `lane_workers` more threads take only the engine requests of up to
//     `lane_max_qubits` qubits, which would otherwise wait behind large ones.
// In plain mode one worker runs every request to completion, in arrival
// order, through the engine: the one-request-at-a-time baseline.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lq_batch.hpp"
#include "lq_engine.hpp"

namespace lqamd {

struct ServerConfig {
    int dtype = 1;
    int batched = 1;                 // 0: plain FIFO through the engine
    uint64_t hbm_budget = 250ull << 30;
    int small_max_qubits = -1;       // -1: the sizes whose states fit the packed kernel's LDS
    int max_batch = 4096;            // requests per packed launch
    int batch_streams = 2;           // packed launches in flight
    // 8: measured on the 1024-request stream, the batched server kept up at
    // the same rates with 4 and 8 workers and its p99 was lower with 8
    int large_workers = 8;
    int prep_threads = 4;
    // engine workers that take only requests of <= lane_max_qubits (oldest
    // first), so that they do not queue behind the large ones: at 1150
    // requests/s the 13-16-qubit engine requests (~0.1-1 ms of GPU work
    // each, 28% of the stream) waited 360-410 ms (p50) behind 17-28-qubit
    // requests in the shared first-come queue
    int lane_workers = 2;
    int lane_max_qubits = 16;
};

struct ServeResult {
    double energy = 0.0;
    std::vector<double> grad;
    double latency_ms = 0.0; // submit -> result ready
    int path = 0;            // 0: packed, 1: engine
    std::string error;
};

class Server {
  public:
    explicit Server(const ServerConfig &cfg);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    uint64_t submit(int n, const OpsList &ops, const std::shared_ptr<Obs> &ham, const std::vector<uint64_t> &tp);
    // Block until request `id` is done; false on timeout. The result is kept
    // until taken.
    bool wait(uint64_t id, double timeout_s, ServeResult *out);

    struct Stats {
        uint64_t submitted = 0, completed = 0, packed = 0, engine = 0, launches = 0;
        uint64_t hbm_in_use = 0;
        uint64_t states_made = 0, states_freed = 0; // engine states allocated / evicted for room
        // packed path, longest since the last reset: a request's wait between
        // its program being ready and its launch, the host time inside one
        // launch call, and one launch's time from that call to the batcher
        // seeing it done
        double ready_wait_max_ms = 0.0, launch_call_max_ms = 0.0, launch_max_ms = 0.0;
        // the GPU segments (upload, kernels, download) of that longest launch
        double slow_launch_seg_ms[3] = {0.0, 0.0, 0.0};
    };
    Stats stats();
    void reset_timing();

  private:
    using Clock = std::chrono::steady_clock;
    struct Req {
        uint64_t id = 0;
        int n = 0;
        OpsList ops;
        std::shared_ptr<Obs> ham;
        std::vector<uint64_t> tp;
        Clock::time_point t0, t_ready;
        BatchProgram prog;
        uint64_t bytes = 0;
        ServeResult res;
        bool done = false;
    };
    using ReqP = std::shared_ptr<Req>;

    ServerConfig cfg_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<ReqP> incoming_, small_ready_, large_ready_;
    std::unordered_map<uint64_t, ReqP> live_;
    uint64_t next_id_ = 1;
    uint64_t hbm_in_use_ = 0;
    // engine states not in use, shared by the workers (oldest first), each
    // counted against the HBM budget until it is freed
    std::vector<std::pair<int, std::unique_ptr<StateVector>>> idle_;
    Stats stats_;
    bool stop_ = false;
    std::vector<std::thread> threads_;
    int device_ = 0;

    void complete(const ReqP &r);
    void prep_loop();
    void batch_loop();
    void large_loop(bool lane);
    void run_engine(StateVector &sv, Req &r);
};

} // namespace lqamd
