// server.hpp -- continuous-batching request loop over a device-resident Batch.
//
// Jobs (a deck, a seed and a step count) are parsed on the submitting thread and
// queued.  One worker thread owns the GPU: it admits queued jobs in arrival order
// while their device memory fits the HBM budget, advances every running job
// together in chunks (one captured step graph per running set, each kernel launch
// covering all replicas), and retires finished jobs -- final properties and
// positions are copied out and their memory released, which lets the next queued
// jobs in.  Time-to-result is measured from submission to retirement.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tild.h"
#include "engine.hpp"
#include "system.hpp"

namespace tild {

struct Job {
    int id = -1;
    std::unique_ptr<System> sys;
    uint32_t seed = 0;
    int nsteps = 0, done_steps = 0;
    int state = TILD_JOB_QUEUED;
    int replica = -1;
    size_t bytes = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    tild_props props{};
    std::vector<float> x;
    std::string err;
    int ns = 0, dim = 3;
    double step_work = 0;   // beads + grid nodes: the work of one step
};

class Server {
public:
    Server(const tild_server_opts& o) : opts_(o) {}
    ~Server();
    int start(std::string& err);
    int submit(const std::string& text, const std::string& base, uint32_t seed, int nsteps, int* job,
               std::string& err);
    int poll(int job, int* state);
    int wait(int job, double timeout_s, std::string& err);
    int drain(double timeout_s, std::string& err);
    int result(int job, tild_props* props, float* x, size_t n, double* latency, std::string& err);
    void stats(tild_server_stats* s);

private:
    void loop();
    double now() const;

    tild_server_opts opts_;
    std::unique_ptr<Batch> batch_;
    std::thread worker_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    bool stop_ = false;
    std::deque<int> queue_;
    std::map<int, std::unique_ptr<Job>> jobs_;
    std::vector<int> running_;
    int admitting_ = 0;            // jobs taken off the queue whose replica is being built
    std::vector<Job*> pending_;    // those jobs, in admission order (worker thread only)
    int next_id_ = 0;
    size_t bytes_in_use_ = 0, peak_bytes_ = 0;
    double t0_ = 0, busy_s_ = 0, admit_s_ = 0, retire_s_ = 0;
    hipEvent_t ev_[2] = {nullptr, nullptr};   // around each chunk's launches: its device time
    double bead_steps_ = 0;
    std::string fatal_;
};

}  // namespace tild
