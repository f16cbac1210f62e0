// server.cpp -- continuous batching of TILD jobs on one device (server.hpp).
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace tild {

double Server::now() const {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    batch_.reset();
    for (hipEvent_t e : ev_) if (e) (void)hipEventDestroy(e);
}

int Server::start(std::string& err) {
    if (opts_.chunk_steps <= 0) opts_.chunk_steps = 100;
    batch_ = std::make_unique<Batch>(opts_.device);
    batch_->set_spread(0);
    if (int rc = batch_->init(err)) return rc;
    for (hipEvent_t& e : ev_)
        if (hipEventCreate(&e) != hipSuccess) { err = "hipEventCreate failed"; return -TILD_ERR_DEVICE; }
    if (opts_.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        hipError_t e = hipMemGetInfo(&fr, &tot);
        if (e != hipSuccess) { err = std::string("hipMemGetInfo: ") + hipGetErrorString(e); return -TILD_ERR_DEVICE; }
        opts_.hbm_budget = (size_t)(0.9 * (double)fr);
    }
    t0_ = now();
    worker_ = std::thread([this] { loop(); });
    return 0;
}

int Server::submit(const std::string& text, const std::string& base, uint32_t seed, int nsteps, int* job,
                   std::string& err) {
    if (nsteps < 1) { err = "a job needs at least one step"; return -TILD_ERR_ARG; }
    auto J = std::make_unique<Job>();
    J->sys = std::make_unique<System>();
    if (int rc = load_system(text, base, *J->sys, err)) return -rc;
    J->bytes = Batch::estimate_bytes(*J->sys);
    J->seed = seed;
    J->nsteps = nsteps;
    J->ns = J->sys->ns;
    J->dim = J->sys->Dim;
    J->step_work = (double)J->sys->ns + (double)J->sys->M;
    std::lock_guard<std::mutex> lk(mu_);
    if (!fatal_.empty()) { err = fatal_; return -TILD_ERR_DEVICE; }
    if (J->bytes > opts_.hbm_budget) {
        err = "job needs about " + std::to_string(J->bytes) + " bytes of device memory, budget is " +
              std::to_string(opts_.hbm_budget);
        return -TILD_ERR_BUDGET;
    }
    J->id = next_id_++;
    J->t_submit = now();
    *job = J->id;
    queue_.push_back(J->id);
    jobs_[J->id] = std::move(J);
    cv_.notify_all();
    return 0;
}

// Worker: take the jobs that fit, enqueue one chunk of the running jobs, build the new
// jobs' replicas while it runs (those not reached before it ends wait for the next
// chunk), wait for it, retire; repeat.  Only this thread touches the Batch.
void Server::loop() {
    (void)hipSetDevice(opts_.device);
    std::string err;
    for (;;) {
        // --- admission (arrival order; the queue head waits until it fits)
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty() || !running_.empty() || !pending_.empty(); });
            if (stop_) break;
            size_t used = bytes_in_use_;
            int nrun = (int)(running_.size() + pending_.size());
            while (!queue_.empty()) {
                Job* J = jobs_[queue_.front()].get();
                if (used + J->bytes > opts_.hbm_budget) break;
                if (opts_.max_replicas > 0 && nrun >= opts_.max_replicas) break;
                used += J->bytes;
                ++nrun;
                pending_.push_back(J);
                queue_.pop_front();
            }
            bytes_in_use_ = used;   // reserved under the estimate until the replica exists
            admitting_ = (int)pending_.size();   // neither queued nor running until added
        }
        // --- one chunk of every running job, cut at the nearest finish: enqueued first,
        // so that the admitted jobs' replicas are built (on the batch's set-up stream)
        // while it runs
        std::vector<int> reps, ids;
        int chunk = opts_.chunk_steps;
        double beads = 0;
        {
            std::lock_guard<std::mutex> lk(mu_);
            std::vector<Job*> run;
            for (int jid : running_) run.push_back(jobs_[jid].get());
            if (opts_.step_work > 0) {
                // least remaining work first, until the chunk's work per step reaches the target
                std::sort(run.begin(), run.end(), [](const Job* a, const Job* b) {
                    const double wa = (a->nsteps - a->done_steps) * a->step_work;
                    const double wb = (b->nsteps - b->done_steps) * b->step_work;
                    return wa != wb ? wa < wb : a->id < b->id;
                });
                double w = 0;
                size_t k = 0;
                while (k < run.size() && (k == 0 || w < 1000.0 * opts_.step_work)) w += run[k++]->step_work;
                run.resize(k);
            }
            for (Job* J : run) {
                chunk = std::min(chunk, J->nsteps - J->done_steps);
                reps.push_back(J->replica);
                ids.push_back(J->id);
                beads += J->ns;
            }
        }
        std::sort(reps.begin(), reps.end());
        const bool stepping = !reps.empty() && chunk > 0;
        int rc = 0;
        if (stepping) {
            (void)hipEventRecord(ev_[0], batch_->stream());
            rc = batch_->step(reps, chunk, err, opts_.use_graph != 0);
            (void)hipEventRecord(ev_[1], batch_->stream());
        }
        // builds run while the chunk does; one that would start after the chunk has
        // finished waits for the next chunk, which goes out first (at least one build a
        // pass, one at a time while nothing runs)
        size_t nbuilt = 0;
        for (; nbuilt < pending_.size(); ++nbuilt) {
            if (nbuilt > 0 && (!stepping || hipEventQuery(ev_[1]) == hipSuccess)) break;
            Job* J = pending_[nbuilt];
            const double ts = now();
            std::string e1;
            const int id = batch_->add(*J->sys, J->seed, e1);
            std::lock_guard<std::mutex> lk(mu_);
            admit_s_ += now() - ts;
            bytes_in_use_ -= J->bytes;
            --admitting_;
            if (id < 0) {
                J->state = TILD_JOB_FAILED;
                J->err = e1;
                J->t_done = now();
                done_cv_.notify_all();
                continue;
            }
            J->replica = id;
            J->bytes = batch_->rep(id)->bytes;
            bytes_in_use_ += J->bytes;
            peak_bytes_ = std::max(peak_bytes_, bytes_in_use_);
            J->t_start = ts;
            J->state = TILD_JOB_RUNNING;
            J->sys.reset();   // the replica holds everything the step needs
            running_.push_back(J->id);   // joins from the next chunk on
        }
        pending_.erase(pending_.begin(), pending_.begin() + nbuilt);
        if (!stepping) continue;
        if (!rc) rc = batch_->sync(err);
        float gpu_ms = 0.f;
        if (!rc) (void)hipEventElapsedTime(&gpu_ms, ev_[0], ev_[1]);
        const double dt = 1e-3 * gpu_ms;
        // --- retire the finished jobs
        std::vector<Job*> fin;
        {
            std::lock_guard<std::mutex> lk(mu_);
            busy_s_ += dt;
            bead_steps_ += beads * std::max(chunk, 0);
            for (int jid : ids) {
                Job* J = jobs_[jid].get();
                if (rc) { J->state = TILD_JOB_FAILED; J->err = err; fin.push_back(J); continue; }
                J->done_steps += chunk;
                if (J->done_steps >= J->nsteps) fin.push_back(J);
            }
            if (rc) fatal_ = err;
        }
        const double tr = now();
        for (Job* J : fin) {
            std::string e2;
            if (J->state != TILD_JOB_FAILED) {
                int r2 = batch_->properties(J->replica, &J->props, e2);
                if (!r2) {
                    J->x.resize((size_t)J->ns * J->dim);
                    r2 = batch_->get_positions(J->replica, J->x.data(), J->x.size(), e2);
                }
                if (r2) { J->state = TILD_JOB_FAILED; J->err = e2; }
            }
            const size_t held = batch_->rep(J->replica) ? batch_->rep(J->replica)->bytes : 0;
            batch_->remove(J->replica);
            std::lock_guard<std::mutex> lk(mu_);
            bytes_in_use_ -= held;
            if (J->state != TILD_JOB_FAILED) J->state = TILD_JOB_DONE;
            J->t_done = now();
            running_.erase(std::find(running_.begin(), running_.end(), J->id));
        }
        if (!fin.empty()) {
            std::lock_guard<std::mutex> lk(mu_);
            retire_s_ += now() - tr;
            done_cv_.notify_all();
        }
        if (rc) {
            // a device error ends the server: fail whatever is still queued or unbuilt
            std::lock_guard<std::mutex> lk(mu_);
            for (int jid : queue_) pending_.push_back(jobs_[jid].get());
            for (Job* J : pending_) {
                J->state = TILD_JOB_FAILED;
                J->err = fatal_;
                J->t_done = now();
            }
            queue_.clear();
            pending_.clear();
            admitting_ = 0;
            done_cv_.notify_all();
            break;
        }
    }
}

int Server::poll(int job, int* state) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) return -TILD_ERR_ARG;
    *state = it->second->state;
    return 0;
}

int Server::wait(int job, double timeout_s, std::string& err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) { err = "no such job"; return -TILD_ERR_ARG; }
    Job* J = it->second.get();
    auto done = [&] { return J->state == TILD_JOB_DONE || J->state == TILD_JOB_FAILED; };
    if (timeout_s < 0) done_cv_.wait(lk, done);
    else if (!done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        err = "timed out waiting for job " + std::to_string(job);
        return -TILD_ERR_ARG;
    }
    if (J->state == TILD_JOB_FAILED) { err = J->err; return -TILD_ERR_DEVICE; }
    return 0;
}

int Server::drain(double timeout_s, std::string& err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto idle = [&] { return queue_.empty() && running_.empty() && admitting_ == 0; };
    if (timeout_s < 0) done_cv_.wait(lk, idle);
    else if (!done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), idle)) {
        err = "timed out draining the server";
        return -TILD_ERR_ARG;
    }
    return 0;
}

int Server::result(int job, tild_props* props, float* x, size_t n, double* latency, std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) { err = "no such job"; return -TILD_ERR_ARG; }
    Job* J = it->second.get();
    if (J->state == TILD_JOB_FAILED) { err = J->err; return -TILD_ERR_DEVICE; }
    if (J->state != TILD_JOB_DONE) { err = "job not finished"; return -TILD_ERR_ARG; }
    if (props) *props = J->props;
    if (x) {
        if (n < J->x.size()) { err = "position buffer too small"; return -TILD_ERR_ARG; }
        std::copy(J->x.begin(), J->x.end(), x);
    }
    if (latency) *latency = J->t_done - J->t_submit;
    return 0;
}

static double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest-rank quantile
    size_t k = (size_t)std::ceil(q * (double)v.size());
    k = std::min(std::max<size_t>(k, 1), v.size());
    return v[k - 1];
}

void Server::stats(tild_server_stats* s) {
    std::lock_guard<std::mutex> lk(mu_);
    *s = tild_server_stats{};
    std::vector<double> lat, wait, svc;
    for (auto& kv : jobs_) {
        const Job* J = kv.second.get();
        s->submitted += 1;
        if (J->state == TILD_JOB_QUEUED) s->queued += 1;
        else if (J->state == TILD_JOB_RUNNING) s->running += 1;
        else if (J->state == TILD_JOB_DONE) {
            s->done += 1;
            lat.push_back(J->t_done - J->t_submit);
            wait.push_back(J->t_start - J->t_submit);
            svc.push_back(J->t_done - J->t_start);
        } else s->failed += 1;
    }
    s->bytes_in_use = bytes_in_use_;
    s->peak_bytes = peak_bytes_;
    s->budget = opts_.hbm_budget;
    s->busy_s = busy_s_;
    s->elapsed_s = now() - t0_;
    s->bead_steps = bead_steps_;
    s->p50_s = quantile(lat, 0.50);
    s->p99_s = quantile(lat, 0.99);
    s->admit_s = admit_s_;
    s->retire_s = retire_s_;
    s->wait50_s = quantile(wait, 0.50);
    s->wait99_s = quantile(wait, 0.99);
    s->svc50_s = quantile(svc, 0.50);
    s->svc99_s = quantile(svc, 0.99);
}

}  // namespace tild
