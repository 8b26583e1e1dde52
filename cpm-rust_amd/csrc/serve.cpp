// Serving loop (see serve.hpp).  One worker thread owns the device work:
//
//   admit:  queued requests, in arrival order, while the HBM budget and the
//           active-request limit allow;
//   round:  every active request advances to the end of its current
//           schedule step, all of them in one batched engine call per mode;
//           in exact replay each chain also stops after round_s on the
//           device's clock (mid-MCS if need be; the next round resumes it),
//           so every chain runs the whole round and none waits for a
//           slower one; in sublattice mode a round is quantum attempts;
//   finish: requests whose step completed take their trace point; those
//           with a readout sample or a complete schedule copy their state
//           to the host, and finished ones release their device memory;
//           the readouts themselves (samples, RunResult, waking waiters) run
//           on a second host thread while the device runs the next round.
//
// A request's result is independent of the batch it runs in and of how
// its MCS are split into rounds: exact replay reproduces the reference's
// serial stream whatever the chunking, and sublattice randomness is keyed
// per model and indexed by absolute MCS.
#include "serve.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "driver.hpp"

namespace cpm {

namespace {

struct Request {
    uint64_t id = 0;
    std::unique_ptr<Model> m;
    int mode = CPM_MODE_EXACT_REPLAY;
    bool include_lattice = false;
    std::vector<Step> sch;
    size_t step = 0;
    uint64_t in_step = 0;  // MCS done within sch[step]
    std::vector<double> trace;
    Result result;
    RequestInfo info;
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct Server {
    Engine* e = nullptr;
    ServeOptions opt;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::vector<Request*> active;  // touched by the worker only
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    uint64_t in_use = 0;
    bool stop = false;
    ServerStats st;
    std::thread worker;
    // the previous round's readouts, running beside the current round
    std::future<void> readouts;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void loop();
    void round();
    void wait_readouts() {
        if (readouts.valid()) readouts.get();  // the task reports its own failures
    }
};

void Server::loop() {
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
            if (stop) {
                lk.unlock();
                wait_readouts();
                return;
            }
            // admission in arrival order; a request that does not fit waits
            // for memory to free up (it fits the budget on its own)
            while (!queue.empty() && active.size() < opt.max_active) {
                Request* r = queue.front();
                if (in_use + r->info.footprint > opt.hbm_budget) break;
                queue.pop_front();
                in_use += r->info.footprint;
                r->info.state = CPM_REQ_ACTIVE;
                r->info.admit_s = now();
                active.push_back(r);
            }
            st.queued = queue.size();
            st.active = active.size();
            st.hbm_in_use = in_use;
        }
        if (!active.empty()) round();
    }
}

void Server::round() {
    const double t_start = now();
    const size_t n = active.size();
    // Exact replay: every request runs to the end of its schedule step or
    // for round_s, whichever comes first (the chain stops on the device's
    // clock, mid-MCS if need be).  Sublattice mode: quantum attempts.
    std::vector<uint64_t> todo(n), mcs_before(n);
    for (size_t i = 0; i < n; ++i) {
        Request* r = active[i];
        const uint64_t left = r->sch[r->step].mcs - r->in_step;
        mcs_before[i] = r->m->mcs;
        if (r->mode == CPM_MODE_EXACT_REPLAY) {
            todo[i] = left;
        } else {
            const double sites = (double)std::max<uint64_t>(1, r->m->n_sites);
            todo[i] = std::min(left, std::max<uint64_t>(1, (uint64_t)((double)opt.quantum / sites)));
        }
    }
    cpm_counters cnt{};
    std::string failure;
    try {
        for (int mode : {CPM_MODE_EXACT_REPLAY, CPM_MODE_SUBLATTICE}) {
            std::vector<Model*> ms;
            std::vector<uint64_t> td;
            for (size_t i = 0; i < n; ++i)
                if (active[i]->mode == mode) {
                    ms.push_back(active[i]->m.get());
                    td.push_back(todo[i]);
                }
            if (!ms.empty())
                run_mcs_each(e, ms.data(), ms.size(), td.data(), mode, &cnt,
                             mode == CPM_MODE_EXACT_REPLAY ? opt.round_s : 0.0);
        }
    } catch (const std::exception& ex) {
        failure = ex.what();
    }
    const double t_device = now();
    std::vector<Request*> finished, sampled;
    std::vector<uint64_t> at_mcs;
    std::vector<double> at_energy;
    std::vector<Sample> samples;
    // the previous round's readout thread may still be appending samples and
    // finishing results
    wait_readouts();
    std::vector<uint64_t> done_mcs(n, 0);  // whole MCS completed, published under the lock
    if (failure.empty()) {
        for (size_t i = 0; i < n; ++i) {
            Request* r = active[i];
            done_mcs[i] = r->m->mcs - mcs_before[i];
            r->in_step += done_mcs[i];
            if (r->in_step < r->sch[r->step].mcs) continue;
            const Step& s = r->sch[r->step];
            if (s.trace) r->trace.push_back(r->m->energy);
            if (s.sample) {
                sampled.push_back(r);
                at_mcs.push_back(r->m->mcs);
                at_energy.push_back(r->m->energy);
            }
            r->in_step = 0;
            if (++r->step == r->sch.size()) finished.push_back(r);
        }
        try {
            // readouts on the device, before finished requests leave it
            std::vector<Model*> sm(sampled.size());
            for (size_t j = 0; j < sampled.size(); ++j) sm[j] = sampled[j]->m.get();
            samples.resize(sampled.size());
            bool lattice_any = false;
            for (Request* r : sampled) lattice_any |= r->include_lattice;
            device_samples(e, sm.data(), sm.size(), lattice_any, at_mcs.data(), at_energy.data(),
                           samples.data());
            for (size_t j = 0; j < sampled.size(); ++j)
                if (!sampled[j]->include_lattice) {
                    samples[j].lattice.clear();
                    samples[j].has_lattice = false;
                }
            parallel_for(finished.size(), [&](size_t j) {
                sync_to_host(finished[j]->m.get());
                device_release(finished[j]->m.get());
            });
        } catch (const std::exception& ex) {
            failure = ex.what();
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    const double t_end = now();
    for (size_t i = 0; i < n; ++i) active[i]->info.mcs_done += done_mcs[i];
    st.rounds += 1;
    st.attempts += cnt.attempts;
    st.busy_s += t_end - t_start;
    st.device_s += t_device - t_start;
    if (!failure.empty()) {
        // every active request fails
        for (Request* r : active) {
            r->info.state = CPM_REQ_FAILED;
            r->info.error = failure;
            r->info.done_s = t_end;
            device_release(r->m.get());
            r->m.reset();
            in_use -= r->info.footprint;
            st.failed += 1;
        }
        active.clear();
        st.active = 0;
        st.hbm_in_use = in_use;
        cv_done.notify_all();
        return;
    }
    // finished requests leave the device now (their memory is free for the
    // next admission) and complete when their readouts do
    for (Request* r : finished) in_use -= r->info.footprint;
    std::vector<Request*> still;
    for (Request* r : active)
        if (r->step < r->sch.size()) still.push_back(r);
    active.swap(still);
    st.active = active.size();
    st.hbm_in_use = in_use;
    if (sampled.empty() && finished.empty()) return;
    readouts = std::async(std::launch::async, [this, sampled = std::move(sampled),
                                               finished = std::move(finished),
                                               samples = std::move(samples)]() mutable {
        std::string err;
        try {
            for (size_t j = 0; j < sampled.size(); ++j)
                sampled[j]->result.samples.push_back(std::move(samples[j]));
            parallel_for(finished.size(), [&](size_t j) {
                Request* r = finished[j];
                finish_result(*r->m, r->trace, r->result);
                r->m.reset();
            });
        } catch (const std::exception& ex) {
            err = ex.what();
        }
        std::lock_guard<std::mutex> lk2(mu);
        const double t_done = now();
        if (!err.empty())
            // a failed readout fails the requests it was for; a sampled
            // request that is still running fails when it finishes
            for (Request* r : sampled)
                if (r->info.state == CPM_REQ_ACTIVE && r->info.error.empty()) r->info.error = err;
        for (Request* r : finished) {
            if (err.empty() && r->info.error.empty()) {
                r->info.state = CPM_REQ_DONE;
                st.completed += 1;
            } else {
                r->info.state = CPM_REQ_FAILED;
                if (r->info.error.empty()) r->info.error = err;
                st.failed += 1;
            }
            r->info.done_s = t_done;
            r->m.reset();
        }
        if (!finished.empty()) cv_done.notify_all();
    });
}

Server* serve_create(Engine* e, const ServeOptions& o) {
    auto* s = new Server();
    s->e = e;
    s->opt = o;
    if (!s->opt.hbm_budget) s->opt.hbm_budget = (uint64_t)(0.8 * (double)device_free_bytes(e));
    if (!s->opt.max_active) s->opt.max_active = 4096;
    if (!s->opt.quantum) s->opt.quantum = 1ull << 18;
    if (!(s->opt.round_s > 0)) s->opt.round_s = 0.03;
    s->st.hbm_budget = s->opt.hbm_budget;
    s->t0 = Clock::now();
    s->worker = std::thread([s] { s->loop(); });
    return s;
}

void serve_destroy(Server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Request* r : s->active) device_release(r->m.get());
    delete s;
}

uint64_t serve_submit(Server* s, Model* m, bool include_lattice, int mode) {
    std::unique_ptr<Model> owned(m);
    if (mode != CPM_MODE_EXACT_REPLAY && mode != CPM_MODE_SUBLATTICE)
        throw Error(CPM_ERR_ARG, "unknown mode");
    if (mode == CPM_MODE_SUBLATTICE && !sublattice_supported(s->e, *m))
        throw Error(CPM_ERR_ARG,
                    "sublattice mode needs an LDS-resident lattice with even periodic sides");
    auto r = std::make_unique<Request>();
    r->sch = schedule(m->cfg);
    r->info.footprint = device_footprint(s->e, *m);
    for (const Step& st : r->sch) r->info.mcs_total += st.mcs;
    if (r->info.footprint > s->opt.hbm_budget)
        throw Error(CPM_ERR_OOM, "request needs " + std::to_string(r->info.footprint) +
                                     " bytes of device memory, above the server's budget of " +
                                     std::to_string(s->opt.hbm_budget));
    r->m = std::move(owned);
    r->mode = mode;
    r->include_lattice = include_lattice;
    uint64_t id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) throw Error(CPM_ERR_ARG, "server is shut down");
        id = r->id = s->next_id++;
        r->info.state = CPM_REQ_QUEUED;
        r->info.submit_s = s->now();
        s->queue.push_back(r.get());
        s->reqs.emplace(id, std::move(r));
        s->st.submitted += 1;
        s->st.queued = s->queue.size();
    }
    s->cv_work.notify_one();
    return id;
}

bool serve_poll(Server* s, uint64_t id, double timeout_s, RequestInfo& info, Result* out) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) throw Error(CPM_ERR_ARG, "unknown request id " + std::to_string(id));
    Request* r = it->second.get();
    auto finished = [&] {
        return s->stop || r->info.state == CPM_REQ_DONE || r->info.state == CPM_REQ_FAILED;
    };
    if (timeout_s > 0 && !finished())
        s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
    info = r->info;
    if (r->info.state == CPM_REQ_FAILED) {
        const std::string msg = r->info.error;
        s->reqs.erase(it);
        throw Error(CPM_ERR_DEVICE, msg);
    }
    if (r->info.state != CPM_REQ_DONE) return false;
    if (out) *out = std::move(r->result);
    s->reqs.erase(it);
    return true;
}

ServerStats serve_stats(Server* s) {
    std::lock_guard<std::mutex> lk(s->mu);
    ServerStats st = s->st;
    st.uptime_s = s->now();
    return st;
}

}  // namespace cpm
