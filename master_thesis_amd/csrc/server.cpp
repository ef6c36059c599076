// server.cpp -- serving of world-generation jobs (tt_server_* in tectonic.h).
//
// Jobs (a world configuration and a step count) are submitted from any thread
// and return an id at once.  One scheduler thread owns the device work:
//   * admission: jobs are taken in submission order while the sum of
//     tt_world_bytes() of the admitted worlds stays within the HBM budget and
//     fewer than max_active worlds are resident (the head of the queue waits
//     until it fits, so jobs are served in order);
//   * an admitted job's world is created and initialised, then joins the
//     server's batch at the next step boundary and stays device-resident until
//     its last step;
//   * the batch advances all resident worlds one step at a time; a world that
//     has run its steps leaves the batch at that step boundary (its slot and
//     bytes are free for the next job), so jobs that arrive while others run
//     overlap with them on the device (continuous batching);
//   * latency = completion time - submission time, on the server's clock.
// A job's result is the same alone and in any mix: batched stepping is
// bit-identical to stepping the world by itself.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tectonic.h"

void tt_set_error(const std::string &msg);
std::string tt_get_error();

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    tt_sim_config cfg{};
    tt_kernel_settings ks{};
    int32_t steps = 0;
    int32_t done = 0;
    size_t bytes = 0;
    tt_world *world = nullptr;
    tt_job_result res{};
};

}  // namespace

struct tt_server {
    std::mutex mu;
    std::condition_variable cv;       // scheduler wake-up
    std::condition_variable doneCv;   // result waiters
    std::deque<Job *> pending;
    std::vector<Job *> active;
    std::map<int64_t, Job *> jobs;    // every job ever submitted (results live here)
    tt_batch *batch = nullptr;
    std::thread sched;
    bool stop = false;
    bool keep = false;
    int64_t nextId = 1;
    uint64_t budget = 0;
    int32_t maxActive = 256;
    size_t inUse = 0, peak = 0;
    int64_t batchSteps = 0;
    Clock::time_point t0 = Clock::now();
    bool anySubmit = false;
    double firstSubmit = 0.0, lastDone = 0.0;
    int64_t submitted = 0, admitted = 0, completed = 0, failed = 0, rejected = 0;
    double cellUpdates = 0.0;
    int32_t schedStatus = TT_OK;
    std::string schedError;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void finish(tt_server *s, Job *j, int32_t status)
{
    // called with the lock held
    j->res.status = status;
    j->res.t_done = s->now();
    s->lastDone = std::max(s->lastDone, j->res.t_done);
    if (status == TT_OK) {
        ++s->completed;
        s->cellUpdates += (double)j->cfg.width * (double)j->cfg.height * (double)j->steps;
    } else {
        ++s->failed;
    }
    s->inUse -= j->bytes;
    s->doneCv.notify_all();
}

void scheduler(tt_server *s)
{
    for (;;) {
        std::vector<Job *> admit;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            s->cv.wait(lock, [&] { return s->stop || !s->pending.empty() || !s->active.empty(); });
            if (s->stop) break;
            while (!s->pending.empty() && (int32_t)(s->active.size() + admit.size()) < s->maxActive &&
                   s->inUse + s->pending.front()->bytes <= s->budget) {
                Job *j = s->pending.front();
                s->pending.pop_front();
                s->inUse += j->bytes;
                s->peak = std::max(s->peak, s->inUse);
                admit.push_back(j);
            }
        }
        // world creation and initialisation outside the lock
        for (Job *j : admit) {
            int st = tt_world_create(&j->cfg, &j->ks, &j->world);
            if (st == TT_OK) st = tt_world_init(j->world);
            if (st == TT_OK && j->steps > 0) st = tt_batch_add(s->batch, j->world);
            std::lock_guard<std::mutex> lock(s->mu);
            j->res.t_admit = s->now();
            ++s->admitted;
            if (st != TT_OK) {
                if (j->world) tt_world_destroy(j->world);
                j->world = nullptr;
                finish(s, j, st);
            } else if (j->steps == 0) {
                tt_step_stats ss;
                tt_world_stats(j->world, &ss);
                j->res.iteration = ss.iteration;
                j->res.mass_ledger = ss.mass_ledger;
                if (!s->keep) {
                    tt_world_destroy(j->world);
                    j->world = nullptr;
                }
                finish(s, j, TT_OK);
            } else {
                s->active.push_back(j);
            }
        }
        std::vector<Job *> act;
        {
            std::lock_guard<std::mutex> lock(s->mu);
            act = s->active;
        }
        if (act.empty()) continue;
        const int st = tt_batch_step(s->batch, 1);
        std::vector<Job *> leaving;
        {
            std::lock_guard<std::mutex> lock(s->mu);
            ++s->batchSteps;
            if (st != TT_OK) {
                s->schedStatus = st;
                s->schedError = tt_get_error();
            }
            for (Job *j : act) {
                ++j->done;
                if (st != TT_OK || j->done >= j->steps) leaving.push_back(j);
            }
        }
        for (Job *j : leaving) {
            tt_batch_remove(s->batch, j->world);
            tt_step_stats ss{};
            tt_world_stats(j->world, &ss);
            std::lock_guard<std::mutex> lock(s->mu);
            s->active.erase(std::find(s->active.begin(), s->active.end(), j));
            j->res.iteration = ss.iteration;
            j->res.mass_ledger = ss.mass_ledger;
            if (!s->keep || st != TT_OK) {
                tt_world_destroy(j->world);
                j->world = nullptr;
            }
            finish(s, j, st);
        }
    }
    // shutdown: release resident worlds, fail queued jobs
    std::lock_guard<std::mutex> lock(s->mu);
    for (Job *j : s->active) {
        tt_batch_remove(s->batch, j->world);
        tt_world_destroy(j->world);
        j->world = nullptr;
        finish(s, j, TT_ERR_STATE);
    }
    s->active.clear();
    for (Job *j : s->pending) {
        s->inUse += j->bytes;   // finish() subtracts it
        finish(s, j, TT_ERR_STATE);
    }
    s->pending.clear();
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest-rank
    size_t r = (size_t)std::ceil(q * (double)v.size());
    if (r < 1) r = 1;
    if (r > v.size()) r = v.size();
    return v[r - 1];
}

}  // namespace

extern "C" {

int tt_server_create(uint64_t budget_bytes, int32_t max_active, tt_server **out)
{
    if (!out || budget_bytes == 0 || max_active < 1) return TT_ERR_INVALID;
    *out = nullptr;
    tt_server *s = new (std::nothrow) tt_server();
    if (!s) return TT_ERR_NOMEM;
    s->budget = budget_bytes;
    s->maxActive = std::min<int32_t>(max_active, 4096);
    const int st = tt_batch_create(&s->batch);
    if (st != TT_OK) {
        delete s;
        return st;
    }
    s->sched = std::thread(scheduler, s);
    *out = s;
    return TT_OK;
}

int tt_server_set_keep(tt_server *s, int32_t keep)
{
    if (!s) return TT_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    s->keep = keep != 0;
    return TT_OK;
}

int tt_server_submit(tt_server *s, const tt_sim_config *cfg, const tt_kernel_settings *ks, int32_t steps, int64_t *id)
{
    if (!s || !cfg || !ks || !id || steps < 0) return TT_ERR_INVALID;
    size_t bytes = 0;
    const int st = tt_world_bytes(cfg, &bytes);
    if (st != TT_OK) return st;
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stop) return TT_ERR_STATE;
    if (bytes > s->budget) {
        ++s->rejected;
        tt_set_error("request needs " + std::to_string(bytes) + " device bytes, budget " + std::to_string(s->budget));
        return TT_ERR_BUDGET;
    }
    Job *j = new Job();
    j->id = s->nextId++;
    j->cfg = *cfg;
    j->ks = *ks;
    j->steps = steps;
    j->bytes = bytes;
    j->res.id = j->id;
    j->res.status = 1;   // not finished
    j->res.t_submit = s->now();
    j->res.bytes = (int64_t)bytes;
    j->res.width = cfg->width;
    j->res.height = cfg->height;
    j->res.steps = steps;
    if (!s->anySubmit) {
        s->anySubmit = true;
        s->firstSubmit = j->res.t_submit;
    }
    ++s->submitted;
    s->jobs[j->id] = j;
    s->pending.push_back(j);
    *id = j->id;
    s->cv.notify_all();
    return TT_OK;
}

int tt_server_result(tt_server *s, int64_t id, int32_t wait_ms, tt_job_result *out)
{
    if (!s || !out) return TT_ERR_INVALID;
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return TT_ERR_INVALID;
    Job *j = it->second;
    if (j->res.status == 1 && wait_ms != 0) {
        auto pred = [&] { return j->res.status != 1; };
        if (wait_ms < 0) s->doneCv.wait(lock, pred);
        else s->doneCv.wait_for(lock, std::chrono::milliseconds(wait_ms), pred);
    }
    *out = j->res;
    return j->res.status == 1 ? TT_ERR_STATE : TT_OK;
}

int tt_server_take(tt_server *s, int64_t id, tt_world **w)
{
    if (!s || !w) return TT_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || it->second->res.status != TT_OK || !it->second->world) return TT_ERR_STATE;
    *w = it->second->world;
    it->second->world = nullptr;
    return TT_OK;
}

int tt_server_drain(tt_server *s, int32_t timeout_ms)
{
    if (!s) return TT_ERR_INVALID;
    std::unique_lock<std::mutex> lock(s->mu);
    auto pred = [&] { return s->pending.empty() && s->active.empty() && s->completed + s->failed == s->submitted; };
    if (timeout_ms < 0) s->doneCv.wait(lock, pred);
    else if (!s->doneCv.wait_for(lock, std::chrono::milliseconds(timeout_ms), pred)) return TT_ERR_STATE;
    if (s->schedStatus != TT_OK) {
        tt_set_error(s->schedError);
        return s->schedStatus;
    }
    return TT_OK;
}

int tt_server_stats(tt_server *s, tt_server_report *out)
{
    if (!s || !out) return TT_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    tt_server_report r{};
    r.submitted = s->submitted;
    r.admitted = s->admitted;
    r.completed = s->completed;
    r.failed = s->failed;
    r.rejected = s->rejected;
    r.active = (int64_t)s->active.size();
    r.pending = (int64_t)s->pending.size();
    r.bytes_in_use = (int64_t)s->inUse;
    r.bytes_peak = (int64_t)s->peak;
    r.budget = (int64_t)s->budget;
    r.batch_steps = s->batchSteps;
    std::vector<double> lat;
    double sum = 0.0, mx = 0.0;
    for (const auto &kv : s->jobs) {
        const tt_job_result &x = kv.second->res;
        if (x.status != TT_OK) continue;
        const double l = x.t_done - x.t_submit;
        lat.push_back(l);
        sum += l;
        mx = std::max(mx, l);
    }
    r.wall_s = s->anySubmit ? std::max(0.0, s->lastDone - s->firstSubmit) : 0.0;
    r.worlds_per_s = r.wall_s > 0.0 ? (double)s->completed / r.wall_s : 0.0;
    r.latency_p50_s = percentile(lat, 0.50);
    r.latency_p99_s = percentile(lat, 0.99);
    r.latency_mean_s = lat.empty() ? 0.0 : sum / (double)lat.size();
    r.latency_max_s = mx;
    r.cell_updates = s->cellUpdates;
    *out = r;
    return TT_OK;
}

int tt_server_destroy(tt_server *s)
{
    if (!s) return TT_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->sched.joinable()) s->sched.join();
    for (auto &kv : s->jobs) {
        if (kv.second->world) tt_world_destroy(kv.second->world);
        delete kv.second;
    }
    if (s->batch) tt_batch_destroy(s->batch);
    delete s;
    return TT_OK;
}

}  // extern "C"
