// sgc_server.cpp -- the served pattern: a persistent request loop over a pool.
//
// Tenants submit update batches and colour queries for graphs resident in one
// sgc_pool, from any thread, as they arrive.  One loop thread drives the GPU:
// each turn it takes every request that has arrived (continuous batching),
// packs the updates of all tenants into one batch -- each graph filling at
// most up to its own epoch boundary, so every graph sees the epoch boundaries
// it would see alone -- applies the batch in one launch, and runs one shared
// epoch step for every graph that is due: its epoch is full, or its oldest
// unpublished update has waited max_delay_us.  A request completes when the
// snapshot holding its last update is published; its update-to-snapshot
// latency is publish time minus arrival time (the submission, or the
// scheduled arrival of a replayed trace), kept raw per request.
//
// The plain mode (SGC_SERVE_PLAIN) is the baseline: the loop serves the same
// requests as the plain per-launch engine would -- the earliest arrival in a
// launch of its own, one epoch step per due graph -- under the same epoch and
// deadline rules.
//
// Graphs are admitted against the pool's HBM budget by the loop: an admission
// that does not fit waits until removals free memory.  Queries read the
// published snapshot on the pool's query stream and never wait for the loop.
//
// Built on the pool's C ABI only (include/streamgc.h).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "streamgc.h"

namespace {

using Clock = std::chrono::steady_clock;

double us_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::micro>(b - a).count();
}

struct Request {
    uint64_t ticket = 0;
    uint32_t gid = 0;
    uint32_t applied = 0;
    Clock::time_point arrival;
    std::vector<sgc_edge_update> recs;  // graph id stamped
};

struct Admission {
    uint64_t ticket = 0;
    int64_t attach = -1;                     // >= 0: a graph already in the pool, served from now on
    sgc_config cfg{};
    uint32_t n = 0;
    std::vector<uint32_t> row_ptr, col, deg0;
};

struct Tenant {
    bool live = false;
    bool removing = false;
    uint32_t epoch_size = 0;
    uint64_t since = 0;                      // updates applied since its last epoch
    std::deque<std::unique_ptr<Request>> queue;   // not yet fully applied (front may be partly applied)
    std::vector<std::unique_ptr<Request>> applied;  // fully applied, waiting for a publish
    Clock::time_point oldest;                // arrival of the oldest unpublished update
    bool has_unpublished = false;
    uint64_t applied_total = 0;              // updates applied through the server, in the tenant's order
    std::vector<uint64_t> watermarks;        // applied_total at each publish
};

}  // namespace

struct sgc_server {
    sgc_pool* pool = nullptr;
    sgc_server_config cfg{};
    std::thread loop;
    std::string err;

    // submission side
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Request>> inbox;
    std::deque<std::unique_ptr<Admission>> admissions;
    std::vector<uint32_t> removals;
    bool stop = false;
    bool flush_req = false;
    bool admit_blocked = false;                // the oldest admission did not fit: retry after a removal
    uint64_t next_ticket = 1;
    uint64_t queued_updates = 0;
    std::map<uint64_t, int64_t> admitted;      // admission ticket -> gid (or negative error)
    std::map<uint64_t, bool> done;             // update tickets completed
    int loop_error = 0;

    // pool structure (tenant table) vs queries
    std::shared_mutex smu;
    std::deque<Tenant> tenants;            // indexed by pool graph id (a deque never relocates)

    // statistics (under mu)
    std::vector<double> latency_us;            // per request, update-to-snapshot
    std::vector<double> query_us;              // per query, API in to colours out
    uint64_t updates_applied = 0, batches = 0, epoch_steps = 0, graph_epochs = 0, requests_done = 0;
    Clock::time_point first_submit, last_publish;
    bool any_submit = false;
    double busy_us = 0;
    double intake_us = 0, pack_us = 0, apply_us = 0, epoch_us = 0;   // loop thread only until read under mu

    void run();
    int turn();
    int intake(bool& flush, bool& work);
    void take_front(uint32_t g, uint32_t take, std::vector<sgc_edge_update>& batch);
    int launch(const std::vector<sgc_edge_update>& batch, const std::vector<uint32_t>& taken);
    int apply_batched(bool& work);
    int apply_plain(bool& work);
    int publish_due(bool flush, bool& work);
    int submit_at(uint32_t gid, const sgc_edge_update* upd, uint32_t count, Clock::time_point arrival,
                  uint64_t* ticket);
    int query_at(uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out, Clock::time_point arrival);
};

// Admissions, removals and the routing of arrived requests to their tenants.
int sgc_server::intake(bool& flush, bool& work) {
    std::vector<std::unique_ptr<Request>> fresh;
    std::vector<std::unique_ptr<Admission>> adm;
    std::vector<uint32_t> rem;
    {
        std::lock_guard<std::mutex> lk(mu);
        while (!inbox.empty()) {
            fresh.push_back(std::move(inbox.front()));
            inbox.pop_front();
        }
        if (!admit_blocked || !removals.empty()) {
            while (!admissions.empty()) {
                adm.push_back(std::move(admissions.front()));
                admissions.pop_front();
            }
        }
        rem.swap(removals);
        flush = flush_req;
    }
    work = !fresh.empty();
    // route arrivals to their tenants
    for (auto& r : fresh) {
        Tenant& T = tenants[r->gid];
        if (!T.has_unpublished) {
            T.oldest = r->arrival;
            T.has_unpublished = true;
        }
        T.queue.push_back(std::move(r));
    }
    // removals wait until the tenant has nothing pending
    for (uint32_t g : rem) tenants[g].removing = true;
    for (uint32_t g = 0; g < tenants.size(); g++) {
        Tenant& T = tenants[g];
        if (!T.live || !T.removing || !T.queue.empty() || !T.applied.empty()) continue;
        std::unique_lock<std::shared_mutex> wl(smu);
        int r = sgc_pool_remove(pool, g);
        if (r) return r;
        T = Tenant();
        work = true;
        std::lock_guard<std::mutex> lk(mu);
        admit_blocked = false;
    }
    // admissions in arrival order while they fit
    std::vector<std::unique_ptr<Admission>> keep;
    for (size_t k = 0; k < adm.size(); k++) {
        Admission& A = *adm[k];
        if (!keep.empty()) {
            keep.push_back(std::move(adm[k]));
            continue;
        }
        uint32_t gid = 0;
        int r;
        {
            std::unique_lock<std::shared_mutex> wl(smu);
            if (A.attach >= 0) {
                gid = (uint32_t)A.attach;
                uint32_t hc = 0;   // fails for a graph id the pool does not hold
                r = sgc_pool_history_count(pool, gid, &hc);
                if (r == SGC_OK && gid < tenants.size() && tenants[gid].live) r = SGC_ERR_INVALID;
            } else {
                r = sgc_pool_add(pool, &A.cfg, A.n, A.row_ptr.data(), A.col.empty() ? nullptr : A.col.data(),
                                 A.deg0.data(), &gid);
            }
            if (r == SGC_OK) {
                if (tenants.size() <= gid) tenants.resize(gid + 1);
                tenants[gid] = Tenant();
                tenants[gid].live = true;
                tenants[gid].epoch_size = A.cfg.epoch_size;
            }
        }
        if (r == SGC_ERR_NOMEM) {
            keep.push_back(std::move(adm[k]));   // waits for memory
            continue;
        }
        std::lock_guard<std::mutex> lk(mu);
        admitted[A.ticket] = r == SGC_OK ? (int64_t)gid : (int64_t)r;
        work = true;
    }
    if (!adm.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        admit_blocked = !keep.empty();
        for (auto it = keep.rbegin(); it != keep.rend(); ++it) admissions.push_front(std::move(*it));
    }
    if (!adm.empty()) cv_done.notify_all();
    return SGC_OK;
}

// Takes up to `take` more updates of tenant g's front request into `batch`.
void sgc_server::take_front(uint32_t g, uint32_t take, std::vector<sgc_edge_update>& batch) {
    Tenant& T = tenants[g];
    Request& R = *T.queue.front();
    batch.insert(batch.end(), R.recs.begin() + R.applied, R.recs.begin() + R.applied + take);
    R.applied += take;
    if (R.applied == R.recs.size()) {
        T.applied.push_back(std::move(T.queue.front()));
        T.queue.pop_front();
    }
}

// One launch over `batch`, then the tenants' update counts move on.
int sgc_server::launch(const std::vector<sgc_edge_update>& batch, const std::vector<uint32_t>& taken) {
    const auto t0 = Clock::now();
    int r = sgc_pool_apply(pool, batch.data(), batch.size(), 0);
    if (r) return r;
    const double dt = us_between(t0, Clock::now());
    for (uint32_t g = 0; g < tenants.size(); g++) {
        tenants[g].since += taken[g];
        tenants[g].applied_total += taken[g];
    }
    std::lock_guard<std::mutex> lk(mu);
    updates_applied += batch.size();
    batches++;
    queued_updates -= batch.size();
    apply_us += dt;
    return SGC_OK;
}

// Continuous batching: one packed batch, round-robin over the tenants, each
// filled at most up to its epoch boundary.
int sgc_server::apply_batched(bool& work) {
    std::vector<sgc_edge_update> batch;
    std::vector<uint32_t> taken(tenants.size(), 0);
    const uint32_t cap = cfg.max_batch;
    bool progress = true;
    while (batch.size() < cap && progress) {
        progress = false;
        for (uint32_t g = 0; g < tenants.size() && batch.size() < cap; g++) {
            Tenant& T = tenants[g];
            if (!T.live || T.queue.empty()) continue;
            const uint64_t room = T.epoch_size - std::min<uint64_t>(T.since + taken[g], T.epoch_size);
            if (room == 0) continue;
            const Request& R = *T.queue.front();
            const uint32_t left = (uint32_t)R.recs.size() - R.applied;
            const uint32_t take = (uint32_t)std::min<uint64_t>({(uint64_t)left, room, (uint64_t)(cap - batch.size())});
            take_front(g, take, batch);
            taken[g] += take;
            progress = true;
        }
    }
    if (batch.empty()) return SGC_OK;
    work = true;
    return launch(batch, taken);
}

// The plain per-launch engine: the request that arrived first (of the tenants
// with room left in their epoch) in one launch of its own, up to its tenant's
// epoch boundary.
int sgc_server::apply_plain(bool& work) {
    int64_t best = -1;
    for (uint32_t g = 0; g < tenants.size(); g++) {
        const Tenant& T = tenants[g];
        if (!T.live || T.queue.empty() || T.since >= T.epoch_size) continue;
        const Request& R = *T.queue.front();
        if (best < 0) {
            best = g;
            continue;
        }
        const Request& B = *tenants[best].queue.front();
        if (R.arrival < B.arrival || (R.arrival == B.arrival && R.ticket < B.ticket)) best = g;
    }
    if (best < 0) return SGC_OK;
    const uint32_t g = (uint32_t)best;
    Tenant& T = tenants[g];
    const Request& R = *T.queue.front();
    const uint32_t take = (uint32_t)std::min<uint64_t>({(uint64_t)(R.recs.size() - R.applied),
                                                         T.epoch_size - T.since, (uint64_t)cfg.max_batch});
    std::vector<sgc_edge_update> batch;
    std::vector<uint32_t> taken(tenants.size(), 0);
    take_front(g, take, batch);
    taken[g] = take;
    work = true;
    return launch(batch, taken);
}

// The epoch steps of the due graphs (epoch full, oldest unpublished update
// older than max_delay_us, or a flush): one shared step (batched) or one step
// per graph (plain).  sgc_pool_epoch returns after the pool's stream has
// synchronised, so the snapshot kernel of the publish has completed when the
// completion time is taken.
int sgc_server::publish_due(bool flush, bool& work) {
    const auto now = Clock::now();
    std::vector<uint32_t> due;
    for (uint32_t g = 0; g < tenants.size(); g++) {
        Tenant& T = tenants[g];
        if (!T.live || T.since == 0) continue;
        const bool full = T.since >= T.epoch_size;
        const bool stale = cfg.max_delay_us && T.has_unpublished && us_between(T.oldest, now) >= cfg.max_delay_us;
        const bool flushing = (flush || T.removing) && T.queue.empty();
        if (full || stale || flushing) due.push_back(g);
    }
    if (due.empty()) return SGC_OK;
    work = true;
    const bool plain = cfg.mode == SGC_SERVE_PLAIN;
    for (size_t k = 0; k < due.size(); k += plain ? 1 : due.size()) {
        const uint32_t cnt = plain ? 1u : (uint32_t)due.size();
        const auto te = Clock::now();
        int r = sgc_pool_epoch(pool, due.data() + k, cnt, nullptr);
        if (r) return r;
        const auto tp = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        epoch_us += us_between(te, tp);
        for (uint32_t j = 0; j < cnt; j++) {
            Tenant& T = tenants[due[k + j]];
            T.since = 0;
            T.watermarks.push_back(T.applied_total);
            for (auto& R : T.applied) {
                latency_us.push_back(us_between(R->arrival, tp));
                done[R->ticket] = true;
                requests_done++;
            }
            T.applied.clear();
            // the front request may be partly applied: its updates are not all published
            T.has_unpublished = !T.queue.empty();
            if (T.has_unpublished) T.oldest = T.queue.front()->arrival;
        }
        epoch_steps++;
        graph_epochs += cnt;
        last_publish = tp;
    }
    cv_done.notify_all();
    return SGC_OK;
}

// One loop turn: admissions and removals, one launch (packed, or one request
// in plain mode), the epoch steps of the due graphs.  Returns 1 if it did
// work, 0 if idle, negative on error.
int sgc_server::turn() {
    bool flush = false, work = false;
    const auto t0 = Clock::now();
    int r = intake(flush, work);
    const auto t1 = Clock::now();
    double applied0;
    {
        std::lock_guard<std::mutex> lk(mu);
        applied0 = apply_us;
    }
    if (r == SGC_OK) r = cfg.mode == SGC_SERVE_PLAIN ? apply_plain(work) : apply_batched(work);
    const auto t2 = Clock::now();
    if (r == SGC_OK) r = publish_due(flush, work);
    if (r) return r;
    if (work) {
        std::lock_guard<std::mutex> lk(mu);
        busy_us += us_between(t0, Clock::now());
        intake_us += us_between(t0, t1);
        pack_us += us_between(t1, t2) - (apply_us - applied0);   // the launch call itself is apply_us
    }
    return work ? 1 : 0;
}

void sgc_server::run() {
    for (;;) {
        int r = turn();
        if (r < 0) {
            std::lock_guard<std::mutex> lk(mu);
            loop_error = r;
            err = std::string("server loop: ") + sgc_pool_last_error(pool);
            stop = true;
            cv_done.notify_all();
            return;
        }
        std::unique_lock<std::mutex> lk(mu);
        bool pending = false;
        for (auto& T : tenants)
            if (T.live && (!T.queue.empty() || T.since)) pending = true;
        if (stop && admit_blocked) {   // stopping: admissions still waiting for memory fail
            for (auto& A : admissions) admitted[A->ticket] = SGC_ERR_NOMEM;
            admissions.clear();
            cv_done.notify_all();
        }
        if (stop && inbox.empty() && !pending && admissions.empty() && removals.empty()) return;
        if (r == 0) {
            if (flush_req && !pending) {
                flush_req = false;
                cv_done.notify_all();
            }
            // idle: wait for arrivals, or for the next stale-update deadline
            auto wake = std::chrono::microseconds(cfg.max_delay_us ? std::max<uint32_t>(cfg.max_delay_us / 4, 50) : 2000);
            cv_work.wait_for(lk, wake, [&] {
                return stop || !inbox.empty() || !removals.empty() || flush_req ||
                       (!admissions.empty() && !admit_blocked);
            });
        }
    }
}

int sgc_server::submit_at(uint32_t gid, const sgc_edge_update* upd, uint32_t count, Clock::time_point arrival,
                          uint64_t* ticket) {
    auto R = std::make_unique<Request>();
    R->gid = gid;
    R->arrival = arrival;
    R->recs.assign(upd, upd + count);
    for (auto& r : R->recs) memcpy(r.pad + 3, &gid, 4);
    {
        std::shared_lock<std::shared_mutex> rl(smu);
        if (gid >= tenants.size() || !tenants[gid].live || tenants[gid].removing) return SGC_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (stop || loop_error) return SGC_ERR_STATE;
    R->ticket = next_ticket++;
    *ticket = R->ticket;
    if (!any_submit) {
        first_submit = R->arrival;
        any_submit = true;
    }
    queued_updates += count;
    inbox.push_back(std::move(R));
    cv_work.notify_all();
    return SGC_OK;
}

// A query's latency runs from its arrival to the colours being out.
int sgc_server::query_at(uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out,
                         Clock::time_point arrival) {
    int r;
    {
        std::shared_lock<std::shared_mutex> rl(smu);
        if (gid >= tenants.size() || !tenants[gid].live) return SGC_ERR_INVALID;
        r = sgc_pool_query(pool, gid, ids, count, out);
    }
    const double us = us_between(arrival, Clock::now());
    std::lock_guard<std::mutex> lk(mu);
    query_us.push_back(us);
    return r;
}

extern "C" {

void sgc_default_server_config(sgc_server_config* c) {
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->max_batch = 32768;
    c->max_delay_us = 0;
}

int sgc_server_create(sgc_pool* pool, const sgc_server_config* cfg, sgc_server** out) {
    if (!pool || !cfg || !out || cfg->max_batch == 0) return SGC_ERR_INVALID;
    if (cfg->mode != SGC_SERVE_BATCHED && cfg->mode != SGC_SERVE_PLAIN) return SGC_ERR_INVALID;
    sgc_server* s = new sgc_server;
    s->pool = pool;
    s->cfg = *cfg;
    s->loop = std::thread([s] { s->run(); });
    *out = s;
    return SGC_OK;
}

int sgc_server_destroy(sgc_server* s) {
    if (!s) return SGC_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->flush_req = true;
    }
    s->cv_work.notify_all();
    if (s->loop.joinable()) s->loop.join();
    const int r = s->loop_error;
    delete s;
    return r;
}

const char* sgc_server_last_error(const sgc_server* s) { return s ? s->err.c_str() : ""; }

int sgc_server_admit(sgc_server* s, const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr,
                     const uint32_t* col_idx, const uint32_t* deg0, uint64_t* ticket) {
    if (!s || !cfg || !row_ptr || !deg0 || !ticket || n == 0 || cfg->epoch_size == 0) return SGC_ERR_INVALID;
    auto A = std::make_unique<Admission>();
    A->cfg = *cfg;
    A->n = n;
    A->row_ptr.assign(row_ptr, row_ptr + n + 1);
    const uint64_t nnz = row_ptr[n];
    if (nnz && !col_idx) return SGC_ERR_INVALID;
    if (nnz) A->col.assign(col_idx, col_idx + nnz);
    A->deg0.assign(deg0, deg0 + n);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return SGC_ERR_STATE;
    if (s->admissions.empty()) s->admit_blocked = false;   // a new head of the line gets a try
    A->ticket = s->next_ticket++;
    *ticket = A->ticket;
    s->admissions.push_back(std::move(A));
    s->cv_work.notify_all();
    return SGC_OK;
}

int sgc_server_attach(sgc_server* s, uint32_t gid, uint32_t epoch_size, uint64_t* ticket) {
    if (!s || !ticket || epoch_size == 0) return SGC_ERR_INVALID;
    auto A = std::make_unique<Admission>();
    A->attach = gid;
    A->cfg.epoch_size = epoch_size;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return SGC_ERR_STATE;
    A->ticket = s->next_ticket++;
    *ticket = A->ticket;
    s->admissions.push_back(std::move(A));
    s->cv_work.notify_all();
    return SGC_OK;
}

int sgc_server_admitted(sgc_server* s, uint64_t ticket, uint32_t timeout_ms, uint32_t* gid) {
    if (!s || !gid) return SGC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] {
        return s->admitted.count(ticket) || s->loop_error;
    });
    if (s->loop_error) return s->loop_error;
    if (!ok) return SGC_ERR_STATE;
    const int64_t v = s->admitted[ticket];
    if (v < 0) return (int)v;
    *gid = (uint32_t)v;
    return SGC_OK;
}

int sgc_server_remove(sgc_server* s, uint32_t gid) {
    if (!s) return SGC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->removals.push_back(gid);
    s->cv_work.notify_all();
    return SGC_OK;
}

int sgc_server_submit(sgc_server* s, uint32_t gid, const sgc_edge_update* upd, uint32_t count, uint64_t* ticket) {
    if (!s || (!upd && count) || !ticket || count == 0) return SGC_ERR_INVALID;
    return s->submit_at(gid, upd, count, Clock::now(), ticket);
}

int sgc_server_wait(sgc_server* s, uint64_t ticket, uint32_t timeout_ms) {
    if (!s) return SGC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] {
        return s->done.count(ticket) || s->loop_error;
    });
    if (s->loop_error) return s->loop_error;
    return ok ? SGC_OK : SGC_ERR_STATE;
}

int sgc_server_flush(sgc_server* s, uint32_t timeout_ms) {
    if (!s) return SGC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->flush_req = true;
    s->cv_work.notify_all();
    const bool ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] {
        return (!s->flush_req && s->inbox.empty()) || s->loop_error;
    });
    if (s->loop_error) return s->loop_error;
    return ok ? SGC_OK : SGC_ERR_STATE;
}

int sgc_server_query(sgc_server* s, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out) {
    if (!s) return SGC_ERR_INVALID;
    return s->query_at(gid, ids, count, out, Clock::now());
}

int sgc_server_get_stats(sgc_server* s, sgc_server_stats* out) {
    if (!s || !out) return SGC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof *out);
    out->updates_applied = s->updates_applied;
    out->batches = s->batches;
    out->epoch_steps = s->epoch_steps;
    out->graph_epochs = s->graph_epochs;
    out->requests_done = s->requests_done;
    out->queries = s->query_us.size();
    out->queued_updates = s->queued_updates;
    out->span_us = s->any_submit && s->epoch_steps ? us_between(s->first_submit, s->last_publish) : 0;
    out->busy_us = s->busy_us;
    out->intake_us = s->intake_us;
    out->pack_us = s->pack_us;
    out->apply_us = s->apply_us;
    out->epoch_us = s->epoch_us;
    return SGC_OK;
}

int sgc_server_latencies(sgc_server* s, int which, double* out, uint64_t cap, uint64_t* count) {
    if (!s || !count || (which != 0 && which != 1)) return SGC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    const std::vector<double>& v = which == 0 ? s->latency_us : s->query_us;
    *count = v.size();
    if (out) std::copy(v.begin(), v.begin() + std::min<uint64_t>(cap, v.size()), out);
    return SGC_OK;
}

int sgc_server_watermarks(sgc_server* s, uint32_t gid, uint64_t* out, uint64_t cap, uint64_t* count) {
    if (!s || !count) return SGC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (gid >= s->tenants.size() || !s->tenants[gid].live) return SGC_ERR_INVALID;
    const std::vector<uint64_t>& w = s->tenants[gid].watermarks;
    *count = w.size();
    if (out) std::copy(w.begin(), w.begin() + std::min<uint64_t>(cap, w.size()), out);
    return SGC_OK;
}

int sgc_server_reset_stats(sgc_server* s) {
    if (!s) return SGC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->latency_us.clear();
    s->query_us.clear();
    s->updates_applied = s->batches = s->epoch_steps = s->graph_epochs = s->requests_done = 0;
    s->any_submit = false;
    s->busy_us = 0;
    return SGC_OK;
}

// Open-loop replay of an arrival trace: update items from this thread,
// queries from a second one, each submitted at its scheduled time (the last
// stretch of each wait spins, so submissions run at most a few microseconds
// late unless the host falls behind), with its scheduled time as arrival.
int sgc_server_replay(sgc_server* s, const sgc_trace_item* items, uint64_t count, const sgc_edge_update* updates,
                      const uint32_t* query_ids, uint32_t timeout_ms, sgc_replay_stats* out) {
    if (!s || (!items && count)) return SGC_ERR_INVALID;
    uint32_t max_q = 0;
    for (uint64_t i = 0; i < count; i++) {
        const sgc_trace_item& it = items[i];
        if (it.kind > 1 || it.count == 0) return SGC_ERR_INVALID;
        if (it.kind == 0 && !updates) return SGC_ERR_INVALID;
        if (it.kind == 1 && !query_ids) return SGC_ERR_INVALID;
        if (it.kind == 1) max_q = std::max(max_q, it.count);
    }
    sgc_replay_stats st{};
    const auto start = Clock::now() + std::chrono::microseconds(200);
    auto wait_until = [](Clock::time_point t) {
        const auto coarse = t - std::chrono::microseconds(100);
        if (Clock::now() < coarse) std::this_thread::sleep_until(coarse);
        while (Clock::now() < t) {
        }
    };
    std::atomic<int> qerr{0};
    std::atomic<uint64_t> nq{0};
    std::thread qthread([&] {
        std::vector<uint16_t> colours(std::max<uint32_t>(max_q, 1));
        for (uint64_t i = 0; i < count && !qerr.load(); i++) {
            const sgc_trace_item& it = items[i];
            if (it.kind != 1) continue;
            const auto at = start + std::chrono::nanoseconds(it.at_ns);
            wait_until(at);
            const int r = s->query_at(it.gid, query_ids + it.offset, it.count, colours.data(), at);
            if (r) qerr = r;
            nq++;
        }
    });
    int err = SGC_OK;
    double late_sum = 0;
    for (uint64_t i = 0; i < count && err == SGC_OK; i++) {
        const sgc_trace_item& it = items[i];
        if (it.kind != 0) continue;
        const auto at = start + std::chrono::nanoseconds(it.at_ns);
        wait_until(at);
        const double late = us_between(at, Clock::now());
        st.max_late_us = std::max(st.max_late_us, late);
        late_sum += late;
        uint64_t ticket = 0;
        err = s->submit_at(it.gid, updates + it.offset, it.count, at, &ticket);
        st.requests++;
        st.updates += it.count;
    }
    qthread.join();
    if (err == SGC_OK) err = qerr.load();
    // the last requests are published by the deadline rule as in the rest of
    // the trace (a flush right away would publish them early); then a flush
    // publishes whatever is left
    if (err == SGC_OK && count) {
        uint64_t last = 0;
        for (uint64_t i = 0; i < count; i++) last = std::max(last, items[i].at_ns);
        wait_until(start + std::chrono::nanoseconds(last) + std::chrono::microseconds(s->cfg.max_delay_us));
    }
    if (err == SGC_OK) err = sgc_server_flush(s, timeout_ms);
    if (err) return err;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        st.span_us = us_between(start, s->last_publish);
    }
    st.mean_late_us = st.requests ? late_sum / (double)st.requests : 0;
    st.queries = nq.load();
    if (out) *out = st;
    return SGC_OK;
}

}  // extern "C"
