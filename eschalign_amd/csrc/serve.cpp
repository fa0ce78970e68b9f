// Serving loop: resident panels, continuous batching, admission by HBM
// footprint (no counterpart in the reference, whose CLI scores one batch at
// a time on one GPU thread, crates/eschalign-cli/src/align.rs:2801-2917).
//
// Requests (reads against one resident panel) are queued as units of at
// most one batch each. A scheduler thread packs queued units that share a
// panel and options into a device batch whenever a slot is free and keeps
// up to max_slots batches in flight, each on its own stream with its own
// scratch and captured graph, so their work overlaps on the device. A new
// slot is created only while the summed footprint of all slots stays within
// the HBM budget; idle slots of other panels are released when memory is
// needed. Scores per read do not depend on what shares its batch, so every
// result equals the result of the request run alone.
#include "../../include/eschalign_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    int panel = 0;
    int both = 1;
    int32_t min_score = 0;
    int64_t max_ties = -1;
    size_t tie_cap = 1;
    size_t n = 0;
    std::vector<uint8_t> codes;
    std::vector<uint64_t> off;
    std::vector<int32_t> len;
    std::vector<esch_read_select> recs;
    std::vector<int32_t> ties;
    size_t parts_left = 0;
    bool done = false;
    esch_status status = ESCH_OK;
    std::string error;
    Clock::time_point t_submit, t_done;
};

struct Unit {  // reads [a, b) of one request
    std::shared_ptr<Request> req;
    size_t a, b;
    size_t bases() const {
        size_t s = 0;
        for (size_t k = a; k < b; ++k) s += (size_t)req->len[k];
        return s;
    }
};

struct Slot {
    int panel = -1;
    esch_batch *batch = nullptr;
    size_t footprint = 0;
    bool busy = false;
    std::vector<Unit> units;
    Clock::time_point last_used;
};

struct PanelEntry {
    esch_scorer *scorer = nullptr;
    size_t n_refs = 0;
    size_t slot_footprint = 0;
};

}  // namespace

struct esch_server {
    esch_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<PanelEntry> panels;
    std::deque<Unit> queue;
    std::map<uint64_t, std::shared_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    uint64_t next_id = 1;
    size_t hbm_in_use = 0, hbm_peak = 0;
    uint64_t completed = 0, batches = 0;
    bool stop = false;
    std::thread th;

    ~esch_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (th.joinable()) th.join();
        for (auto &s : slots) esch_batch_free(s->batch);
        for (auto &p : panels) esch_scorer_free(p.scorer);
    }

    void finish_unit(const Unit &u, esch_status st, const char *err) {
        auto &r = *u.req;
        if (st && !r.status) {
            r.status = st;
            r.error = err ? err : "";
        }
        if (--r.parts_left == 0) {
            r.done = true;
            r.t_done = Clock::now();
            ++completed;
        }
    }

    // Collect finished slots (called with mu held).
    bool reap() {
        bool any = false;
        for (auto &sp : slots) {
            Slot &s = *sp;
            if (!s.busy || !esch_batch_ready(s.batch)) continue;
            size_t n = 0;
            for (auto &u : s.units) n += u.b - u.a;
            const size_t cap = s.units.empty() ? 1 : s.units[0].req->tie_cap;
            std::vector<esch_read_select> recs(std::max<size_t>(n, 1));
            std::vector<int32_t> ties(std::max<size_t>(n, 1) * cap);
            esch_status st = esch_batch_read_select(s.batch, recs.data(), ties.data());
            const char *err = st ? esch_last_error() : nullptr;
            size_t at = 0;
            for (auto &u : s.units) {
                auto &r = *u.req;
                for (size_t k = u.a; k < u.b; ++k, ++at) {
                    if (!st) {
                        r.recs[k] = recs[at];
                        memcpy(&r.ties[k * cap], &ties[at * cap], cap * sizeof(int32_t));
                    }
                }
                finish_unit(u, st, err);
            }
            s.units.clear();
            s.busy = false;
            s.last_used = Clock::now();
            any = true;
        }
        return any;
    }

    // A free slot for `panel`, creating one within the budget if needed
    // (mu held). nullptr when the budget or max_slots forbids it for now.
    Slot *free_slot(int panel) {
        for (auto &sp : slots)
            if (!sp->busy && sp->panel == panel) return sp.get();
        const size_t need = panels[panel].slot_footprint;
        auto busy_count = [&] {
            return (int)std::count_if(slots.begin(), slots.end(), [](auto &s) { return s->busy; });
        };
        if (busy_count() >= cfg.max_slots) return nullptr;
        // release idle slots of other panels, least recently used first
        while (hbm_in_use + need > cfg.hbm_budget || (int)slots.size() >= cfg.max_slots) {
            auto victim = slots.end();
            for (auto it = slots.begin(); it != slots.end(); ++it)
                if (!(*it)->busy && (victim == slots.end() || (*it)->last_used < (*victim)->last_used))
                    victim = it;
            if (victim == slots.end()) return nullptr;
            hbm_in_use -= (*victim)->footprint;
            esch_batch_free((*victim)->batch);
            slots.erase(victim);
        }
        auto s = std::make_unique<Slot>();
        s->panel = panel;
        s->footprint = need;
        if (esch_batch_create(panels[panel].scorer, cfg.max_batch_reads, cfg.max_batch_bases, &s->batch))
            return nullptr;
        hbm_in_use += need;
        hbm_peak = std::max(hbm_peak, hbm_in_use);
        slots.push_back(std::move(s));
        return slots.back().get();
    }

    static bool compatible(const Request &x, const Request &y) {
        return x.panel == y.panel && x.both == y.both && x.min_score == y.min_score &&
               x.max_ties == y.max_ties && x.tie_cap == y.tie_cap;
    }

    // Start as many batches as slots allow (mu held).
    bool admit() {
        bool any = false;
        while (!queue.empty()) {
            const Unit head = queue.front();
            Slot *s = free_slot(head.req->panel);
            if (!s) break;
            // pack compatible queued units, in arrival order, up to capacity
            std::vector<Unit> take;
            size_t reads = 0, bases = 0;
            for (auto it = queue.begin(); it != queue.end();) {
                const size_t r = it->b - it->a, bs = it->bases();
                if (compatible(*it->req, *head.req) && reads + r <= cfg.max_batch_reads &&
                    bases + bs <= cfg.max_batch_bases) {
                    take.push_back(*it);
                    reads += r;
                    bases += bs;
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
            std::vector<const uint8_t *> ptrs;
            std::vector<size_t> lens;
            ptrs.reserve(reads);
            lens.reserve(reads);
            for (auto &u : take)
                for (size_t k = u.a; k < u.b; ++k) {
                    ptrs.push_back(u.req->codes.data() + u.req->off[k]);
                    lens.push_back((size_t)u.req->len[k]);
                }
            const Request &q = *head.req;
            esch_status st = esch_batch_load(s->batch, ptrs.data(), lens.data(), reads, q.both, 0);
            if (!st) st = esch_batch_run(s->batch, ESCH_RUN_SCORES32 | ESCH_RUN_SELECT, q.min_score, q.max_ties, q.tie_cap);
            if (st) {
                const char *err = esch_last_error();
                for (auto &u : take) finish_unit(u, st, err);
                cv_done.notify_all();
                continue;
            }
            s->units = std::move(take);
            s->busy = true;
            ++batches;
            any = true;
        }
        return any;
    }

    void loop() {
        (void)hipSetDevice(cfg.ordinal);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            const bool reaped = reap();
            if (reaped) cv_done.notify_all();
            admit();
            const bool busy = std::any_of(slots.begin(), slots.end(), [](auto &s) { return s->busy; });
            if (busy) {
                // poll the device; new requests also wake the loop
                cv_work.wait_for(lk, std::chrono::microseconds(100));
            } else if (queue.empty()) {
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            } else {
                cv_work.wait_for(lk, std::chrono::microseconds(200));
            }
        }
    }
};

extern "C" {

esch_status esch_server_new(const esch_server_config *cfg, esch_server **out) {
    if (!cfg || !out || cfg->max_slots < 1 || cfg->max_batch_reads == 0 || cfg->max_batch_bases == 0) {
        return ESCH_INVALID_ARG;
    }
    *out = nullptr;
    if (!esch_device_visible()) return ESCH_UNAVAILABLE;
    auto s = std::make_unique<esch_server>();
    s->cfg = *cfg;
    esch_server *raw = s.get();
    s->th = std::thread([raw] { raw->loop(); });
    *out = s.release();
    return ESCH_OK;
}

void esch_server_free(esch_server *s) { delete s; }

esch_status esch_server_add_panel(esch_server *s, const uint8_t *const *refs, const size_t *lens, size_t n_refs,
                                  int32_t ma, int32_t mi, int32_t o, int32_t e, esch_mode mode, int *panel_id) {
    if (!s || !panel_id) return ESCH_INVALID_ARG;
    esch_scorer *sc = nullptr;
    if (esch_status st = esch_scorer_new(s->cfg.ordinal, refs, lens, n_refs, ma, mi, o, e, mode, &sc)) return st;
    PanelEntry p;
    p.scorer = sc;
    p.n_refs = n_refs;
    p.slot_footprint = esch_batch_footprint(sc, s->cfg.max_batch_reads, s->cfg.max_batch_bases);
    std::lock_guard<std::mutex> g(s->mu);
    if (p.slot_footprint > s->cfg.hbm_budget) {
        esch_scorer_free(sc);
        return ESCH_CAPACITY;
    }
    s->panels.push_back(p);
    *panel_id = (int)s->panels.size() - 1;
    return ESCH_OK;
}

esch_scorer *esch_server_scorer(esch_server *s, int panel_id) {
    if (!s) return nullptr;
    std::lock_guard<std::mutex> g(s->mu);
    if (panel_id < 0 || panel_id >= (int)s->panels.size()) return nullptr;
    return s->panels[panel_id].scorer;
}

esch_status esch_server_submit(esch_server *s, int panel_id, const uint8_t *const *reads, const size_t *lens,
                               size_t n, int both_strands, int32_t min_score, int64_t max_ties, size_t tie_cap,
                               uint64_t *request_id) {
    if (!s || !request_id || (n && (!reads || !lens))) return ESCH_INVALID_ARG;
    auto r = std::make_shared<Request>();
    r->t_submit = Clock::now();
    r->both = both_strands ? 1 : 0;
    r->min_score = min_score;
    r->max_ties = max_ties;
    r->tie_cap = std::max<size_t>(tie_cap, 1);
    r->n = n;
    r->off.resize(n);
    r->len.resize(n);
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) {
        if (lens[k] > s->cfg.max_batch_bases) return ESCH_CAPACITY;
        r->off[k] = total;
        r->len[k] = (int32_t)lens[k];
        total += lens[k];
    }
    r->codes.resize(std::max<size_t>(total, 1));
    for (size_t k = 0; k < n; ++k)
        if (lens[k]) memcpy(r->codes.data() + r->off[k], reads[k], lens[k]);
    r->recs.resize(n);
    r->ties.resize(n * r->tie_cap);
    std::lock_guard<std::mutex> g(s->mu);
    if (panel_id < 0 || panel_id >= (int)s->panels.size()) return ESCH_INVALID_ARG;
    r->panel = panel_id;
    r->id = s->next_id++;
    // cut into units of at most one batch
    std::vector<Unit> units;
    size_t a = 0;
    while (a < n) {
        size_t b = a, bases = 0;
        while (b < n && b - a < s->cfg.max_batch_reads && bases + (size_t)r->len[b] <= s->cfg.max_batch_bases)
            bases += (size_t)r->len[b++];
        units.push_back(Unit{r, a, b});
        a = b;
    }
    r->parts_left = units.size();
    s->reqs[r->id] = r;
    if (units.empty()) {
        r->done = true;
        r->t_done = Clock::now();
        ++s->completed;
    }
    for (auto &u : units) s->queue.push_back(u);
    *request_id = r->id;
    s->cv_work.notify_all();
    return ESCH_OK;
}

esch_status esch_server_wait(esch_server *s, uint64_t id, int64_t timeout_ms) {
    if (!s) return ESCH_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return ESCH_INVALID_ARG;
    auto r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return ESCH_TIMEOUT;
    return r->status;
}

esch_status esch_server_result(esch_server *s, uint64_t id, esch_read_select *out, int32_t *ties,
                               double *latency_s) {
    if (!s) return ESCH_INVALID_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return ESCH_INVALID_ARG;
    auto r = it->second;
    if (!r->done) return ESCH_TIMEOUT;
    if (out && r->n) memcpy(out, r->recs.data(), r->n * sizeof(esch_read_select));
    if (ties && r->n) memcpy(ties, r->ties.data(), r->ties.size() * sizeof(int32_t));
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    const esch_status st = r->status;
    s->reqs.erase(it);
    return st;
}

esch_status esch_server_stats(esch_server *s, esch_server_stats_t *out) {
    if (!s || !out) return ESCH_INVALID_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    out->completed = s->completed;
    out->batches = s->batches;
    out->queued_units = s->queue.size();
    out->hbm_in_use = s->hbm_in_use;
    out->hbm_peak = s->hbm_peak;
    out->slots = s->slots.size();
    out->busy_slots = (size_t)std::count_if(s->slots.begin(), s->slots.end(), [](auto &x) { return x->busy; });
    return ESCH_OK;
}

}  // extern "C"
