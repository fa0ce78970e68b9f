// Serving loop: resident panels, continuous batching, admission by HBM
// footprint (no counterpart in the reference, whose CLI scores one batch at
// a time on one GPU thread, crates/eschalign-cli/src/align.rs:2801-2917).
//
// Requests (reads against one resident panel) are cut into units of at most
// one batch each, and at most max_batch_cells DP cells each, so that a large
// request becomes many short device batches; units are cut from the
// request's reads in length order, so each holds reads of near-equal length. A scheduler thread packs queued
// units that share a panel and options into a device batch whenever fewer
// than max_slots batches are in flight, taking first the units of the
// request with the least queued work (shortest remaining work first; ties by
// arrival): a small request waits for at most the batches already in flight,
// not for every unit of a large request queued before it. Each slot (a
// device batch of one panel, with its own stream, scratch and captured
// graphs) is kept for reuse while idle and released, least recently used
// first, only when a new slot would not fit the HBM budget.
//
// Only the scheduler thread makes HIP calls, and it makes them without the
// mutex held: slot creation, staging, launches, result copies and frees
// (hipFree waits for the device) never block submit, wait or result.
// Scores per read do not depend on what shares its batch, so every result
// equals the result of the request run alone.
#include "../../include/eschalign_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <utility>
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
    std::vector<uint32_t> perm;  // read indices, longest first: units are ranges of it
    std::vector<esch_read_select> recs;
    std::vector<int32_t> ties;
    size_t parts_left = 0;
    uint64_t cells_queued = 0;  // DP cells of this request's units still queued
    bool done = false;
    esch_status status = ESCH_OK;
    std::string error;
    Clock::time_point t_submit, t_done;
};

struct Unit {  // reads perm[a .. b) of one request
    std::shared_ptr<Request> req;
    size_t a = 0, b = 0;
    size_t bases = 0;
    uint64_t cells = 0;
};

enum class SlotState {
    Idle,  // no batch in flight; reusable by its panel
    Held,  // the scheduler works on it with the mutex released
    Busy,  // a batch in flight on its stream
};

struct Slot {
    int panel = -1;
    esch_batch *batch = nullptr;
    size_t footprint = 0;
    SlotState state = SlotState::Idle;
    std::vector<Unit> units;
    std::vector<esch_read_select> recs;  // result staging, reused
    std::vector<int32_t> ties;
    Clock::time_point last_used;
};

struct PanelEntry {
    esch_scorer *scorer = nullptr;
    size_t n_refs = 0;
    uint64_t ref_bases = 0;
    size_t slot_footprint = 0;
};

}  // namespace

struct esch_server {
    esch_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<PanelEntry> panels;
    std::vector<Unit> queue;
    std::map<uint64_t, std::shared_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    uint64_t next_id = 1;
    size_t hbm_in_use = 0, hbm_peak = 0;  // slot footprints, counted from reservation to release
    uint64_t completed = 0, batches = 0, swapped_batches = 0;
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

    // (mu held)
    void finish_unit(const Unit &u, esch_status st, const std::string &err) {
        auto &r = *u.req;
        if (st && !r.status) {
            r.status = st;
            r.error = err;
        }
        if (--r.parts_left == 0) {
            r.done = true;
            r.t_done = Clock::now();
            ++completed;
        }
    }

    int in_flight() const {
        return (int)std::count_if(slots.begin(), slots.end(),
                                  [](const auto &s) { return s->state != SlotState::Idle; });
    }

    static bool compatible(const Request &x, const Request &y) {
        return x.panel == y.panel && x.both == y.both && x.min_score == y.min_score &&
               x.max_ties == y.max_ties && x.tie_cap == y.tie_cap;
    }

    // Copy out the results of every finished batch and complete its units.
    // Called and returns with the lock held; releases it around the copies.
    bool reap(std::unique_lock<std::mutex> &lk) {
        std::vector<Slot *> ready;
        for (auto &sp : slots)
            if (sp->state == SlotState::Busy && esch_batch_ready(sp->batch)) {
                sp->state = SlotState::Held;
                ready.push_back(sp.get());
            }
        if (ready.empty()) return false;
        lk.unlock();
        std::vector<std::pair<esch_status, std::string>> res;
        for (Slot *s : ready) {
            size_t n = 0;
            for (auto &u : s->units) n += u.b - u.a;
            const size_t cap = s->units.empty() ? 1 : s->units[0].req->tie_cap;
            s->recs.resize(std::max<size_t>(n, 1));
            s->ties.resize(std::max<size_t>(n, 1) * cap);
            const esch_status st = esch_batch_read_select(s->batch, s->recs.data(), s->ties.data());
            res.emplace_back(st, st ? esch_last_error() : "");
        }
        lk.lock();
        for (size_t i = 0; i < ready.size(); ++i) {
            Slot &s = *ready[i];
            const esch_status st = res[i].first;
            const size_t cap = s.units.empty() ? 1 : s.units[0].req->tie_cap;
            size_t at = 0;
            for (auto &u : s.units) {
                auto &r = *u.req;
                for (size_t k = u.a; k < u.b && !st; ++k) {
                    const size_t q = r.perm[k], j = at + (k - u.a);
                    r.recs[q] = s.recs[j];
                    memcpy(&r.ties[q * cap], &s.ties[j * cap], cap * sizeof(int32_t));
                }
                at += u.b - u.a;
                finish_unit(u, st, res[i].second);
            }
            s.units.clear();
            s.state = SlotState::Idle;
            s.last_used = Clock::now();
        }
        return true;
    }

    // Pick the next batch: the queued unit of the request with the least
    // queued work whose panel can get a slot, and the compatible units that
    // fit beside it in the same order (mu held). Reserves the slot (Held),
    // marks idle slots to release in `evict`, and removes the taken units
    // from the queue. nullptr when nothing can start now.
    Slot *pick(std::vector<Unit> &take, std::vector<std::unique_ptr<Slot>> &evict, bool &fresh) {
        fresh = false;
        if (queue.empty() || in_flight() >= cfg.max_slots) return nullptr;
        std::vector<size_t> order(queue.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        auto key = [&](size_t i) {
            const Unit &u = queue[i];
            return std::make_tuple(u.req->cells_queued, u.req->id, u.a);
        };
        std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return key(x) < key(y); });
        std::set<int> tried;
        for (size_t oi = 0; oi < order.size(); ++oi) {
            const Unit &head = queue[order[oi]];
            const int panel = head.req->panel;
            if (!tried.insert(panel).second) continue;
            Slot *s = nullptr;
            for (auto &sp : slots)
                if (sp->state == SlotState::Idle && sp->panel == panel) {
                    s = sp.get();
                    break;
                }
            if (!s) {
                // a new slot; release idle slots of other panels, least
                // recently used first, while it would not fit the budget
                const size_t need = panels[panel].slot_footprint;
                size_t freed = 0;
                std::vector<Slot *> victims;
                std::vector<Slot *> idle;
                for (auto &sp : slots)
                    if (sp->state == SlotState::Idle) idle.push_back(sp.get());
                std::sort(idle.begin(), idle.end(),
                          [](Slot *x, Slot *y) { return x->last_used < y->last_used; });
                for (Slot *v : idle) {
                    if (hbm_in_use - freed + need <= cfg.hbm_budget) break;
                    victims.push_back(v);
                    freed += v->footprint;
                }
                if (hbm_in_use - freed + need > cfg.hbm_budget) continue;  // wait for a slot to finish
                // the victims are freed before the new slot is created, on
                // this thread, so their memory counts as released now
                for (Slot *v : victims) {
                    auto it = std::find_if(slots.begin(), slots.end(), [&](auto &p) { return p.get() == v; });
                    evict.push_back(std::move(*it));
                    slots.erase(it);
                }
                hbm_in_use -= freed;
                auto ns = std::make_unique<Slot>();
                ns->panel = panel;
                ns->footprint = need;
                hbm_in_use += need;
                hbm_peak = std::max(hbm_peak, hbm_in_use);
                slots.push_back(std::move(ns));
                s = slots.back().get();
                fresh = true;
            }
            // pack compatible queued units, least queued work first
            size_t reads = 0, bases = 0;
            uint64_t cells = 0;
            std::vector<char> taken(queue.size(), 0);
            for (size_t oj = oi; oj < order.size(); ++oj) {
                const Unit &u = queue[order[oj]];
                const size_t r = u.b - u.a;
                if (!compatible(*u.req, *head.req)) continue;
                if (!take.empty() &&
                    (reads + r > cfg.max_batch_reads || bases + u.bases > cfg.max_batch_bases ||
                     (cfg.max_batch_cells && cells + u.cells > cfg.max_batch_cells)))
                    continue;
                take.push_back(u);
                taken[order[oj]] = 1;
                reads += r;
                bases += u.bases;
                cells += u.cells;
            }
            std::vector<Unit> rest;
            rest.reserve(queue.size() - take.size());
            for (size_t i = 0; i < queue.size(); ++i)
                if (!taken[i]) rest.push_back(std::move(queue[i]));
            queue.swap(rest);
            for (auto &u : take) u.req->cells_queued -= u.cells;
            s->state = SlotState::Held;
            return s;
        }
        return nullptr;
    }

    // Start as many batches as the slots allow. Called and returns with the
    // lock held; releases it around the device work.
    bool admit(std::unique_lock<std::mutex> &lk) {
        bool any = false;
        for (;;) {
            std::vector<Unit> take;
            std::vector<std::unique_ptr<Slot>> evict;
            bool fresh = false;
            Slot *s = pick(take, evict, fresh);
            if (!s) return any;
            const PanelEntry pe = panels[s->panel];
            const size_t cap_reads = cfg.max_batch_reads, cap_bases = cfg.max_batch_bases;
            lk.unlock();
            for (auto &v : evict) esch_batch_free(v->batch);
            esch_status st = ESCH_OK;
            std::string err;
            if (fresh) {
                st = esch_batch_create(pe.scorer, cap_reads, cap_bases, &s->batch);
                if (st) err = esch_last_error();
            }
            if (!st) {
                std::vector<const uint8_t *> ptrs;
                std::vector<size_t> lens;
                for (auto &u : take)
                    for (size_t k = u.a; k < u.b; ++k) {
                        const size_t q = u.req->perm[k];
                        ptrs.push_back(u.req->codes.data() + u.req->off[q]);
                        lens.push_back((size_t)u.req->len[q]);
                    }
                const Request &q = *take[0].req;
                st = esch_batch_load(s->batch, ptrs.data(), lens.data(), ptrs.size(), q.both, 0);
                if (!st)
                    st = esch_batch_run(s->batch, ESCH_RUN_SCORES32 | ESCH_RUN_SELECT, q.min_score, q.max_ties,
                                        q.tie_cap);
                if (st) err = esch_last_error();
            }
            const bool swapped = !st && esch_batch_swapped(s->batch);
            lk.lock();
            if (st) {
                for (auto &u : take) finish_unit(u, st, err);
                if (fresh && !s->batch) {
                    hbm_in_use -= s->footprint;
                    slots.erase(std::find_if(slots.begin(), slots.end(), [&](auto &p) { return p.get() == s; }));
                } else {
                    s->state = SlotState::Idle;
                    s->last_used = Clock::now();
                }
                cv_done.notify_all();
                continue;
            }
            s->units = std::move(take);
            s->state = SlotState::Busy;
            ++batches;
            swapped_batches += swapped ? 1 : 0;
            any = true;
        }
    }

    void loop() {
        (void)hipSetDevice(cfg.ordinal);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            if (reap(lk)) cv_done.notify_all();
            admit(lk);
            if (stop) break;
            const bool busy = in_flight() > 0;
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
    for (size_t i = 0; i < n_refs; ++i) p.ref_bases += lens[i];
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
    // units of near-equal lengths: the device batches reads in chunks of
    // equal-length rows, so a unit cut from the length order carries no
    // padding and routes nothing to the long-read kernel
    r->perm.resize(n);
    for (size_t k = 0; k < n; ++k) r->perm[k] = (uint32_t)k;
    std::stable_sort(r->perm.begin(), r->perm.end(),
                     [&](uint32_t x, uint32_t y) { return r->len[x] > r->len[y]; });
    std::lock_guard<std::mutex> g(s->mu);
    if (panel_id < 0 || panel_id >= (int)s->panels.size()) return ESCH_INVALID_ARG;
    r->panel = panel_id;
    r->id = s->next_id++;
    // cut into units of at most one batch (reads, bases, DP cells) along the
    // length order; a unit cut short by a limit holds a multiple of 64 reads
    // when it can (128 rows with both strands: whole chunks), and a read
    // over the cell limit alone is a unit of its own
    const uint64_t per_base = s->panels[panel_id].ref_bases * (r->both ? 2 : 1);
    std::vector<Unit> units;
    size_t a = 0;
    while (a < n) {
        size_t b = a, bases = 0;
        uint64_t cells = 0;
        while (b < n && b - a < s->cfg.max_batch_reads) {
            const size_t len = (size_t)r->len[r->perm[b]];
            const uint64_t c = (uint64_t)len * per_base;
            if (bases + len > s->cfg.max_batch_bases) break;
            if (b > a && s->cfg.max_batch_cells && cells + c > s->cfg.max_batch_cells) break;
            bases += len;
            cells += c;
            ++b;
        }
        if (b < n && b - a > 64) b = a + (b - a) / 64 * 64;
        Unit u;
        u.req = r;
        u.a = a;
        u.b = b;
        for (size_t k = a; k < b; ++k) {
            u.bases += (size_t)r->len[r->perm[k]];
            u.cells += (uint64_t)r->len[r->perm[k]] * per_base;
        }
        r->cells_queued += u.cells;
        units.push_back(std::move(u));
        a = b;
    }
    r->parts_left = units.size();
    s->reqs[r->id] = r;
    if (units.empty()) {
        r->done = true;
        r->t_done = Clock::now();
        ++s->completed;
    }
    for (auto &u : units) s->queue.push_back(std::move(u));
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
    out->busy_slots = (size_t)s->in_flight();
    out->swapped_batches = s->swapped_batches;
    return ESCH_OK;
}

}  // extern "C"
