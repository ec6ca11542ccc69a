// serve_host.hip - the serving loop over batched films (cz_server_* of
// include/canvaz_film.h).
//
// A request asks for `steps` frames of a seeded canvas.  A request is
// admitted when its arrival time has passed, a batch slot is free and the
// canvas fits the device-memory budget, first come first served or, with
// max_wait_s > 0, least work first with the aged first (an aged canvas that
// does not fit evicts long-running ones to host memory); an admitted canvas
// joins the running batch between two frames, takes one frame per batched
// frame (with its own substeps, exactly as it would alone) and leaves the
// batch once its steps are done.  Films are recycled: a finished canvas's
// film goes to a pool, and a later canvas that fits its capacity reuses it
// (re-dimensioned and reseeded on the device) instead of allocating;
// pooled films count against the budget and are released when a new
// allocation needs their room.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "film_internal.h"

using namespace czi;

namespace {

using Clock = std::chrono::steady_clock;

struct Slot {
    cz_film* film;
    uint32_t req;        // index into the request array
    uint32_t left;       // frames still to run
    double since;        // when it was admitted or resumed
};

// A canvas evicted to host memory, waiting to resume.
struct Parked {
    uint32_t req;
    uint32_t left;
    double since;        // when it was evicted
    czi::FilmSnapshot snap;
};

uint64_t film_bytes_for(uint32_t w, uint32_t h) {
    // the allocation pattern of film_create: per cell, per 16-tile, per
    // 64-tile and per pool cell, plus the fixed deposit tables
    const uint64_t n = (uint64_t)w * h;
    const uint64_t nt16 = (uint64_t)((w + 15u) / 16u) * ((h + 15u) / 16u);
    const uint64_t nt64 = (uint64_t)((w + 63u) / 64u) * ((h + 63u) / 64u);
    const uint64_t np = (uint64_t)((w + 3u) / 4u) * ((h + 3u) / 4u);
    const uint64_t per_cell = 4 + 4 + 8 + 16 + 128 + 16 + 32 + 4 + 8 + 16 + 16 + 4 + 4 + 4 + 4 + 4 + 64 + 16 + 16 +
                              128 + 16 + 4 + 8 + 16 + 7 * 4 + 4;
    return n * per_cell + nt16 * (8 + 8 + 4 + 4 + 4 + 4 + 1 + 1) + nt64 * (8 + 4 + 4 + 4) + np * 8 +
           (uint64_t)CZ_MAX_DEPOSITS * (sizeof(cz_deposit) + 16) + (uint64_t)CZ_MAX_TILE_ITEMS * 4 + 64 * 1024;
}

bool fits(const cz_film* f, uint32_t w, uint32_t h) {
    const uint64_t n = (uint64_t)w * h;
    const uint64_t nt16 = (uint64_t)((w + 15u) / 16u) * ((h + 15u) / 16u);
    const uint64_t nt64 = (uint64_t)((w + 63u) / 64u) * ((h + 63u) / 64u);
    const uint64_t np = (uint64_t)((w + 3u) / 4u) * ((h + 3u) / 4u);
    return n <= f->cap_n && nt16 <= f->cap_nt16 && nt64 <= f->cap_nt64 && np <= f->cap_np;
}

}  // namespace

struct cz_server {
    cz_server_config cfg;
    cz_batch* batch = nullptr;
    std::vector<cz_film*> pool;      // idle films
    std::vector<std::pair<uint64_t, cz_film*>> kept;  // finished canvases kept for reading
    uint64_t held = 0;               // device bytes of the active and pooled films
};

namespace {

void release(cz_server* s, cz_film* f) {
    s->held -= f->bytes;
    cz_film_destroy(f);
}

// A film for a w x h canvas: the smallest pooled film that fits, else a new
// one if the budget allows (releasing pooled films to make room).
int acquire(cz_server* s, const cz_request& r, cz_film** out, cz_server_stats& st) {
    *out = nullptr;
    cz_film* best = nullptr;
    size_t bi = 0;
    for (size_t k = 0; k < s->pool.size(); k++) {
        cz_film* f = s->pool[k];
        if (fits(f, r.w, r.h) && (!best || f->cap_n < best->cap_n)) {
            best = f;
            bi = k;
        }
    }
    if (best) {
        s->pool.erase(s->pool.begin() + (long)bi);
        int rc = film_reshape(best, r.w, r.h, &r.settings);
        if (rc) {
            release(s, best);
            return rc;
        }
        *out = best;
        st.films_reused++;
        return CZ_OK;
    }
    const uint64_t need = film_bytes_for(r.w, r.h);
    if (need > s->cfg.hbm_budget) return set_err(CZ_ENOMEM, "request %llu needs %llu bytes, over the budget",
                                                 (unsigned long long)r.id, (unsigned long long)need);
    // release pooled films, largest first, until the new one fits
    std::sort(s->pool.begin(), s->pool.end(), [](cz_film* a, cz_film* b) { return a->bytes < b->bytes; });
    while (s->held + need > s->cfg.hbm_budget && !s->pool.empty()) {
        release(s, s->pool.back());
        s->pool.pop_back();
    }
    if (s->held + need > s->cfg.hbm_budget) return CZ_OK;  // wait for an active canvas to finish
    cz_film* f = nullptr;
    int rc = film_create(r.w, r.h, &r.settings, nullptr, nullptr, false, &f);
    if (rc) return rc;
    s->held += f->bytes;
    *out = f;
    st.films_created++;
    return CZ_OK;
}

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

}  // namespace

extern "C" {

void cz_server_config_default(cz_server_config* c) {
    memset(c, 0, sizeof(*c));
    c->hbm_budget = 200ull << 30;
    c->max_members = 256;
    c->frame_dt = 1.0f / 60.0f;
    c->keep_films = 0;
    c->max_wait_s = 0.0f;
    c->host_budget = 64ull << 30;
}

int cz_server_create(const cz_server_config* cfg, cz_server** out) {
    if (!cfg || !out || cfg->max_members == 0 || !(cfg->frame_dt > 0.0f)) return set_err(CZ_EINVAL, "cz_server_create: bad config");
    *out = nullptr;
    cz_server* s = new cz_server();
    s->cfg = *cfg;
    int rc = cz_batch_create(cfg->max_members, &s->batch);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return CZ_OK;
}

void cz_server_destroy(cz_server* s) {
    if (!s) return;
    if (s->batch) cz_batch_destroy(s->batch);
    for (cz_film* f : s->pool) cz_film_destroy(f);
    for (auto& k : s->kept) cz_film_destroy(k.second);
    delete s;
}

int cz_server_run(cz_server* s, const cz_request* reqs, uint32_t n, cz_result* res, cz_server_stats* st) {
    if (!s || (n > 0 && (!reqs || !res))) return set_err(CZ_EINVAL, "cz_server_run: bad arguments");
    for (uint32_t k = 1; k < n; k++)
        if (reqs[k].arrival_s < reqs[k - 1].arrival_s) return set_err(CZ_EINVAL, "cz_server_run: requests out of order");
    cz_server_stats stats;
    memset(&stats, 0, sizeof(stats));
    for (uint32_t k = 0; k < n; k++) {
        memset(&res[k], 0, sizeof(res[k]));
        res[k].id = reqs[k].id;
        res[k].arrival_s = reqs[k].arrival_s;
        res[k].status = 1;  // pending
    }
    const double max_wait = s->cfg.max_wait_s > 0.0f ? (double)s->cfg.max_wait_s : 0.0;
    std::vector<uint32_t> waiting;  // arrived, not yet admitted (arrival order)
    std::vector<Parked> parked;     // evicted, waiting to resume
    std::vector<Slot> active;
    uint64_t host_bytes = 0;
    uint32_t next = 0, done = 0;
    const Clock::time_point t0 = Clock::now();
    // on an error: every running canvas leaves the batch and its film is
    // released, so that neither the batch nor the server keeps it
    auto fail = [&](int rc) -> int {
        const std::string msg = cz_last_error();
        for (Slot& a : active) {
            cz_batch_remove(s->batch, a.film);
            release(s, a.film);
        }
        active.clear();
        parked.clear();
        set_err(rc, "%s", msg.c_str());
        return rc;
    };
    // admit (or resume) a canvas onto film f
    auto start = [&](cz_film* f, uint32_t k, uint32_t left, const FilmSnapshot* snap, double now) -> int {
        int rc = snap ? film_restore(f, *snap) : cz_film_seed_blobs(f, &reqs[k].scene);
        if (!rc) rc = cz_batch_add(s->batch, f);
        if (rc) {
            release(s, f);
            return rc;
        }
        active.push_back({f, k, left, now});
        return CZ_OK;
    };
    // take the running canvas active[q] out to host memory
    auto evict = [&](size_t q, double now) -> int {
        Slot a = active[q];
        int rc = cz_batch_remove(s->batch, a.film);
        if (rc) return rc;
        active[q] = active.back();
        active.pop_back();
        Parked p;
        p.req = a.req;
        p.left = a.left;
        p.since = now;
        rc = film_snapshot(a.film, p.snap);
        if (rc) {
            release(s, a.film);
            return rc;
        }
        s->pool.push_back(a.film);
        host_bytes += p.snap.data.size();
        stats.peak_host_bytes = std::max<uint64_t>(stats.peak_host_bytes, host_bytes);
        parked.push_back(std::move(p));
        stats.evictions++;
        return CZ_OK;
    };
    int rc = CZ_OK;
    while (done < n) {
        double now = since(t0);
        while (next < n && reqs[next].arrival_s <= now) waiting.push_back(next++);
        if (max_wait <= 0.0) {
            // first come, first served: the head waits until it fits
            size_t w = 0;
            while (w < waiting.size() && active.size() < s->cfg.max_members) {
                const uint32_t k = waiting[w];
                const cz_request& r = reqs[k];
                if (r.steps == 0) {
                    res[k].admit_s = res[k].done_s = since(t0);
                    res[k].status = 0;
                    w++;
                    done++;
                    continue;
                }
                cz_film* f = nullptr;
                rc = acquire(s, r, &f, stats);
                if (rc) {
                    res[k].status = rc;
                    res[k].admit_s = res[k].done_s = since(t0);
                    w++;
                    done++;
                    rc = CZ_OK;
                    continue;
                }
                if (!f) break;  // not enough memory until a canvas finishes
                rc = start(f, k, r.steps, nullptr, now);
                if (rc) return fail(rc);
                res[k].admit_s = since(t0);
                stats.admitted++;
                w++;
            }
            waiting.erase(waiting.begin(), waiting.begin() + (long)w);
        } else {
            // least work first; the aged (waiting longer than max_wait) before
            // all others, oldest first, and allowed to evict
            struct Cand {
                bool parked, aged;
                double key;
                uint64_t work;
                uint32_t idx;
            };
            std::vector<Cand> cand;
            for (size_t w = 0; w < waiting.size(); w++) {
                const cz_request& r = reqs[waiting[w]];
                cand.push_back({false, now - r.arrival_s >= max_wait, r.arrival_s,
                                (uint64_t)r.w * r.h * r.steps, (uint32_t)w});
            }
            for (size_t q = 0; q < parked.size(); q++) {
                const cz_request& r = reqs[parked[q].req];
                cand.push_back({true, now - parked[q].since >= max_wait, parked[q].since,
                                (uint64_t)r.w * r.h * parked[q].left, (uint32_t)q});
            }
            std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
                if (a.aged != b.aged) return a.aged;
                if (a.aged) return a.key < b.key;
                if (a.work != b.work) return a.work < b.work;
                return a.key < b.key;
            });
            std::vector<char> took_w(waiting.size(), 0), took_p(parked.size(), 0);
            for (const Cand& c : cand) {
                if (active.size() >= s->cfg.max_members) break;
                const uint32_t k = c.parked ? parked[c.idx].req : waiting[c.idx];
                const cz_request& r = reqs[k];
                if (!c.parked && r.steps == 0) {
                    res[k].admit_s = res[k].done_s = since(t0);
                    res[k].status = 0;
                    took_w[c.idx] = 1;
                    done++;
                    continue;
                }
                cz_film* f = nullptr;
                rc = acquire(s, r, &f, stats);
                if (!rc && !f && c.aged) {
                    // plan the evictions: running canvases that have run for
                    // max_wait since they started and have more work left
                    // than the candidate (so two canvases never evict each
                    // other in turn), most work left first, each within the
                    // host budget, until the candidate's film fits
                    std::vector<std::pair<uint64_t, cz_film*>> order;
                    for (const Slot& a : active) {
                        const uint64_t work = (uint64_t)a.film->W * a.film->H * a.left;
                        if (now - a.since >= max_wait && work > c.work) order.push_back({work, a.film});
                    }
                    std::sort(order.begin(), order.end(),
                              [](const std::pair<uint64_t, cz_film*>& x, const std::pair<uint64_t, cz_film*>& y) {
                                  return x.first > y.first;
                              });
                    const uint64_t need = film_bytes_for(r.w, r.h);
                    uint64_t freeable = 0, to_host = host_bytes;
                    for (cz_film* pf : s->pool) freeable += pf->bytes;
                    std::vector<cz_film*> plan;
                    for (const auto& o : order) {
                        if (s->held - freeable + need <= s->cfg.hbm_budget) break;
                        const uint64_t sb = film_snapshot_bytes(o.second);
                        if (to_host + sb > s->cfg.host_budget) continue;
                        plan.push_back(o.second);
                        to_host += sb;
                        freeable += o.second->bytes;
                    }
                    if (s->held - freeable + need > s->cfg.hbm_budget) break;  // nothing passes it
                    for (cz_film* victim : plan) {
                        size_t q = 0;
                        while (active[q].film != victim) q++;
                        rc = evict(q, now);
                        if (rc) return fail(rc);
                        took_p.push_back(0);
                    }
                    rc = acquire(s, r, &f, stats);
                }
                if (rc) {
                    if (c.parked) return fail(rc);  // it fitted before: a device failure
                    res[k].status = rc;
                    res[k].admit_s = res[k].done_s = since(t0);
                    took_w[c.idx] = 1;
                    done++;
                    rc = CZ_OK;
                    continue;
                }
                if (!f) {
                    if (c.aged) break;  // nothing passes an aged canvas that does not fit
                    continue;           // a smaller one may fit
                }
                if (c.parked) {
                    Parked& p = parked[c.idx];
                    rc = start(f, k, p.left, &p.snap, now);
                    if (rc) return fail(rc);
                    host_bytes -= p.snap.data.size();
                    std::vector<uint8_t>().swap(p.snap.data);
                    took_p[c.idx] = 1;
                    stats.resumes++;
                } else {
                    rc = start(f, k, r.steps, nullptr, now);
                    if (rc) return fail(rc);
                    res[k].admit_s = since(t0);
                    took_w[c.idx] = 1;
                    stats.admitted++;
                }
            }
            size_t o = 0;
            for (size_t w = 0; w < waiting.size(); w++)
                if (!took_w[w]) waiting[o++] = waiting[w];
            waiting.resize(o);
            std::vector<Parked> keep;
            for (size_t q = 0; q < parked.size(); q++)
                if (q >= took_p.size() || !took_p[q]) keep.push_back(std::move(parked[q]));
            parked.swap(keep);
        }
        stats.peak_members = std::max<uint64_t>(stats.peak_members, active.size());
        stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, s->held);
        if (active.empty()) {
            if (next >= n && waiting.empty() && parked.empty()) break;
            if (!waiting.empty() || !parked.empty()) {
                const cz_request& r = reqs[!waiting.empty() ? waiting[0] : parked[0].req];
                return fail(set_err(CZ_ENOMEM,
                                    "cz_server_run: request %llu (%ux%u, %llu bytes) does not fit an empty server "
                                    "(%zu waiting, %zu evicted, %llu bytes held in %zu pooled films, budget %llu)",
                                    (unsigned long long)r.id, r.w, r.h, (unsigned long long)film_bytes_for(r.w, r.h),
                                    waiting.size(), parked.size(), (unsigned long long)s->held, s->pool.size(),
                                    (unsigned long long)s->cfg.hbm_budget));
            }
            // idle until the next arrival
            double wait = reqs[next].arrival_s - since(t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
            continue;
        }
        rc = cz_batch_step(s->batch, s->cfg.frame_dt);
        if (rc) return fail(rc);
        stats.frames++;
        bool finishing = false;
        for (Slot& a : active) {
            a.left--;
            stats.cell_steps += (uint64_t)a.film->W * a.film->H;
            if (a.left == 0) finishing = true;
        }
        if (!finishing) continue;
        rc = cz_batch_sync(s->batch);
        if (rc) return fail(rc);
        const double t_done = since(t0);
        for (size_t q = 0; q < active.size();) {
            Slot& a = active[q];
            if (a.left > 0) {
                q++;
                continue;
            }
            rc = cz_batch_remove(s->batch, a.film);
            if (rc) return fail(rc);
            res[a.req].done_s = t_done;
            res[a.req].status = 0;
            res[a.req].cell_steps = (uint64_t)a.film->W * a.film->H * reqs[a.req].steps;
            if (s->cfg.keep_films) {
                // kept for reading, outside the budget (which covers the
                // films the loop steps and pools)
                s->held -= a.film->bytes;
                s->kept.push_back({reqs[a.req].id, a.film});
            } else {
                s->pool.push_back(a.film);
            }
            active[q] = active.back();
            active.pop_back();
            done++;
        }
    }
    rc = cz_batch_sync(s->batch);
    stats.wall_s = since(t0);
    if (st) *st = stats;
    return rc;
}

cz_film* cz_server_film(cz_server* s, uint64_t id) {
    if (!s) return nullptr;
    for (auto& k : s->kept)
        if (k.first == id) return k.second;
    return nullptr;
}

}  // extern "C"
