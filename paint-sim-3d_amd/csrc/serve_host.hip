// serve_host.hip - the serving loop over batched films (cz_server_* of
// include/canvaz_film.h).
//
// A request asks for `steps` frames of a seeded canvas.  Requests are
// admitted first come, first served, when their arrival time has passed,
// a batch slot is free and the canvas fits the device-memory budget; an
// admitted canvas joins the running batch between two frames, takes one
// frame per batched frame (with its own substeps, exactly as it would alone)
// and leaves the batch once its steps are done.  Canvas state stays resident
// from admission to completion.  Films are recycled: a finished canvas's
// film goes to a pool, and a later canvas that fits its capacity reuses it
// (re-dimensioned and reseeded on the device) instead of allocating;
// pooled films count against the budget and are released when a new
// allocation needs their room.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
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
    uint64_t held = 0;               // device bytes of every film the server owns
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
    std::deque<uint32_t> waiting;
    std::vector<Slot> active;
    uint32_t next = 0, done = 0;
    const Clock::time_point t0 = Clock::now();
    int rc = CZ_OK;
    while (done < n) {
        double now = since(t0);
        while (next < n && reqs[next].arrival_s <= now) waiting.push_back(next++);
        // admission, first come first served
        while (!waiting.empty() && active.size() < s->cfg.max_members) {
            const uint32_t k = waiting.front();
            const cz_request& r = reqs[k];
            if (r.steps == 0) {
                res[k].admit_s = res[k].done_s = since(t0);
                res[k].status = 0;
                waiting.pop_front();
                done++;
                continue;
            }
            cz_film* f = nullptr;
            rc = acquire(s, r, &f, stats);
            if (rc) {
                res[k].status = rc;
                res[k].admit_s = res[k].done_s = since(t0);
                waiting.pop_front();
                done++;
                rc = CZ_OK;
                continue;
            }
            if (!f) break;  // not enough memory until a canvas finishes
            rc = cz_film_seed_blobs(f, &r.scene);
            if (!rc) rc = cz_batch_add(s->batch, f);
            if (rc) {
                release(s, f);
                return rc;
            }
            res[k].admit_s = since(t0);
            active.push_back({f, k, r.steps});
            waiting.pop_front();
            stats.admitted++;
        }
        stats.peak_members = std::max<uint64_t>(stats.peak_members, active.size());
        stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, s->held);
        if (active.empty()) {
            if (next >= n && waiting.empty()) break;
            if (!waiting.empty()) return set_err(CZ_ENOMEM, "cz_server_run: request does not fit an empty server");
            // idle until the next arrival
            double wait = reqs[next].arrival_s - since(t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
            continue;
        }
        rc = cz_batch_step(s->batch, s->cfg.frame_dt);
        if (rc) return rc;
        stats.frames++;
        bool finishing = false;
        for (Slot& a : active) {
            a.left--;
            stats.cell_steps += (uint64_t)a.film->W * a.film->H;
            if (a.left == 0) finishing = true;
        }
        if (!finishing) continue;
        rc = cz_batch_sync(s->batch);
        if (rc) return rc;
        const double t_done = since(t0);
        for (size_t q = 0; q < active.size();) {
            Slot& a = active[q];
            if (a.left > 0) {
                q++;
                continue;
            }
            rc = cz_batch_remove(s->batch, a.film);
            if (rc) return rc;
            res[a.req].done_s = t_done;
            res[a.req].status = 0;
            res[a.req].cell_steps = (uint64_t)a.film->W * a.film->H * reqs[a.req].steps;
            if (s->cfg.keep_films) s->kept.push_back({reqs[a.req].id, a.film});
            else s->pool.push_back(a.film);
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
