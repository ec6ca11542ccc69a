// film_host.hip - host side of the film: the C ABI of include/canvaz_film.h
// over the gfx950 kernels.  Mirrors `Film` (src/film/mod.rs:445-1462):
// buffers, update_params / choose_substeps, upload_deposits binning,
// encode_frame order, stats readback, read/write_state, dry_now, clear.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "film_internal.h"
#include "host_ref.h"

using namespace cz;

static thread_local char g_err[512] = "";

int cz_set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

namespace {

template <typename T>
int dalloc(cz_film* f, T** p, size_t count) {
    void* v = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess) return set_err(CZ_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    f->allocs.push_back(v);
    f->bytes += bytes;
    *p = (T*)v;
    e = hipMemsetAsync(v, 0, bytes, f->stream);
    if (e != hipSuccess) return set_err(CZ_EHIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    return CZ_OK;
}

// RAII-free launch helper: optional event pair around a launch, by label.
// Profiling of one launch: the two events go into the kernel's own dispatch
// (hipExtLaunchKernelGGL), which stamps them when the kernel starts and
// ends, so the time read back is the kernel's, without the event packets a
// hipEventRecord on either side would add to it.
struct Launch {
    cz_film* f;
    const char* label;
    hipEvent_t a = nullptr, b = nullptr;
    Launch(cz_film* ff, const char* l) : f(ff), label(l) {
        if (f->prof_on && (f->prof_only.empty() || f->prof_only == l)) {
            a = f->take_event();
            b = f->take_event();
            if (!a || !b) {
                if (a) f->ev_pool.push_back(a);
                if (b) f->ev_pool.push_back(b);
                a = b = nullptr;
            }
        }
    }
    ~Launch() {
        if (a && b) f->prof_pending.push_back({label, a, b});
    }
};

void fold_profile(cz_film* f) {
    for (auto& s : f->prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) {
            auto& acc = f->prof_acc[s.label];
            acc.first += (double)ms * 1000.0;
            acc.second += 1;
        }
        f->ev_pool.push_back(s.a);
        f->ev_pool.push_back(s.b);
    }
    f->prof_pending.clear();
    fold_tuners(f->tuner_pending, f->ev_pool);
}

}  // namespace

namespace czi {

ShapeTuner* device_substep_tuner(int device) {
    static std::mutex m;
    static std::map<int, ShapeTuner*> books;  // held until the process ends
    std::lock_guard<std::mutex> g(m);
    ShapeTuner*& t = books[device];
    if (!t) {
        t = new ShapeTuner();
        t->ns = kSubstepShapes;
    }
    return t;
}

void fold_tuners(std::vector<TunerSlot>& pending, std::vector<hipEvent_t>& pool) {
    for (auto& s : pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) s.t->record(s.bucket, s.shape, (double)ms * 1000.0, s.items);
        pool.push_back(s.a);
        pool.push_back(s.b);
    }
    pending.clear();
}

void update_params(cz_film* f, float frame_dt, uint32_t substeps) {
    if (f->ov.no_implicit) f->s.implicit_hydro = 0;
    cz_host_state hs = {f->h_max, f->mu_min, f->h_max_pending, 0u};
    cz_update_params_host(&f->s, &f->ov, f->W, f->H, frame_dt, substeps, &hs, &f->P, &f->implicit_steps);
}

void fold_frame_out(cz_film* f, const float* o) {
    f->h_max = o[0];
    f->mu_min = (o[1] > 1e8f) ? 1.0f : o[1];
    f->s_max = o[8];
    f->active_tiles = cz_f2bits(o[2]);
    f->active_hint = f->active_tiles;
    for (int k = 0; k < 4; k++) f->cg_report[k] = o[4 + k];
    f->h_max_pending = 0.0f;
    f->stats_pending = false;
    uint64_t cells = cz_f2bits(o[3]);
    f->cnt_act_cells += cells;
    f->cnt_act_cell_substeps += cells * (uint64_t)f->substeps_last;
    f->cnt_act_cell_implicit += cells * (uint64_t)f->implicit_last;
    for (int k = 0; k < czi::kSubstepShapes; k++) f->cnt_shape_cell_substeps[k] += cells * (uint64_t)f->shape_subs_last[k];
    if (f->one_launch_last) {
        f->cnt_one_frames += 1;
        f->cnt_one_cells += cells;
        f->cnt_one_cell_substeps += cells * (uint64_t)f->substeps_last;
    }
    f->cnt_frames += 1;
    if (cz_f2bits(o[15]) != 0u) f->sync_fault = true;
}

int wait_idle(cz_film* f) {
    HIP_TRY(hipStreamSynchronize(f->stream));
    fold_profile(f);
    return CZ_OK;
}

int reset_solver_state(cz_film* f) {
    uint64_t n = f->n;
    HIP_TRY(hipMemsetAsync(f->pres, 0, 8 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->flux, 0, 16 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->acc, 0, 16 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->alpha, 0, 4 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->delta_a, 0, 4 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->delta_b, 0, 4 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->tmp, 0, 4 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->acc_s, 0, 64 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->strata_h2, 0, 16 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->strata_pig2, 0, 128 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->alpha_s, 0, 16 * n, f->stream));
    HIP_TRY(hipMemsetD32Async((hipDeviceptr_t)f->tile_moved, 1, f->nt16, f->stream));
    f->active_hint = f->nt16;
    return wait_idle(f);
}

void frame_done(cz_film* f, uint32_t substeps, bool parity_done) {
    if (!parity_done && (substeps & 1u)) f->delta_cur_a = !f->delta_cur_a;
    f->stats_pending = true;
    f->substeps_last = substeps;
    f->implicit_last = f->s.implicit_hydro ? f->implicit_steps : 0u;
    f->one_launch_last = f->one_launch_now;
    for (int k = 0; k < czi::kSubstepShapes; k++) f->shape_subs_last[k] = f->shape_subs_now[k];
    f->frame++;
    f->deposit_count = 0;
    f->contact_count = 0;  // the contacts are consumed by the frame
}

}  // namespace czi

using namespace czi;

namespace {

// poll_stats (src/film/mod.rs:935-973) on the frame scalars of the last frame.
int poll_stats(cz_film* f) {
    if (!f->stats_pending) return CZ_OK;
    HIP_TRY(hipStreamSynchronize(f->stream));
    fold_profile(f);
    fold_frame_out(f, f->frame_out_h);
    if (f->sync_fault) return set_err(CZ_EHIP, "a grid barrier of the one-launch substeps timed out: the frame is invalid");
    return CZ_OK;
}

// upload_deposits (src/film/mod.rs:1175-1224)
int upload_deposits(cz_film* f, const cz_deposit* deps, uint32_t n, uint32_t* sent) {
    int rc = wait_idle(f);  // the pinned staging may still feed the last frame's copies
    if (rc) return rc;
    uint32_t count = std::min<uint32_t>(n, CZ_MAX_DEPOSITS);
    f->deposit_count = count;
    // the tile binning on the device for many deposits (k_dep_rects, k_dep_scan,
    // k_dep_fill in the frame); the volume and height accounting stays here
    const bool on_dev = count > 0 && (f->bin_mode == 1 || (f->bin_mode < 0 && count >= czi::kDevBinMin));
    f->bin_on_device = on_dev;
    uint32_t nt = f->t64x * f->t64y;
    std::vector<uint32_t> cnt(on_dev ? 0 : nt, 0u);
    std::vector<int32_t> rect(on_dev ? 0 : (size_t)count * 4, -1);
    float tile_m = 64.0f * f->dx;
    for (uint32_t di = 0; di < count; di++) {
        const cz_deposit& d = deps[di];
        f->h_max_pending = fmaxf(f->h_max_pending, czh::max_height(d) + f->h_max);
        float rb = fmaxf(d.r0, d.r1);
        journal_touch_rect(f, fminf(d.p0[0], d.p1[0]) - rb, fminf(d.p0[1], d.p1[1]) - rb, fmaxf(d.p0[0], d.p1[0]) + rb,
                           fmaxf(d.p0[1], d.p1[1]) + rb);
        bool inside = fmaxf(d.p0[0], d.p1[0]) + rb > 0.0f && fmaxf(d.p0[1], d.p1[1]) + rb > 0.0f &&
                      fminf(d.p0[0], d.p1[0]) - rb < f->s.width_m && fminf(d.p0[1], d.p1[1]) - rb < f->s.height_m;
        if (!inside) continue;
        if ((d.kind & CZ_DEPOSIT_RECYCLED) == 0u)
            f->total_deposited += (double)d.vol * (double)czh::inside_fraction(d, f->s.width_m, f->s.height_m);
        if (on_dev) continue;
        float minx = fminf(d.p0[0], d.p1[0]) - rb, maxx = fmaxf(d.p0[0], d.p1[0]) + rb;
        float miny = fminf(d.p0[1], d.p1[1]) - rb, maxy = fmaxf(d.p0[1], d.p1[1]) + rb;
        int64_t mx = (int64_t)f->t64x - 1, my = (int64_t)f->t64y - 1;
        // ((x / tile_m).floor().max(0.0) as i64).min(m): Rust's float-to-int
        // cast saturates, so the clamp to m comes first here
        int64_t tx0 = (int64_t)fminf(fmaxf(floorf(minx / tile_m), 0.0f), (float)mx);
        int64_t tx1 = (int64_t)fminf(fmaxf(floorf(maxx / tile_m), 0.0f), (float)mx);
        int64_t ty0 = (int64_t)fminf(fmaxf(floorf(miny / tile_m), 0.0f), (float)my);
        int64_t ty1 = (int64_t)fminf(fmaxf(floorf(maxy / tile_m), 0.0f), (float)my);
        if (maxx < 0.0f || maxy < 0.0f) continue;
        rect[4 * di] = (int32_t)tx0; rect[4 * di + 1] = (int32_t)tx1;
        rect[4 * di + 2] = (int32_t)ty0; rect[4 * di + 3] = (int32_t)ty1;
        for (int64_t ty = ty0; ty <= ty1; ty++)
            for (int64_t tx = tx0; tx <= tx1; tx++) cnt[(size_t)(ty * f->t64x + tx)]++;
    }
    if (count > 0) {  // the deposit pass runs only when there are deposits
        memcpy(f->dep_h, deps, (size_t)count * sizeof(cz_deposit));
        HIP_TRY(hipMemcpyAsync(f->deposits, f->dep_h, (size_t)count * sizeof(cz_deposit), hipMemcpyHostToDevice, f->stream));
        if (on_dev) {
            *f->dep_n_h = count;
            HIP_TRY(hipMemcpyAsync(f->dep_n, f->dep_n_h, sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
        }
    }
    if (count > 0 && !on_dev) {
        uint32_t cap = CZ_MAX_TILE_ITEMS, items = 0;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t take = std::min(cnt[t], cap - items);
            f->tstart_h[t] = items;
            f->tcount_h[t] = take;
            items += take;
        }
        std::vector<uint32_t> fill(nt, 0u);
        for (uint32_t di = 0; di < count; di++) {
            if (rect[4 * di] < 0) continue;
            for (int32_t ty = rect[4 * di + 2]; ty <= rect[4 * di + 3]; ty++)
                for (int32_t tx = rect[4 * di]; tx <= rect[4 * di + 1]; tx++) {
                    uint32_t t = (uint32_t)ty * f->t64x + (uint32_t)tx;
                    if (fill[t] < f->tcount_h[t]) f->titems_h[f->tstart_h[t] + fill[t]++] = di;
                }
        }
        HIP_TRY(hipMemcpyAsync(f->tile_start, f->tstart_h, nt * sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
        HIP_TRY(hipMemcpyAsync(f->tile_count, f->tcount_h, nt * sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
        if (items > 0) HIP_TRY(hipMemcpyAsync(f->tile_items, f->titems_h, items * sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
    }
    if (sent) *sent = count;
    return CZ_OK;
}

#define LAUNCH(label, kern, grid, block, ...)                                                    \
    do {                                                                                         \
        Launch _l(f, label);                                                                     \
        if (_l.a)                                                                                \
            hipExtLaunchKernelGGL(kern, grid, block, 0, f->stream, _l.a, _l.b, 0u, __VA_ARGS__); \
        else                                                                                     \
            hipLaunchKernelGGL(kern, grid, block, 0, f->stream, __VA_ARGS__);                    \
    } while (0)

// encode_frame (src/film/mod.rs:1322-1462): the launches of one frame.
int record_frame(cz_film* f, uint32_t substeps) {
    drop_stale_error();
    const dim3 b256(256);
    const dim3 gt(f->tile_grid());
    const uint32_t n64 = f->t64x * f->t64y;
    FilmDev F = f->dev();
    // the journal pass comes first (src/film/mod.rs:1335-1338)
    if (f->j_gen != 0u) {
        Launch _l(f, "journal");
        journal_record(f);
    }
    if (f->contact_count > 0) {
        Launch _l(f, "brush");
        tools_record_brush(f);
    }
    if (f->deposit_count > 0) {
        if (f->bin_on_device) {
            const uint32_t nt64 = f->t64x * f->t64y;
            LAUNCH("dep_bin", k_dep_rects, dim3(kDepRectBlocks), b256, F, f->s.width_m, f->s.height_m, 64.0f * f->dx);
            LAUNCH("dep_bin", k_dep_scan, dim3(1), dim3(1024), F);
            LAUNCH("dep_bin", k_dep_fill, dim3((nt64 + 3u) / 4u), b256, F);
        }
        const dim3 gcell((f->W + 15u) / 16u, (f->H + 15u) / 16u);
        LAUNCH("deposit", k_deposit, gcell, b256, F);
    }
    if (f->knife_active && f->knife_grid[0] > 0u) {
        Launch _l(f, "knife");
        tools_record_knife(f);
    }
    LAUNCH("mark", k_mark<OneFilmAll>, dim3((f->nt16 + 3u) / 4u), b256, OneFilmAll{F, f->nt16});
    LAUNCH("compact", k_compact_bits, dim3((f->nt16 + 255u) / 256u), b256, F);
    LAUNCH("compact_list", k_compact_list, dim3((f->nt16 + 255u) / 256u), b256, F);
    LAUNCH("cure", k_cure<OneFilmCure>, dim3(std::min(f->nt16, czi::kMaxTileBlocks)), b256, OneFilmCure{F});
    const bool implicit = f->s.implicit_hydro && f->implicit_steps > 0;
    if (implicit) {
        for (uint32_t it = 0; it < f->implicit_steps; it++) {
            const OneFilm w{f->dev()};
            const OneCanvas c1{w.F};
            LAUNCH("lap", k_lap_tail<OneFilm>, gt, b256, w, 0, it == 0 ? 1 : 0);
            LAUNCH("cg_setup", k_cg_setup<OneFilm>, gt, b256, w);
            LAUNCH("cg_lap", k_cg_lap<OneFilm>, gt, b256, w, 1u);
            LAUNCH("cg_init", k_cg_init<OneFilm>, gt, b256, w);
            LAUNCH("cg_reduce", k_cg_reduce<OneCanvas>, dim3(1), b256, c1, 0);
            for (uint32_t k = 0; k < f->s.implicit_iters; k++) {
                LAUNCH("cg_lap", k_cg_lap<OneFilm>, gt, b256, w, 0u);
                LAUNCH("cg_iter_a", k_cg_iter_a<OneFilm>, gt, b256, w);
                LAUNCH("cg_reduce", k_cg_reduce<OneCanvas>, dim3(1), b256, c1, 1);
                LAUNCH("cg_iter_b", k_cg_iter_b<OneFilm>, gt, b256, w);
                LAUNCH("cg_reduce", k_cg_reduce<OneCanvas>, dim3(1), b256, c1, 2);
                LAUNCH("cg_iter_c", k_cg_iter_c<OneFilm>, gt, b256, w);
            }
            LAUNCH("cg_lap", k_cg_lap<OneFilm>, gt, b256, w, 4u);
            LAUNCH("cg_finish", k_cg_finish<OneFilm>, gt, b256, w);
            LAUNCH("cg_reduce", k_cg_reduce<OneCanvas>, dim3(1), b256, c1, 3);
            LAUNCH("cg_rescale", k_cg_rescale<OneFilm>, gt, b256, w);
        }
    }
    const uint32_t pg = f->persist_grid();
    f->one_launch_now = pg > 0u;
    for (uint32_t& v : f->shape_subs_now) v = 0u;
    if (pg > 0u) {
        // every substep and the last nose in one launch
        const FilmDev Fa = f->dev();
        f->delta_cur_a = !f->delta_cur_a;
        const FilmDev Fb = f->dev();
        f->delta_cur_a = !f->delta_cur_a;
        LAUNCH("substeps", k_substeps, dim3(pg), b256, Fa, Fb, substeps, implicit ? 0 : 1);
        // the host's offset-buffer parity as the per-pass launches leave it:
        // one flip per nose (substeps 1.. and the last one)
        if (substeps & 1u) f->delta_cur_a = !f->delta_cur_a;
    }
    // substeps: [nose of the previous substep +] lap + plug, flux, limiter + apply,
    // in the launch shape the substep tuner picks for the active-set size (it
    // times the whole substep)
    const dim3 b1024(1024);
    const uint32_t nsub = pg > 0u ? 0u : substeps;
    std::vector<int> shp(nsub), bks(nsub);
    std::vector<char> tim(nsub), to_main(nsub, 1);
    const bool slump = f->s.slump != 0u;
    for (uint32_t s = 0; s < nsub; s++) {
        bool timed = false;
        // profiled frames time every launch: their substeps carry the
        // events' gaps, so they take the tuner's choice and add no samples
        shp[s] = f->flux_tuner.choose(std::min(f->active_hint, f->nt16), !f->capturing && !f->prof_on, bks[s], timed);
        tim[s] = timed ? 1 : 0;
        // the fused shape carries the nose into its substep; with the slump
        // passes between the substeps the passes run separately
        if (slump && shp[s] == 5) {
            if (timed) f->flux_tuner.reissue(bks[s], 5, 4);
            shp[s] = 4;
        }
    }
    // The fused shape (5) reads the heights before its apply from one of F.h
    // and h_sub and writes the new ones to the other.  A run of fused
    // substeps starts from F.h (the frame's start or a separate-pass
    // substep) and must end in F.h (the next separate-pass substep or the
    // frame's end), so it alternates over an even number of substeps: a run
    // of odd length runs its first substep as separate passes (shape 4).
    for (uint32_t a = 0; a < nsub;) {
        if (shp[a] != 5) {
            a++;
            continue;
        }
        uint32_t b = a;
        while (b + 1 < nsub && shp[b + 1] == 5) b++;
        if (((b - a + 1) & 1u) != 0u) {
            if (tim[a]) f->flux_tuner.reissue(bks[a], 5, 4);
            shp[a++] = 4;
        }
        for (uint32_t s = a; s <= b && a <= b; s++) to_main[s] = ((b - s) & 1u) == 0u ? 1 : 0;
        a = b + 1;
    }
    bool cur_main = true;  // where the current heights are
    for (uint32_t s = 0; s < nsub; s++) {
        const int bk = bks[s];
        const bool timed = tim[s] != 0;
        const int shape = shp[s];
        hipEvent_t ea = nullptr, eb = nullptr;
        if (timed) {
            ea = f->take_event();
            eb = f->take_event();
            if (ea && eb) hipEventRecord(ea, f->stream);
        }
        const OneFilm wl{f->dev()};
        // with slump the previous substep's nose ran before its slump passes
        const int nose = (s > 0 && !slump) ? 1 : 0, clear = (s == 0 && !implicit) ? 1 : 0;
        if (shape == 5) {
            if (cur_main == (to_main[s] != 0)) return set_err(CZ_EINVAL, "fused substep: height buffers collide");
            float* hr = cur_main ? f->h : f->h_sub;
            float* hw = to_main[s] ? f->h : f->h_sub;
            LAUNCH("substep_fused", k_substep_fused<OneFilm>, gt, b1024, wl, (const float*)hr, hw, nose, clear);
            if (nose) f->delta_cur_a = !f->delta_cur_a;
            cur_main = to_main[s] != 0;
        } else {
        if (shape >= 3)
            LAUNCH("lap_plug_w", k_lap_plug_w<OneFilm>, gt, b1024, wl, nose, 1, clear);
        else
            LAUNCH("lap_plug", k_lap_plug<OneFilm>, gt, b256, wl, nose, 1, clear);
        if (nose) f->delta_cur_a = !f->delta_cur_a;
        const OneFilm w{f->dev()};
        switch (shape) {
            case 1:
                LAUNCH("flux", k_fluxes_faces<OneFilm>, dim3(f->tile_grid(4)), b256, w);
                LAUNCH("apply", k_limit_apply<OneFilm>, gt, b256, w);
                break;
            case 2:
                LAUNCH("flux_apply", k_flux_apply<OneFilm>, gt, b256, w);
                break;
            case 3:
                LAUNCH("flux_w", k_fluxes_w<OneFilm>, gt, b1024, w);
                LAUNCH("apply_w", k_limit_apply_w<OneFilm>, gt, b1024, w);
                break;
            case 4:
                LAUNCH("flux_apply_w", k_flux_apply_w<OneFilm>, gt, b1024, w);
                break;
            case 6:
                LAUNCH("flux", k_fluxes_faces<OneFilm>, dim3(f->tile_grid(4)), b256, w);
                LAUNCH("apply_w", k_limit_apply_w<OneFilm>, gt, b1024, w);
                break;
            default:
                LAUNCH("flux", k_fluxes<OneFilm>, gt, b256, w);
                LAUNCH("apply", k_limit_apply<OneFilm>, gt, b256, w);
                break;
        }
        cur_main = true;
        if (slump) {
            // this substep's nose, then shed and slump (+ commit_tmp), as
            // encode_frame orders them (src/film/mod.rs:1420-1429); in the
            // block size the substep's shape took
            if (shape >= 3)
                LAUNCH("lap", k_lap_tail_w<OneFilm>, gt, b1024, OneFilm{f->dev()}, 1, 0);
            else
                LAUNCH("lap", k_lap_tail<OneFilm>, gt, b256, OneFilm{f->dev()}, 1, 0);
            f->delta_cur_a = !f->delta_cur_a;
            const OneFilm ws{f->dev()};
            LAUNCH("shed", k_shed<OneFilm>, gt, b256, ws);
            LAUNCH("slump", k_slump<OneFilm>, gt, b256, ws);
        }
        }
        if (ea && eb) {
            hipEventRecord(eb, f->stream);
            f->tuner_pending.push_back({f->flux_tuner.t, bk, shape, ea, eb, (double)std::min(f->active_hint, f->nt16)});
        }
        f->shape_subs_now[shape]++;
    }
    // nose of the last substep (with slump: done) + the laplacian strata_split
    // reads, in 1024-thread blocks when the last substep took a wide shape
    // (a few hundred active tiles or fewer: 8-10 us against 12-14)
    if (pg == 0u) {
        if (nsub > 0 && shp[nsub - 1] >= 3)
            LAUNCH("lap", k_lap_tail_w<OneFilm>, gt, b1024, OneFilm{f->dev()}, slump ? 0 : 1, 0);
        else
            LAUNCH("lap", k_lap_tail<OneFilm>, gt, b256, OneFilm{f->dev()}, slump ? 0 : 1, 0);
        if (!slump) f->delta_cur_a = !f->delta_cur_a;
    }
    const OneFilm w{f->dev()};
    LAUNCH("split", k_strata_split<OneFilm>, dim3(f->tile_grid(4)), b256, w);
    LAUNCH("reduce_pool", k_reduce_pool<OneFilmAll>, dim3(n64), b256, OneFilmAll{w.F, n64});
    LAUNCH("strata_limit", k_strata_limit<OneFilm>, gt, b256, w);
    LAUNCH("strata_update", k_strata_update<OneFilm>, gt, b256, w);
    LAUNCH("strata_fill", k_strata_fill<OneFilm>, gt, b256, w);
    // the display pass of every frame (shade_full, src/film/mod.rs:1449-1456)
    if (f->shade_on)
        LAUNCH("shade", k_shade<OneFilmAll>, dim3(std::min(f->nt16, czi::kMaxTileBlocks)), b256, OneFilmAll{w.F, f->nt16});
    LAUNCH("frame_stats", k_frame_stats<OneCanvas>, dim3(1), dim3(1024), OneCanvas{w.F});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(f->frame_out_h, f->frame_out, 16 * sizeof(float), hipMemcpyDeviceToHost, f->stream));
    return CZ_OK;
}

// The frame as a hipGraph, captured once per frame shape and replayed: the
// shape is everything the launches depend on (substeps, implicit steps,
// deposit pass, offset-buffer parity, params, grid sizes).
struct GraphKey {
    uint32_t substeps, implicit_steps, implicit_on, has_dep, parity, grid1, grid4, flux_shape, persist, shade, journal;
    uint32_t brush, knife_gx, knife_gy;
    cz_params P;
    bool operator==(const GraphKey& o) const { return memcmp(this, &o, sizeof(GraphKey)) == 0; }
};

int encode_direct(cz_film* f, uint32_t substeps) {
    int rc = record_frame(f, substeps);
    if (rc) return rc;
    frame_done(f, substeps, true);
    return CZ_OK;
}

int encode_frame(cz_film* f, uint32_t substeps) {
    if (f->sync_fault) return set_err(CZ_EHIP, "the film's last one-launch substeps failed; it takes no more frames");
    int jrc = journal_flush(f);  // flush_journal (src/main.rs:1894) before the frame
    if (jrc) return jrc;
    const uint64_t flux_items = std::min(f->active_hint, f->nt16);
    const uint32_t pg = f->persist_grid();
    if (!f->use_graphs || f->prof_on || flux_items > czi::kGraphMaxTiles ||
        (pg == 0u && f->flux_tuner.learning(flux_items)))
        return encode_direct(f, substeps);
    GraphKey key;
    memset(&key, 0, sizeof(key));
    key.substeps = substeps;
    key.implicit_steps = f->implicit_steps;
    key.implicit_on = f->s.implicit_hydro;
    key.has_dep = f->deposit_count == 0 ? 0u : (f->bin_on_device ? 2u : 1u);
    key.parity = f->delta_cur_a ? 1u : 0u;
    key.grid1 = f->tile_grid(1);
    key.grid4 = f->tile_grid(4);
    int bk = 0;
    bool timed = false;
    key.flux_shape = pg > 0u ? 0u : (uint32_t)f->flux_tuner.choose(flux_items, false, bk, timed);
    key.persist = pg;
    key.shade = f->shade_on ? 1u : 0u;
    key.journal = f->j_gen != 0u ? 1u : 0u;
    key.brush = f->contact_count > 0 ? 1u : 0u;
    key.knife_gx = f->knife_active ? f->knife_grid[0] : 0u;
    key.knife_gy = f->knife_active ? f->knife_grid[1] : 0u;
    key.P = f->P;
    hipGraphExec_t exec = nullptr;
    for (size_t k = 0; k < f->graphs.size(); k++) {
        if (memcmp(&f->graphs[k].first, &key, sizeof(key)) == 0) {
            exec = f->graphs[k].second;
            // move to the back (most recently used)
            auto e = f->graphs[k];
            f->graphs.erase(f->graphs.begin() + (long)k);
            f->graphs.push_back(e);
            break;
        }
    }
    bool start_a = f->delta_cur_a;
    if (!exec) {
        // a shape is captured the second time it comes: one-off shapes (a
        // changing active set, a deposit frame) run direct
        GraphKeyStore ks;
        memcpy(&ks, &key, sizeof(key));
        bool seen = false;
        for (const GraphKeyStore& o : f->graph_seen) seen = seen || memcmp(&o, &ks, sizeof(ks)) == 0;
        if (!seen) {
            if (f->graph_seen.size() >= 64) f->graph_seen.erase(f->graph_seen.begin());
            f->graph_seen.push_back(ks);
            return encode_direct(f, substeps);
        }
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
        f->capturing = true;
        int rc = record_frame(f, substeps);
        f->capturing = false;
        hipError_t e = hipStreamEndCapture(f->stream, &graph);
        if (rc) return rc;
        if (e != hipSuccess) return set_err(CZ_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return set_err(CZ_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        if (f->graphs.size() >= 16) {
            hipGraphExecDestroy(f->graphs.front().second);
            f->graphs.erase(f->graphs.begin());
        }
        f->graph_captures++;
        GraphKeyStore kk;
        static_assert(sizeof(GraphKeyStore) == sizeof(GraphKey), "graph key size");
        memcpy(&kk, &key, sizeof(key));
        f->graphs.push_back({kk, exec});
        f->delta_cur_a = start_a;  // record_frame flipped it while capturing
    }
    HIP_TRY(hipGraphLaunch(exec, f->stream));
    f->graph_replays++;
    frame_done(f, substeps, false);
    return CZ_OK;
}

struct BufDesc {
    void* ptr;
    uint64_t bytes;
};

BufDesc buffer_of(const cz_film* f, int which) {
    uint64_t n = f->n;
    switch (which) {
    case CZ_PART_H: return {f->h, 4 * n};
    case CZ_PART_M: return {f->m, 4 * n};
    case CZ_PART_RHEO: return {f->rheo, 8 * n};
    case CZ_PART_STRATA_H: return {f->strata_h, 16 * n};
    case CZ_PART_STRATA_PIG: return {f->strata_pig, 128 * n};
    case CZ_PART_STRATA_CURE: return {f->strata_cure, 16 * n};
    case CZ_PART_SUB_REFL: return {f->sub_refl, 32 * n};
    case CZ_PART_LAY: return {f->lay, 4 * n};
    case CZ_BUF_DELTA: return {f->delta_cur(), 4 * n};
    case CZ_BUF_FNORM: return {f->fnorm, 8 * n};
    case CZ_BUF_PRES: return {f->pres, 8 * n};
    case CZ_BUF_H_FRAME: return {f->h_frame, 4 * n};
    case CZ_BUF_CG_M: return {f->cg_m, 16 * n};
    case CZ_BUF_CG_P: return {f->cg_p, 4 * n};
    case CZ_BUF_TILE_MOVED: return {f->tile_moved, 4 * (uint64_t)f->nt16};
    case CZ_BUF_TILE_FLAGS: return {f->tile_flags, 4 * (uint64_t)f->nt16};
    case CZ_BUF_TMP: return {f->tmp, 4 * n};
    case CZ_BUF_FLUX: return {f->flux, 16 * n};
    case CZ_BUF_ACC: return {f->acc, 16 * n};
    case CZ_BUF_ALPHA: return {f->alpha, 4 * n};
    case CZ_BUF_STATS: return {f->stats_out, 8 * (uint64_t)f->t64x * f->t64y};
    case CZ_BUF_POOL: return {f->pool_out, 8 * (uint64_t)f->pw * f->ph};
    case CZ_BUF_TILE_LIST: return {f->tile_list, 4 * (uint64_t)f->nt16};
    case CZ_BUF_ACC_S: return {f->acc_s, 64 * n};
    case CZ_BUF_ALPHA_S: return {f->alpha_s, 16 * n};
    case CZ_BUF_DEP_START: return {f->tile_start, 4 * (uint64_t)f->t64x * f->t64y};
    case CZ_BUF_DEP_COUNT: return {f->tile_count, 4 * (uint64_t)f->t64x * f->t64y};
    case CZ_BUF_DEP_ITEMS: return {f->tile_items, 4 * (uint64_t)CZ_MAX_TILE_ITEMS};
    case CZ_BUF_COLOR: return {f->color, f->color ? 8 * n : 0};
    case CZ_BUF_DETAIL: return {f->detail, f->detail ? 8 * n : 0};
    case CZ_BUF_JMETA: return {f->j_cap ? f->j_meta : nullptr, 16 * (uint64_t)f->j_cap};
    case CZ_BUF_JMARK: return {f->j_cap ? f->j_mark : nullptr, f->j_cap ? 4 * (uint64_t)f->nt16 : 0};
    case CZ_BUF_JDATA: return {f->j_cap ? f->j_data : nullptr, 53 * 256 * 4 * (uint64_t)f->j_cap};
    case CZ_BUF_N_CELL: return {f->n_cell, 4 * n};
    default: return {nullptr, 0};
    }
}

// The initial contents of Film::new / clear (src/film/mod.rs:686-693, 1009-1028).
int init_state(cz_film* f) {
    std::vector<float> wv((size_t)f->n);
    cz_weave(f->W, f->H, f->dx, f->s.weave_um * 1e-6f, wv.data());
    double wsum = 0.0;
    for (float v : wv) wsum += (double)v;
    f->weave_volume = wsum * (double)f->dx * (double)f->dx;
    HIP_TRY(hipMemcpyAsync(f->m, wv.data(), wv.size() * 4, hipMemcpyHostToDevice, f->stream));
    std::vector<uint32_t> white((size_t)f->n * 8);
    uint32_t words[8];
    cz_substrate_words(&f->s, f->cmf_h, words);
    for (uint64_t c = 0; c < f->n; c++) memcpy(&white[c * 8], words, sizeof(words));
    HIP_TRY(hipMemcpyAsync(f->sub_refl, white.data(), white.size() * 4, hipMemcpyHostToDevice, f->stream));
    std::vector<float2> ones((size_t)f->n, make_float2(1.0f, 0.0f));
    HIP_TRY(hipMemcpyAsync(f->rheo, ones.data(), ones.size() * 8, hipMemcpyHostToDevice, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    return CZ_OK;
}

void destroy(cz_film* f) {
    if (!f) return;
    batch_detach(f);
    // the last frames' tuner samples go to the device's tuner
    if (f->stream && hipStreamSynchronize(f->stream) == hipSuccess) fold_tuners(f->tuner_pending, f->ev_pool);
    for (void* p : f->allocs) hipFree(p);
    if (f->frame_out_h) hipHostFree(f->frame_out_h);
    if (f->dep_h) hipHostFree(f->dep_h);
    if (f->tstart_h) hipHostFree(f->tstart_h);
    if (f->tcount_h) hipHostFree(f->tcount_h);
    if (f->titems_h) hipHostFree(f->titems_h);
    if (f->dep_n_h) hipHostFree(f->dep_n_h);
    journal_free_host(f);
    tools_free_host(f);
    for (auto& s : f->prof_pending) {
        hipEventDestroy(s.a);
        hipEventDestroy(s.b);
    }
    for (auto& t : f->tuner_pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    for (auto e : f->ev_pool) hipEventDestroy(e);
    for (auto& g : f->graphs) hipGraphExecDestroy(g.second);
    if (f->stream) hipStreamDestroy(f->stream);
    delete f;
}

}  // namespace

extern "C" {

const char* cz_last_error(void) { return g_err; }
const char* cz_version(void) { return "canvaz-film-mi355x 0.1 (gfx950)"; }

}  // extern "C"

namespace {

bool set_dims(cz_film* f, uint32_t w, uint32_t h) {
    f->W = w; f->H = h; f->n = w * h;
    f->tw = (w + 15u) / 16u; f->th = (h + 15u) / 16u; f->nt16 = f->tw * f->th;
    f->t64x = (w + 63u) / 64u; f->t64y = (h + 63u) / 64u;
    f->pw = (w + 3u) / 4u; f->ph = (h + 3u) / 4u;
    f->dx = f->s.width_m / (float)w;
    return f->n <= f->cap_n && f->nt16 <= f->cap_nt16 && f->t64x * f->t64y <= f->cap_nt64 &&
           f->pw * f->ph <= f->cap_np;
}

// Host state of a film that starts afresh (Film::new): lag-1 stats,
// counters, parities, params of update_params(1/60, 1).
void reset_host_state(cz_film* f) {
    f->h_max = 0.0f; f->s_max = 0.0f; f->mu_min = 1.0f; f->h_max_pending = 0.0f;
    f->implicit_steps = 1; f->substeps_last = 1; f->implicit_last = 0; f->frame = 0; f->active_tiles = 0;
    f->active_hint = f->nt16;
    f->total_deposited = 0.0;
    for (int k = 0; k < 4; k++) f->cg_report[k] = 0.0f;
    f->stats_pending = false;
    f->cnt_act_cells = f->cnt_act_cell_substeps = f->cnt_frames = f->cnt_act_cell_implicit = 0;
    f->cnt_one_frames = f->cnt_one_cells = f->cnt_one_cell_substeps = 0;
    for (uint64_t& v : f->cnt_shape_cell_substeps) v = 0;
    f->delta_cur_a = true;
    f->deposit_count = 0;
    memset(&f->P, 0, sizeof(f->P));
    f->P.size[0] = f->W; f->P.size[1] = f->H;
    f->P.tiles[0] = f->t64x; f->P.tiles[1] = f->t64y;
    f->P.dx = f->dx; f->P.inv_dx = 1.0f / f->dx;
    update_params(f, 1.0f / 60.0f, 1);
}

}  // namespace

namespace czi {

int film_create(uint32_t w, uint32_t h, const cz_settings* s, const float* pigtab, const float* cmf, bool init,
                cz_film** out) {
    if (!out || w == 0 || h == 0) return set_err(CZ_EINVAL, "cz_film_create: bad size %ux%u", w, h);
    if ((uint64_t)w * h > (1ull << 31)) return set_err(CZ_EINVAL, "cz_film_create: %ux%u cells exceed 2^31", w, h);
    if ((w + 15u) / 16u > 21845u) return set_err(CZ_EINVAL, "cz_film_create: width %u exceeds 349520 cells", w);
    *out = nullptr;
    cz_film* f = new cz_film();
    if (s) f->s = *s; else cz_settings_default(&f->s);
    cz_overrides_from_env(&f->ov);
    f->cap_n = w * h;
    f->cap_nt16 = ((w + 15u) / 16u) * ((h + 15u) / 16u);
    f->cap_nt64 = ((w + 63u) / 64u) * ((h + 63u) / 64u);
    f->cap_np = ((w + 3u) / 4u) * ((h + 3u) / 4u);
    set_dims(f, w, h);
    if (pigtab) memcpy(f->pigtab_h, pigtab, sizeof(f->pigtab_h));
    else {
        float refl[12 * 16], sc[12], ti[12], vi[12], yi[12];
        cz_default_pigments(refl, sc, ti, vi, yi);
        cz_pigment_table(refl, sc, ti, vi, yi, f->pigtab_h);
    }
    if (cmf) memcpy(f->cmf_h, cmf, sizeof(f->cmf_h));
    else cz_cmf_table(f->cmf_h);
    int rc = CZ_OK;
    hipError_t e = hipGetDevice(&f->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        destroy(f);
        return set_err(CZ_EHIP, "cz_film_create: %s", hipGetErrorString(e));
    }
    f->flux_tuner.t = czi::device_substep_tuner(f->device);
    size_t n = f->n, nt16 = f->nt16, nt64 = (size_t)f->t64x * f->t64y, np = (size_t)f->pw * f->ph;
#define A(p, cnt) if (!rc) rc = dalloc(f, &f->p, (cnt))
    A(h, n); A(m, n); A(rheo, n); A(strata_h, n); A(strata_pig, 8 * n); A(strata_cure, n); A(sub_refl, 8 * n);
    A(lay, n); A(pres, n); A(flux, n); A(acc, n); A(alpha, n); A(h_sub, n); A(delta_a, n); A(delta_b, n); A(tmp, n);
    A(acc_s, 4 * n); A(alpha_s, n); A(strata_h2, n); A(strata_pig2, 8 * n); A(strata_cure2, n); A(h_frame, n);
    A(fnorm, n); A(cg_m, n); A(cg_d, n); A(cg_x, n); A(cg_r, n); A(cg_p, n); A(cg_q, n); A(cg_t, n); A(cg_t2, n);
    A(cg_part, nt16); A(cg_part2, nt16); A(cg_s, 8); A(tile_flags, nt16); A(tile_list, nt16); A(tile_moved, nt16);
    A(active_count, 4); A(cure_list, nt16); A(cure_count, 4); A(grid_sync, 4); A(tile_bits, nt16); A(dep_rect, 4 * (size_t)CZ_MAX_DEPOSITS); A(dep_cnt, nt64); A(dep_n, 4); A(cmf, CZ_NBAND * 3); A(compact_cnt, 2 * ((nt16 + 255) / 256) + 2); A(stats_out, nt64); A(smax_out, nt64); A(pool_out, 2 * np); A(tsum_a, n);
    A(pigtab, CZ_NPIG * CZ_PIG_STRIDE); A(deposits, CZ_MAX_DEPOSITS); A(tile_start, nt64); A(tile_count, nt64);
    A(tile_items, CZ_MAX_TILE_ITEMS); A(frame_out, 16);
    A(mass_ws, (cz_mass_sums_workspace(n) + 15) / 16); A(mass_out, 1); A(n_cell, n); A(lmu_cell, n);
#undef A
    if (!rc) {
        e = hipHostMalloc((void**)&f->frame_out_h, 16 * sizeof(float), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&f->dep_h, CZ_MAX_DEPOSITS * sizeof(cz_deposit), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&f->tstart_h, nt64 * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&f->tcount_h, nt64 * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&f->titems_h, CZ_MAX_TILE_ITEMS * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&f->dep_n_h, 4, hipHostMallocDefault);
        if (e != hipSuccess) rc = set_err(CZ_ENOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
    }
    if (!rc) {
        e = hipMemcpyAsync(f->pigtab, f->pigtab_h, sizeof(f->pigtab_h), hipMemcpyHostToDevice, f->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(f->cmf, f->cmf_h, sizeof(f->cmf_h), hipMemcpyHostToDevice, f->stream);
        if (e != hipSuccess) rc = set_err(CZ_EHIP, "table upload: %s", hipGetErrorString(e));
    }
    if (!rc && init) rc = init_state(f);
    if (!rc && !init) {
        e = hipStreamSynchronize(f->stream);
        if (e != hipSuccess) rc = set_err(CZ_EHIP, "cz_film_create: %s", hipGetErrorString(e));
    }
    if (rc) {
        destroy(f);
        return rc;
    }
    reset_host_state(f);
    *out = f;
    return CZ_OK;
}

int film_reshape(cz_film* f, uint32_t w, uint32_t h, const cz_settings* s) {
    if (f->batch) return set_err(CZ_EINVAL, "reshape: the film is a batch member");
    int rc = wait_idle(f);
    if (rc) return rc;
    const cz_settings keep = f->s;
    f->s = *s;
    if (w == 0 || h == 0 || !set_dims(f, w, h)) {
        f->s = keep;
        set_dims(f, f->W, f->H);
        return set_err(CZ_ESIZE, "reshape: %ux%u exceeds the film's capacity", w, h);
    }
    for (auto& g : f->graphs) hipGraphExecDestroy(g.second);
    f->graphs.clear();
    f->graph_seen.clear();
    f->weave_volume = 0.0;
    reset_host_state(f);
    // buffers a fresh film starts with at zero and may read at cells outside
    // the active tiles before writing them (the others are written first or
    // cleared with the state parts)
    const uint64_t n = f->n;
    HIP_TRY(hipMemsetAsync(f->fnorm, 0, 8 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->cg_m, 0, 16 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->cg_p, 0, 4 * n, f->stream));
    HIP_TRY(hipMemsetAsync(f->h_frame, 0, 4 * n, f->stream));
    rc = journal_reset(f);
    if (!rc) rc = tools_reset(f);
    if (rc) return rc;
    return wait_idle(f);
}

namespace {
struct SnapPart {
    void* p;
    uint64_t b;
};
std::vector<SnapPart> snapshot_parts(cz_film* f) {
    const uint64_t n = f->n;
    std::vector<SnapPart> v;
    for (int k = 0; k < CZ_NPARTS; k++) {
        BufDesc d = buffer_of(f, k);
        v.push_back({d.ptr, d.bytes});
    }
    v.push_back({f->delta_a, 4 * n});
    v.push_back({f->delta_b, 4 * n});
    v.push_back({f->fnorm, 8 * n});
    v.push_back({f->pres, 8 * n});
    v.push_back({f->tmp, 4 * n});
    v.push_back({f->h_frame, 4 * n});
    v.push_back({f->cg_m, 16 * n});
    v.push_back({f->cg_p, 4 * n});
    v.push_back({f->acc, 16 * n});
    v.push_back({f->n_cell, 4 * n});
    v.push_back({f->lmu_cell, 4 * n});
    v.push_back({f->tile_moved, 4 * (uint64_t)f->nt16});
    v.push_back({f->tile_flags, 4 * (uint64_t)f->nt16});
    return v;
}
}  // namespace

uint64_t film_snapshot_bytes(cz_film* f) {
    uint64_t total = 0;
    for (const SnapPart& q : snapshot_parts(f)) total += q.b;
    return total;
}

int film_snapshot(cz_film* f, FilmSnapshot& out) {
    if (f->batch) return set_err(CZ_EINVAL, "snapshot: the film is a batch member");
    int rc = wait_idle(f);
    if (rc) return rc;
    if (f->stats_pending) fold_frame_out(f, f->frame_out_h);
    const std::vector<SnapPart> parts = snapshot_parts(f);
    uint64_t total = 0;
    for (const SnapPart& q : parts) total += q.b;
    out.data.resize(total);
    uint64_t off = 0;
    for (const SnapPart& q : parts) {
        HIP_TRY(hipMemcpyAsync(out.data.data() + off, q.p, q.b, hipMemcpyDeviceToHost, f->stream));
        off += q.b;
    }
    out.W = f->W;
    out.H = f->H;
    out.s = f->s;
    out.h_max = f->h_max;
    out.mu_min = f->mu_min;
    out.h_max_pending = f->h_max_pending;
    out.s_max = f->s_max;
    out.implicit_steps = f->implicit_steps;
    out.substeps_last = f->substeps_last;
    out.implicit_last = f->implicit_last;
    out.frame = f->frame;
    out.active_tiles = f->active_tiles;
    out.active_hint = f->active_hint;
    out.delta_cur_a = f->delta_cur_a;
    out.total_deposited = f->total_deposited;
    out.weave_volume = f->weave_volume;
    return wait_idle(f);
}

int film_restore(cz_film* f, const FilmSnapshot& snap) {
    if (f->batch) return set_err(CZ_EINVAL, "restore: the film is a batch member");
    if (f->W != snap.W || f->H != snap.H) return set_err(CZ_ESIZE, "restore: the film is %ux%u, the snapshot %ux%u",
                                                         f->W, f->H, snap.W, snap.H);
    int rc = wait_idle(f);
    if (rc) return rc;
    const std::vector<SnapPart> parts = snapshot_parts(f);
    uint64_t total = 0;
    for (const SnapPart& q : parts) total += q.b;
    if (total != snap.data.size()) return set_err(CZ_ESIZE, "restore: snapshot of %llu bytes, the film needs %llu",
                                                  (unsigned long long)snap.data.size(), (unsigned long long)total);
    uint64_t off = 0;
    for (const SnapPart& q : parts) {
        HIP_TRY(hipMemcpyAsync(q.p, snap.data.data() + off, q.b, hipMemcpyHostToDevice, f->stream));
        off += q.b;
    }
    f->h_max = snap.h_max;
    f->mu_min = snap.mu_min;
    f->h_max_pending = snap.h_max_pending;
    f->s_max = snap.s_max;
    f->implicit_steps = snap.implicit_steps;
    f->substeps_last = snap.substeps_last;
    f->implicit_last = snap.implicit_last;
    f->frame = snap.frame;
    f->active_tiles = snap.active_tiles;
    f->active_hint = snap.active_hint;
    f->delta_cur_a = snap.delta_cur_a;
    f->total_deposited = snap.total_deposited;
    f->weave_volume = snap.weave_volume;
    f->stats_pending = false;
    return wait_idle(f);
}

}  // namespace czi

extern "C" {

int cz_film_create(uint32_t w, uint32_t h, const cz_settings* s, const float* pigtab, const float* cmf, cz_film** out) {
    return film_create(w, h, s, pigtab, cmf, true, out);
}

int cz_film_reshape(cz_film* f, uint32_t w, uint32_t h, const cz_settings* s) {
    if (!f || !s) return set_err(CZ_EINVAL, "null argument");
    return film_reshape(f, w, h, s);
}

void cz_film_destroy(cz_film* f) { destroy(f); }

int cz_film_dims(const cz_film* f, uint32_t* w, uint32_t* h, float* dx) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    if (w) *w = f->W;
    if (h) *h = f->H;
    if (dx) *dx = f->dx;
    return CZ_OK;
}

int cz_film_set_settings(cz_film* f, const cz_settings* s) {
    if (!f || !s) return set_err(CZ_EINVAL, "null argument");
    if (s->width_m != f->s.width_m) return set_err(CZ_EINVAL, "width_m is fixed at creation");
    if (f->batch && s->slump) return set_err(CZ_EUNSUP, "set_settings: a batch member takes no slump pass");
    f->s = *s;
    return CZ_OK;
}

int cz_film_get_settings(const cz_film* f, cz_settings* s) {
    if (!f || !s) return set_err(CZ_EINVAL, "null argument");
    *s = f->s;
    return CZ_OK;
}

int cz_film_set_overrides(cz_film* f, const cz_overrides* o) {
    if (!f || !o) return set_err(CZ_EINVAL, "null argument");
    f->ov = *o;
    f->ov_from_env = false;
    return CZ_OK;
}

// gravity_vec (src/film/mod.rs:861-867)
int cz_film_gravity_vec(const cz_film* f, float out[3]) {
    if (!f || !out) return set_err(CZ_EINVAL, "null argument");
    cz_gravity_vec_host(&f->s, out);
    return CZ_OK;
}

int cz_film_update_params(cz_film* f, float frame_dt, uint32_t substeps) {
    if (!f || substeps == 0 || !(frame_dt > 0.0f)) return set_err(CZ_EINVAL, "update_params: bad arguments");
    update_params(f, frame_dt, substeps);
    return CZ_OK;
}

// choose_substeps (src/film/mod.rs:919-933)
uint32_t cz_film_choose_substeps(const cz_film* f, float frame_dt) {
    cz_host_state hs = {f->h_max, f->mu_min, f->h_max_pending, 0u};
    return cz_choose_substeps_host(&f->s, f->W, frame_dt, &hs);
}

int cz_film_get_params(const cz_film* f, cz_params* p) {
    if (!f || !p) return set_err(CZ_EINVAL, "null argument");
    *p = f->P;
    return CZ_OK;
}

int cz_film_set_params(cz_film* f, const cz_params* p) {
    if (!f || !p) return set_err(CZ_EINVAL, "null argument");
    if (p->size[0] != f->W || p->size[1] != f->H) return set_err(CZ_EINVAL, "set_params: size mismatch");
    f->P = *p;
    return CZ_OK;
}

int cz_film_upload_deposits(cz_film* f, const cz_deposit* deps, uint32_t n, uint32_t* sent) {
    if (f && f->batch) return set_err(CZ_EINVAL, "upload_deposits: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f || (n > 0 && !deps)) return set_err(CZ_EINVAL, "upload_deposits: bad arguments");
    return upload_deposits(f, deps, n, sent);
}

int cz_film_encode_frame(cz_film* f, uint32_t substeps) {
    if (f && f->batch) return set_err(CZ_EINVAL, "encode_frame: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f || substeps == 0) return set_err(CZ_EINVAL, "encode_frame: bad arguments");
    return encode_frame(f, substeps);
}

// encode_sim (src/main.rs:1875-1909) with stats of the previous frame.
int cz_film_step(cz_film* f, const cz_deposit* deps, uint32_t n, float frame_dt, uint32_t* sent) {
    if (f && f->batch) return set_err(CZ_EINVAL, "step: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f || (n > 0 && !deps) || !(frame_dt > 0.0f)) return set_err(CZ_EINVAL, "step: bad arguments");
    int rc = poll_stats(f);
    if (rc) return rc;
    if (f->ov_from_env) cz_overrides_from_env(&f->ov);
    rc = upload_deposits(f, deps, n, sent);
    if (rc) return rc;
    uint32_t substeps = cz_film_choose_substeps(f, frame_dt);
    update_params(f, frame_dt, substeps);
    return encode_frame(f, substeps);
}

int cz_film_sync(cz_film* f) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    return poll_stats(f) ? CZ_EHIP : wait_idle(f);
}

int cz_film_stats(cz_film* f, cz_stats* st) {
    if (!f || !st) return set_err(CZ_EINVAL, "null argument");
    int rc = wait_idle(f);
    if (rc) return rc;
    const float* o = f->frame_out_h;
    st->h_max = f->stats_pending ? o[0] : f->h_max;
    st->mu_min = f->stats_pending ? ((o[1] > 1e8f) ? 1.0f : o[1]) : f->mu_min;
    st->active_tiles = f->stats_pending ? cz_f2bits(o[2]) : f->active_tiles;
    st->substeps_last = f->substeps_last;
    st->implicit_steps = f->implicit_steps;
    st->frame = f->frame;
    for (int k = 0; k < 4; k++) st->cg_report[k] = f->stats_pending ? o[4 + k] : f->cg_report[k];
    st->total_deposited = f->total_deposited;
    st->weave_volume = f->weave_volume;
    st->act_cells = f->cnt_act_cells;
    st->act_cell_substeps = f->cnt_act_cell_substeps;
    st->frames_counted = f->cnt_frames;
    st->act_cell_implicit = f->cnt_act_cell_implicit;
    st->one_launch_frames = f->cnt_one_frames;
    st->one_launch_cells = f->cnt_one_cells;
    st->one_launch_cell_substeps = f->cnt_one_cell_substeps;
    for (int k = 0; k < czi::kSubstepShapes; k++) st->flux_shape_cell_substeps[k] = f->cnt_shape_cell_substeps[k];
    return CZ_OK;
}

int cz_film_pool(cz_film* f, float* out, uint64_t n_floats) {
    if (!f || !out) return set_err(CZ_EINVAL, "null argument");
    uint64_t need = 2ull * f->pw * f->ph;
    if (n_floats < need) return set_err(CZ_ESIZE, "pool: need %llu floats", (unsigned long long)need);
    HIP_TRY(hipMemcpyAsync(out, f->pool_out, need * 4, hipMemcpyDeviceToHost, f->stream));
    return wait_idle(f);
}

// compute_stats (src/main.rs:1679-1819): host fold in FP64 of h and m.
int cz_film_mass_report(cz_film* f, cz_mass_report* r) {
    if (!f || !r) return set_err(CZ_EINVAL, "null argument");
    int rc = cz_mass_sums(f->h, f->m, f->n, f->mass_ws, cz_mass_sums_workspace(f->cap_n), f->mass_out, f->stream);
    if (rc) return rc;
    cz_mass_totals ms;
    HIP_TRY(hipMemcpyAsync(&ms, f->mass_out, sizeof(ms), hipMemcpyDeviceToHost, f->stream));
    rc = wait_idle(f);
    if (rc) return rc;
    // src/main.rs:1721-1726,1812-1818
    const double dx = (double)f->dx;
    const double vol = ms.sum_h * dx * dx, volm = ms.sum_m * dx * dx - f->weave_volume;
    r->wet_ml = vol * 1e6;
    r->cured_ml = volm * 1e6;
    r->deposited_ml = f->total_deposited * 1e6;
    r->err_pct = f->total_deposited > 0.0 ? (vol + volm - f->total_deposited) / f->total_deposited * 100.0 : 0.0;
    r->hmax = ms.h_max;
    r->pad = 0;
    r->wet_cells = ms.wet_cells;
    r->nan_cells = ms.nan_cells;
    return CZ_OK;
}

int cz_film_state_sizes(const cz_film* f, uint64_t sizes[CZ_NPARTS]) {
    if (!f || !sizes) return set_err(CZ_EINVAL, "null argument");
    for (int k = 0; k < CZ_NPARTS; k++) sizes[k] = buffer_of(f, k).bytes;
    return CZ_OK;
}

int cz_film_read_state(cz_film* f, void* const parts[CZ_NPARTS]) {
    if (!f || !parts) return set_err(CZ_EINVAL, "null argument");
    for (int k = 0; k < CZ_NPARTS; k++) {
        BufDesc b = buffer_of(f, k);
        if (!parts[k]) return set_err(CZ_EINVAL, "read_state: part %d is null", k);
        HIP_TRY(hipMemcpyAsync(parts[k], b.ptr, b.bytes, hipMemcpyDeviceToHost, f->stream));
    }
    return wait_idle(f);
}

// write_state (src/film/mod.rs:1303-1320)
int cz_film_write_state(cz_film* f, const void* const parts[CZ_NPARTS], const uint64_t sizes[CZ_NPARTS]) {
    if (f && f->batch) return set_err(CZ_EINVAL, "write_state: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f || !parts || !sizes) return set_err(CZ_EINVAL, "null argument");
    for (int k = 0; k < CZ_NPARTS; k++) {
        if (!parts[k] || sizes[k] != buffer_of(f, k).bytes)
            return set_err(CZ_ESIZE, "write_state: part %d has %llu bytes, expected %llu", k,
                           (unsigned long long)sizes[k], (unsigned long long)buffer_of(f, k).bytes);
    }
    int rc = wait_idle(f);
    if (rc) return rc;
    for (int k = 0; k < CZ_NPARTS; k++) {
        BufDesc b = buffer_of(f, k);
        HIP_TRY(hipMemcpyAsync(b.ptr, parts[k], b.bytes, hipMemcpyHostToDevice, f->stream));
    }
    return reset_solver_state(f);
}

int cz_film_buffer_size(const cz_film* f, int which, uint64_t* bytes) {
    if (!f || !bytes) return set_err(CZ_EINVAL, "null argument");
    BufDesc b = buffer_of(f, which);
    if (!b.ptr) return set_err(CZ_EINVAL, "unknown buffer %d", which);
    *bytes = b.bytes;
    return CZ_OK;
}

int cz_film_read_buffer(cz_film* f, int which, void* host, uint64_t bytes) {
    if (!f || !host) return set_err(CZ_EINVAL, "null argument");
    BufDesc b = buffer_of(f, which);
    if (!b.ptr) return set_err(CZ_EINVAL, "unknown buffer %d", which);
    if (bytes != b.bytes) return set_err(CZ_ESIZE, "buffer %d has %llu bytes", which, (unsigned long long)b.bytes);
    HIP_TRY(hipMemcpyAsync(host, b.ptr, bytes, hipMemcpyDeviceToHost, f->stream));
    return wait_idle(f);
}

int cz_film_write_buffer(cz_film* f, int which, const void* host, uint64_t bytes) {
    if (f && f->batch) return set_err(CZ_EINVAL, "write_buffer: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f || !host) return set_err(CZ_EINVAL, "null argument");
    BufDesc b = buffer_of(f, which);
    if (!b.ptr) return set_err(CZ_EINVAL, "unknown buffer %d", which);
    if (bytes != b.bytes) return set_err(CZ_ESIZE, "buffer %d has %llu bytes", which, (unsigned long long)b.bytes);
    int rc = wait_idle(f);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(b.ptr, host, bytes, hipMemcpyHostToDevice, f->stream));
    if (which == CZ_BUF_DELTA) HIP_TRY(hipMemcpyAsync(f->delta_other(), host, bytes, hipMemcpyHostToDevice, f->stream));
    return wait_idle(f);
}

int cz_film_device_ptr(cz_film* f, int which, void** ptr, uint64_t* bytes) {
    if (!f || !ptr) return set_err(CZ_EINVAL, "null argument");
    BufDesc b = buffer_of(f, which);
    if (!b.ptr) return set_err(CZ_EINVAL, "unknown buffer %d", which);
    *ptr = b.ptr;
    if (bytes) *bytes = b.bytes;
    return CZ_OK;
}

// dry_now (src/film/mod.rs:1230-1244): one cure pass at 4 / frame_dt.
int cz_film_dry_now(cz_film* f) {
    if (f && f->batch) return set_err(CZ_EINVAL, "dry_now: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f) return set_err(CZ_EINVAL, "null film");
    int rc = wait_idle(f);
    if (rc) return rc;
    cz_params saved = f->P;
    f->P.cure[0] = 4.0f / fmaxf(f->P.frame_dt, 1e-4f);
    drop_stale_error();
    hipLaunchKernelGGL(k_cure_clear<OneFilmAll>, dim3(f->nt16), dim3(256), 0, f->stream, OneFilmAll{f->dev(), f->nt16}, 0);
    HIP_TRY(hipGetLastError());
    f->P = saved;
    return wait_idle(f);
}

// clear (src/film/mod.rs:1009-1028)
int cz_film_clear(cz_film* f) {
    if (f && f->batch) return set_err(CZ_EINVAL, "clear: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f) return set_err(CZ_EINVAL, "null film");
    int rc = wait_idle(f);
    if (rc) return rc;
    uint64_t n = f->n;
    struct Z { void* p; uint64_t b; } zs[] = {
        {f->h, 4 * n}, {f->m, 4 * n}, {f->pres, 8 * n}, {f->flux, 16 * n}, {f->acc, 16 * n}, {f->alpha, 4 * n},
        {f->delta_a, 4 * n}, {f->delta_b, 4 * n}, {f->tmp, 4 * n}, {f->acc_s, 64 * n}, {f->strata_h, 16 * n},
        {f->strata_pig, 128 * n}, {f->strata_cure, 16 * n}, {f->strata_h2, 16 * n}, {f->strata_pig2, 128 * n},
        {f->alpha_s, 16 * n}, {f->tile_flags, 4 * (uint64_t)f->nt16}, {f->lay, 4 * n}};
    for (auto& z : zs) HIP_TRY(hipMemsetAsync(z.p, 0, z.b, f->stream));
    rc = init_state(f);
    if (rc) return rc;
    f->total_deposited = 0.0;
    f->h_max = 0.0f;
    f->s_max = 0.0f;
    f->h_max_pending = 0.0f;
    return CZ_OK;
}

void* cz_film_stream(cz_film* f) { return f ? (void*)f->stream : nullptr; }

int cz_film_shade_output(cz_film* f, int on) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    int rc = wait_idle(f);
    if (rc) return rc;
    if (on && !f->color) {
        rc = dalloc(f, &f->color, f->cap_n);
        if (!rc) rc = dalloc(f, &f->detail, f->cap_n);
        if (rc) return rc;
        for (auto& g : f->graphs) hipGraphExecDestroy(g.second);  // the descriptors change
        f->graphs.clear();
        f->graph_seen.clear();
    }
    f->shade_on = on != 0;
    return CZ_OK;
}

int cz_film_bin_on_device(cz_film* f, int mode) {
    if (!f || mode < -1 || mode > 1) return set_err(CZ_EINVAL, "bin_on_device: -1, 0 or 1");
    f->bin_mode = mode;
    return CZ_OK;
}

int cz_film_substep_launch(cz_film* f, int mode) {
    if (!f || mode < -1 || mode > 1) return set_err(CZ_EINVAL, "substep_launch: -1, 0 or 1");
    f->substep_mode = mode;
    return CZ_OK;
}

int cz_film_flux_shape(cz_film* f, int shape) {
    if (!f || shape < -1 || shape >= czi::kSubstepShapes) return set_err(CZ_EINVAL, "flux_shape: -1 (tuned) or 0-6");
    f->flux_tuner.fixed = shape;
    return CZ_OK;
}

int cz_film_graph_counts(const cz_film* f, uint64_t* captures, uint64_t* replays) {
    if (!f || !captures || !replays) return set_err(CZ_EINVAL, "graph_counts: null argument");
    *captures = f->graph_captures;
    *replays = f->graph_replays;
    return CZ_OK;
}

int cz_film_use_graphs(cz_film* f, int on) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    int rc = wait_idle(f);
    if (rc) return rc;
    f->use_graphs = on != 0;
    return CZ_OK;
}

int cz_film_profile(cz_film* f, int on) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    int rc = wait_idle(f);
    if (rc) return rc;
    f->prof_on = on != 0;
    f->prof_acc.clear();
    return CZ_OK;
}

int cz_film_profile_only(cz_film* f, const char* label) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    f->prof_only = label ? label : "";
    return CZ_OK;
}

int cz_film_reset_counters(cz_film* f) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    int rc = poll_stats(f);
    if (rc) return rc;
    f->cnt_act_cells = f->cnt_act_cell_substeps = f->cnt_frames = f->cnt_act_cell_implicit = 0;
    f->cnt_one_frames = f->cnt_one_cells = f->cnt_one_cell_substeps = 0;
    for (uint64_t& v : f->cnt_shape_cell_substeps) v = 0;
    return CZ_OK;
}

int cz_film_profile_read(cz_film* f, const char* label, double* avg_us, uint64_t* launches) {
    if (!f || !label) return set_err(CZ_EINVAL, "null argument");
    int rc = wait_idle(f);
    if (rc) return rc;
    auto it = f->prof_acc.find(label);
    if (it == f->prof_acc.end() || it->second.second == 0) {
        if (avg_us) *avg_us = 0.0;
        if (launches) *launches = 0;
        return CZ_OK;
    }
    if (avg_us) *avg_us = it->second.first / (double)it->second.second;
    if (launches) *launches = it->second.second;
    return CZ_OK;
}

}  // extern "C"
