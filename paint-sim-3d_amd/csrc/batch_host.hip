// batch_host.hip - many films stepped together (the C ABI cz_batch_* of
// include/canvaz_film.h).
//
// Each member keeps its own buffers, parameters and lag-1 statistics, and
// takes exactly the steps `cz_film_step` (encode_sim, src/main.rs:1875-1909,
// without deposits) would take on it alone.  What changes is the launch
// shape: a batched frame is one launch per kernel phase over the union of
// the members' tiles, so many small active sets fill the device together
// instead of one launch each.  The union lists are built on the device from
// each member's compacted tile list (k_batch_plan); the members still
// stepping at substep s (or implicit step i) form a prefix of the list,
// because the host orders the members by their substep (implicit step)
// counts for the frame.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "film_internal.h"
#include "host_ref.h"

using namespace czi;

struct cz_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t max_members = 0;
    std::vector<cz_film*> members;
    // device
    FilmDev* desc = nullptr;         // [max_members][CZ_VARIANTS]
    uint32_t* plan = nullptr;        // order | sub | iorder | imp, max_members each
    uint32_t* lims = nullptr;        // lim_sub[kMaxSub + 1] | lim_imp[kMaxImp] | n_all16 | n_all64 | n_cure
    uint32_t* imp_members = nullptr;
    uint32_t* imp_bases = nullptr;
    float* frame_out = nullptr;      // [max_members][16]
    uint32_t* items_sub = nullptr;
    uint32_t* items_imp = nullptr;
    uint32_t* items_cure = nullptr;
    uint32_t* all16 = nullptr;       // every 16-tile of every member
    uint32_t* all64 = nullptr;       // every 64-tile of every member
    double* part = nullptr;
    double* part2 = nullptr;
    uint64_t cap_tiles = 0, cap_t64 = 0;
    // host, pinned
    FilmDev* desc_h = nullptr;
    uint32_t* plan_h = nullptr;
    uint32_t* lims_h = nullptr;
    float* frame_out_h = nullptr;
    uint32_t n16 = 0, n64 = 0;       // totals of the static lists
    bool lists_dirty = true;
    bool stats_pending = false;
    uint64_t frames = 0;
    bool use_graphs = true;
    std::vector<std::pair<std::vector<uint32_t>, hipGraphExec_t>> graphs;
    std::vector<std::vector<uint32_t>> graph_seen;  // shapes met once, not yet captured
    uint64_t graph_captures = 0, graph_replays = 0;
    // flux pass shape per launch (0 = LDS window per tile, 1 = block per
    // (tile, face)), tuned online
    czi::ShapeTuner flux_tuner;
    std::vector<czi::TunerSlot> tuner_pending;
    std::vector<hipEvent_t> ev_pool;
};

namespace {

constexpr uint32_t kMaxSub = 64u;   // settings.max_substeps is clamped to this by update_params
constexpr uint32_t kMaxImp = 8u;
constexpr uint32_t kMaxMembers = 1024u;  // k_batch_plan's LDS table
constexpr uint32_t kGridCap = 2048u;

void free_dev(void* p) {
    if (p) hipFree(p);
}

void drop_graphs(cz_batch* b) {
    for (auto& g : b->graphs) hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

// Capacity of the per-tile arrays; captured graphs hold their addresses, so
// a reallocation drops them.
int grow(cz_batch* b, uint64_t tiles, uint64_t t64) {
    if (tiles > b->cap_tiles || t64 > b->cap_t64) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        drop_graphs(b);
    }
    if (tiles > b->cap_tiles) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        free_dev(b->items_sub); free_dev(b->items_imp); free_dev(b->items_cure); free_dev(b->all16);
        free_dev(b->part); free_dev(b->part2);
        b->items_sub = b->items_imp = b->items_cure = b->all16 = nullptr;
        b->part = b->part2 = nullptr;
        uint64_t cap = std::max<uint64_t>(tiles, b->cap_tiles * 2);
        b->cap_tiles = 0;
        HIP_TRY(hipMalloc(&b->items_sub, cap * 4));
        HIP_TRY(hipMalloc(&b->items_imp, cap * 4));
        HIP_TRY(hipMalloc(&b->items_cure, cap * 4));
        HIP_TRY(hipMalloc(&b->all16, cap * 4));
        HIP_TRY(hipMalloc(&b->part, cap * 8));
        HIP_TRY(hipMalloc(&b->part2, cap * 8));
        b->cap_tiles = cap;
    }
    if (t64 > b->cap_t64) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        free_dev(b->all64);
        b->all64 = nullptr;
        uint64_t cap = std::max<uint64_t>(t64, b->cap_t64 * 2);
        b->cap_t64 = 0;
        HIP_TRY(hipMalloc(&b->all64, cap * 4));
        b->cap_t64 = cap;
    }
    return CZ_OK;
}

// Rebuild the static all-tile lists after a membership change.
int rebuild_lists(cz_batch* b) {
    uint64_t t16 = 0, t64 = 0;
    for (cz_film* f : b->members) {
        t16 += f->nt16;
        t64 += (uint64_t)f->t64x * f->t64y;
    }
    int rc = grow(b, std::max<uint64_t>(t16, 1), std::max<uint64_t>(t64, 1));
    if (rc) return rc;
    std::vector<uint32_t> l16, l64;
    l16.reserve(t16);
    l64.reserve(t64);
    for (uint32_t m = 0; m < (uint32_t)b->members.size(); m++) {
        cz_film* f = b->members[m];
        for (uint32_t t = 0; t < f->nt16; t++) l16.push_back((m << CZ_ITEM_TILE_BITS) | t);
        for (uint32_t t = 0; t < f->t64x * f->t64y; t++) l64.push_back((m << CZ_ITEM_TILE_BITS) | t);
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (!l16.empty()) HIP_TRY(hipMemcpy(b->all16, l16.data(), l16.size() * 4, hipMemcpyHostToDevice));
    if (!l64.empty()) HIP_TRY(hipMemcpy(b->all64, l64.data(), l64.size() * 4, hipMemcpyHostToDevice));
    b->n16 = (uint32_t)t16;
    b->n64 = (uint32_t)t64;
    b->lims_h[kMaxSub + 1 + kMaxImp] = b->n16;
    b->lims_h[kMaxSub + 1 + kMaxImp + 1] = b->n64;
    HIP_TRY(hipMemcpy(b->lims + kMaxSub + 1 + kMaxImp, b->lims_h + kMaxSub + 1 + kMaxImp, 8, hipMemcpyHostToDevice));
    b->lists_dirty = false;
    return CZ_OK;
}

// Fold the last batched frame's scalars into every member (poll_stats).
int poll(cz_batch* b) {
    if (!b->stats_pending) return CZ_OK;
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (uint32_t m = 0; m < (uint32_t)b->members.size(); m++) fold_frame_out(b->members[m], b->frame_out_h + 16u * m);
    fold_tuners(b->tuner_pending, b->ev_pool);
    b->stats_pending = false;
    return CZ_OK;
}

hipEvent_t take_event(cz_batch* b) {
    if (!b->ev_pool.empty()) {
        hipEvent_t e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

uint32_t grid_for(uint64_t est, uint64_t total, uint32_t per) {
    est = est * per;
    est = est + est / 4u + 16u;
    uint64_t p2 = 64;
    while (p2 < est) p2 <<= 1;
    uint64_t cap = std::min<uint64_t>(total * per, (uint64_t)kGridCap * per);
    return (uint32_t)std::max<uint64_t>(1u, std::min(p2, cap));
}

// The launches of one batched frame (encode_frame, src/film/mod.rs:1322-1462,
// for every member at once).  key receives the launch shape.
struct FrameShape {
    uint32_t members = 0, max_sub = 0, max_imp = 0, iters = 0;
    std::vector<uint32_t> m_imp;   // members with more than i implicit steps
    std::vector<uint32_t> g_sub, g_flux, g_imp;   // g_flux: the flux shape of each substep
    std::vector<int> f_bucket;
    std::vector<uint8_t> f_timed;
    std::vector<double> f_items;  // active tiles of substep s (estimate), the tuner's normaliser
    uint32_t g_all = 0, g16 = 0, g64 = 0, g_cure = 0;
    std::vector<uint32_t> key() const {
        std::vector<uint32_t> k = {members, max_sub, max_imp, iters, g_all, g16, g64, g_cure};
        k.insert(k.end(), m_imp.begin(), m_imp.end());
        k.insert(k.end(), g_sub.begin(), g_sub.end());
        k.insert(k.end(), g_flux.begin(), g_flux.end());
        k.insert(k.end(), g_imp.begin(), g_imp.end());
        return k;
    }
};

int record(cz_batch* b, const FrameShape& sh) {
    drop_stale_error();
    const dim3 b256(256);
    hipStream_t st = b->stream;
    const uint32_t* lim_sub = b->lims;
    const uint32_t* lim_imp = b->lims + kMaxSub + 1;
    const uint32_t* n_all16 = b->lims + kMaxSub + 1 + kMaxImp;
    const uint32_t* n_all64 = n_all16 + 1;
    auto many = [&](const uint32_t* items, const uint32_t* cnt, uint32_t variant) {
        return ManyFilms{b->desc, items, cnt, variant};
    };
    const ManyCanvases all_c{b->desc, nullptr, nullptr, 0u};
    hipLaunchKernelGGL(k_mark<ManyFilms>, dim3(sh.g16), b256, 0, st, many(b->all16, n_all16, 0));
    hipLaunchKernelGGL(k_compact_tiles<ManyCanvases>, dim3(sh.members), dim3(1024), 0, st, all_c);
    BatchPlan P;
    P.films = b->desc;
    P.order = b->plan;
    P.sub = b->plan + b->max_members;
    P.iorder = b->plan + 2 * b->max_members;
    P.imp = b->plan + 3 * b->max_members;
    P.members = sh.members;
    P.max_sub = sh.max_sub;
    P.max_imp = sh.max_imp;
    P.items_sub = b->items_sub;
    P.items_imp = b->items_imp;
    P.lim_sub = b->lims;
    P.lim_imp = b->lims + kMaxSub + 1;
    P.imp_members = b->imp_members;
    P.imp_bases = b->imp_bases;
    P.items_cure = b->items_cure;
    P.n_cure = b->lims + kMaxSub + 1 + kMaxImp + 2;
    hipLaunchKernelGGL(k_batch_plan, dim3(sh.members), b256, 0, st, P);
    hipLaunchKernelGGL(k_cure<ManyFilms>, dim3(sh.g_cure), b256, 0, st, many(b->items_cure, P.n_cure, 0));
    for (uint32_t i = 0; i < sh.max_imp; i++) {
        const ManyFilms w = many(b->items_imp, lim_imp + i, 0);
        const ManyCanvases c{b->desc, b->imp_members, b->imp_bases, 0u};
        const dim3 g(sh.g_imp[i]), gc(sh.m_imp[i]);
        hipLaunchKernelGGL(k_lap_tail<ManyFilms>, g, b256, 0, st, w, 0, i == 0 ? 1 : 0);
        hipLaunchKernelGGL(k_cg_setup<ManyFilms>, g, b256, 0, st, w);
        hipLaunchKernelGGL(k_cg_lap<ManyFilms>, g, b256, 0, st, w, 1u);
        hipLaunchKernelGGL(k_cg_init<ManyFilms>, g, b256, 0, st, w);
        hipLaunchKernelGGL(k_cg_reduce<ManyCanvases>, gc, b256, 0, st, c, 0);
        for (uint32_t k = 0; k < sh.iters; k++) {
            hipLaunchKernelGGL(k_cg_lap<ManyFilms>, g, b256, 0, st, w, 0u);
            hipLaunchKernelGGL(k_cg_iter_a<ManyFilms>, g, b256, 0, st, w);
            hipLaunchKernelGGL(k_cg_reduce<ManyCanvases>, gc, b256, 0, st, c, 1);
            hipLaunchKernelGGL(k_cg_iter_b<ManyFilms>, g, b256, 0, st, w);
            hipLaunchKernelGGL(k_cg_reduce<ManyCanvases>, gc, b256, 0, st, c, 2);
            hipLaunchKernelGGL(k_cg_iter_c<ManyFilms>, g, b256, 0, st, w);
        }
        hipLaunchKernelGGL(k_cg_lap<ManyFilms>, g, b256, 0, st, w, 4u);
        hipLaunchKernelGGL(k_cg_finish<ManyFilms>, g, b256, 0, st, w);
        hipLaunchKernelGGL(k_cg_reduce<ManyCanvases>, gc, b256, 0, st, c, 3);
        hipLaunchKernelGGL(k_cg_rescale<ManyFilms>, g, b256, 0, st, w);
    }
    const dim3 b1024(1024);
    for (uint32_t s = 0; s < sh.max_sub; s++) {
        const uint32_t v_lap = s > 0 ? ((s - 1u) & 1u) : 0u, v = s & 1u;
        const dim3 g(sh.g_sub[s]);
        const uint32_t shape = sh.g_flux[s];
        hipEvent_t ea = nullptr, eb = nullptr;
        if (sh.f_timed[s]) {
            ea = take_event(b);
            eb = take_event(b);
            if (ea && eb) hipEventRecord(ea, st);
        }
        // the substep in the shape the tuner picked (it times the whole substep)
        const ManyFilms wl = many(b->items_sub, lim_sub + s, v_lap), w = many(b->items_sub, lim_sub + s, v);
        if (shape >= 3u)
            hipLaunchKernelGGL(k_lap_plug_w<ManyFilms>, g, b1024, 0, st, wl, s > 0 ? 1 : 0, 1, s == 0 ? 2 : 0);
        else
            hipLaunchKernelGGL(k_lap_plug<ManyFilms>, g, b256, 0, st, wl, s > 0 ? 1 : 0, 1, s == 0 ? 2 : 0);
        switch (shape) {
            case 1u:
                hipLaunchKernelGGL(k_fluxes_faces<ManyFilms>, dim3(std::min<uint32_t>(4u * sh.g_sub[s], 4u * kGridCap)),
                                   b256, 0, st, w);
                hipLaunchKernelGGL(k_limit_apply<ManyFilms>, g, b256, 0, st, w);
                break;
            case 2u:
                hipLaunchKernelGGL(k_flux_apply<ManyFilms>, g, b256, 0, st, w);
                break;
            case 3u:
                hipLaunchKernelGGL(k_fluxes_w<ManyFilms>, g, b1024, 0, st, w);
                hipLaunchKernelGGL(k_limit_apply_w<ManyFilms>, g, b1024, 0, st, w);
                break;
            case 4u:
                hipLaunchKernelGGL(k_flux_apply_w<ManyFilms>, g, b1024, 0, st, w);
                break;
            default:
                hipLaunchKernelGGL(k_fluxes<ManyFilms>, g, b256, 0, st, w);
                hipLaunchKernelGGL(k_limit_apply<ManyFilms>, g, b256, 0, st, w);
                break;
        }
        if (ea && eb) {
            hipEventRecord(eb, st);
            b->tuner_pending.push_back({&b->flux_tuner, sh.f_bucket[s], (int)shape, ea, eb, sh.f_items[s]});
        }
    }
    const uint32_t* all = lim_sub + sh.max_sub;
    const dim3 ga(sh.g_all);
    hipLaunchKernelGGL(k_lap_tail<ManyFilms>, ga, b256, 0, st, many(b->items_sub, all, 2), 1, 0);
    const ManyFilms w = many(b->items_sub, all, 3);
    hipLaunchKernelGGL(k_strata_split<ManyFilms>, dim3(std::min<uint32_t>(4u * sh.g_all, 4u * kGridCap)), b256, 0, st, w);
    hipLaunchKernelGGL(k_reduce_pool<ManyFilms>, dim3(sh.g64), b256, 0, st, many(b->all64, n_all64, 3));
    hipLaunchKernelGGL(k_strata_limit<ManyFilms>, ga, b256, 0, st, w);
    hipLaunchKernelGGL(k_strata_update<ManyFilms>, ga, b256, 0, st, w);
    hipLaunchKernelGGL(k_strata_fill<ManyFilms>, ga, b256, 0, st, w);
    hipLaunchKernelGGL(k_frame_stats<ManyCanvases>, dim3(sh.members), dim3(1024), 0, st,
                       ManyCanvases{b->desc, nullptr, nullptr, 3u});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(b->frame_out_h, b->frame_out, (size_t)sh.members * 16 * sizeof(float),
                           hipMemcpyDeviceToHost, st));
    return CZ_OK;
}

void destroy_batch(cz_batch* b) {
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    for (cz_film* f : b->members) f->batch = nullptr;
    for (void* p : {(void*)b->desc, (void*)b->plan, (void*)b->lims, (void*)b->imp_members, (void*)b->imp_bases,
                    (void*)b->frame_out, (void*)b->items_sub, (void*)b->items_imp, (void*)b->items_cure, (void*)b->all16, (void*)b->all64,
                    (void*)b->part, (void*)b->part2})
        free_dev(p);
    for (void* p : {(void*)b->desc_h, (void*)b->plan_h, (void*)b->lims_h, (void*)b->frame_out_h})
        if (p) hipHostFree(p);
    for (auto& g : b->graphs) hipGraphExecDestroy(g.second);
    for (auto& t : b->tuner_pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    for (auto e : b->ev_pool) hipEventDestroy(e);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace

namespace czi {
// A member film is being destroyed: take it out of its batch first.
void batch_detach(cz_film* f) {
    if (f && f->batch) cz_batch_remove((cz_batch*)f->batch, f);
}
}  // namespace czi

extern "C" {

int cz_batch_create(uint32_t max_members, cz_batch** out) {
    if (!out || max_members == 0 || max_members > kMaxMembers)
        return set_err(CZ_EINVAL, "cz_batch_create: max_members must be 1..%u", kMaxMembers);
    *out = nullptr;
    cz_batch* b = new cz_batch();
    b->max_members = max_members;
    // the 1024-thread shapes lose on the thousands of tiles a batch steps at
    // once (measured 1.6-2x slower), so its tuner keeps to shapes 0-2
    b->flux_tuner.ns = 3;
    int rc = CZ_OK;
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->desc, (size_t)max_members * CZ_VARIANTS * sizeof(FilmDev));
    if (e == hipSuccess) e = hipMalloc(&b->plan, (size_t)max_members * 4 * 4);
    if (e == hipSuccess) e = hipMalloc(&b->lims, (kMaxSub + 1 + kMaxImp + 3) * 4);
    if (e == hipSuccess) e = hipMalloc(&b->imp_members, (size_t)max_members * 4);
    if (e == hipSuccess) e = hipMalloc(&b->imp_bases, (size_t)max_members * 4);
    if (e == hipSuccess) e = hipMalloc(&b->frame_out, (size_t)max_members * 16 * 4);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->desc_h, (size_t)max_members * CZ_VARIANTS * sizeof(FilmDev), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->plan_h, (size_t)max_members * 4 * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->lims_h, (kMaxSub + 1 + kMaxImp + 3) * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->frame_out_h, (size_t)max_members * 16 * 4, hipHostMallocDefault);
    if (e != hipSuccess) rc = set_err(CZ_ENOMEM, "cz_batch_create: %s", hipGetErrorString(e));
    if (!rc) {
        memset(b->lims_h, 0, (kMaxSub + 1 + kMaxImp + 3) * 4);
        e = hipMemset(b->lims, 0, (kMaxSub + 1 + kMaxImp + 3) * 4);
        if (e != hipSuccess) rc = set_err(CZ_EHIP, "cz_batch_create: %s", hipGetErrorString(e));
    }
    if (rc) {
        destroy_batch(b);
        return rc;
    }
    *out = b;
    return CZ_OK;
}

void cz_batch_destroy(cz_batch* b) { destroy_batch(b); }

int cz_batch_add(cz_batch* b, cz_film* f) {
    if (!b || !f) return set_err(CZ_EINVAL, "null argument");
    if (f->batch) return set_err(CZ_EINVAL, "cz_batch_add: the film is already a batch member");
    if (f->j_gen != 0u) return set_err(CZ_EINVAL, "cz_batch_add: the film keeps an undo journal (generation 0 first)");
    if (f->knife_active) return set_err(CZ_EINVAL, "cz_batch_add: the film has a knife set (cz_film_set_knife(NULL) first)");
    if (f->s.slump) return set_err(CZ_EUNSUP, "cz_batch_add: the film runs the slump pass, which batches do not");
    if (b->members.size() >= b->max_members) return set_err(CZ_ENOMEM, "cz_batch_add: batch is full");
    if (f->device != b->device) return set_err(CZ_EINVAL, "cz_batch_add: film and batch are on different devices");
    if (f->nt16 > CZ_ITEM_TILE_MASK + 1u) return set_err(CZ_EINVAL, "cz_batch_add: canvas has too many tiles");
    if (!b->members.empty() && f->s.implicit_iters != b->members[0]->s.implicit_iters)
        return set_err(CZ_EINVAL, "cz_batch_add: members must share implicit_iters");
    int rc = poll(b);
    if (rc) return rc;
    // finish the film's own work and fold its pending frame scalars
    HIP_TRY(hipStreamSynchronize(f->stream));
    if (f->stats_pending) fold_frame_out(f, f->frame_out_h);
    rc = wait_idle(f);
    if (rc) return rc;
    f->batch = b;
    b->members.push_back(f);
    b->lists_dirty = true;
    return CZ_OK;
}

int cz_batch_remove(cz_batch* b, cz_film* f) {
    if (!b || !f) return set_err(CZ_EINVAL, "null argument");
    auto it = std::find(b->members.begin(), b->members.end(), f);
    if (it == b->members.end()) return set_err(CZ_EINVAL, "cz_batch_remove: not a member");
    int rc = poll(b);
    if (rc) return rc;
    b->members.erase(it);
    f->batch = nullptr;
    b->lists_dirty = true;
    return CZ_OK;
}

uint32_t cz_batch_size(const cz_batch* b) { return b ? (uint32_t)b->members.size() : 0u; }

int cz_batch_graph_counts(const cz_batch* b, uint64_t* captures, uint64_t* replays) {
    if (!b || !captures || !replays) return set_err(CZ_EINVAL, "graph_counts: null argument");
    *captures = b->graph_captures;
    *replays = b->graph_replays;
    return CZ_OK;
}

int cz_batch_use_graphs(cz_batch* b, int on) {
    if (!b) return set_err(CZ_EINVAL, "null batch");
    b->use_graphs = on != 0;
    return CZ_OK;
}

int cz_batch_flux_shape(cz_batch* b, int shape) {
    if (!b || shape < -1 || shape >= czi::kBatchShapes) return set_err(CZ_EINVAL, "flux_shape: -1 (tuned) or 0-4");
    b->flux_tuner.fixed = shape;
    return CZ_OK;
}

void* cz_batch_stream(cz_batch* b) { return b ? (void*)b->stream : nullptr; }

int cz_batch_step(cz_batch* b, float frame_dt) {
    if (!b || !(frame_dt > 0.0f)) return set_err(CZ_EINVAL, "cz_batch_step: bad arguments");
    const uint32_t nm = (uint32_t)b->members.size();
    if (nm == 0) return CZ_OK;
    int rc = poll(b);
    if (rc) return rc;
    if (b->lists_dirty) {
        rc = rebuild_lists(b);
        if (rc) return rc;
    }
    // per member: choose_substeps and update_params on its own lag-1 state
    std::vector<uint32_t> sub(nm), imp(nm);
    for (uint32_t m = 0; m < nm; m++) {
        cz_film* f = b->members[m];
        if (f->ov_from_env) cz_overrides_from_env(&f->ov);
        sub[m] = cz_film_choose_substeps(f, frame_dt);
        if (sub[m] > kMaxSub) return set_err(CZ_EINVAL, "cz_batch_step: %u substeps exceed %u", sub[m], kMaxSub);
        update_params(f, frame_dt, sub[m]);
        imp[m] = f->s.implicit_hydro ? f->implicit_steps : 0u;
        if (imp[m] > kMaxImp) return set_err(CZ_EINVAL, "cz_batch_step: %u implicit steps exceed %u", imp[m], kMaxImp);
    }
    std::vector<uint32_t> order(nm), iorder(nm);
    std::iota(order.begin(), order.end(), 0u);
    std::iota(iorder.begin(), iorder.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return sub[a] > sub[c]; });
    std::stable_sort(iorder.begin(), iorder.end(), [&](uint32_t a, uint32_t c) { return imp[a] > imp[c]; });
    FrameShape sh;
    sh.members = nm;
    sh.iters = b->members[0]->s.implicit_iters;
    uint64_t hint_all = 0;
    for (uint32_t k = 0; k < nm; k++) {
        b->plan_h[k] = order[k];
        b->plan_h[b->max_members + k] = sub[order[k]];
        b->plan_h[2 * b->max_members + k] = iorder[k];
        b->plan_h[3 * b->max_members + k] = imp[iorder[k]];
        sh.max_sub = std::max(sh.max_sub, sub[k]);
        sh.max_imp = std::max(sh.max_imp, imp[k]);
        hint_all += std::min(b->members[k]->active_hint, b->members[k]->nt16);
    }
    // per substep: active-set estimate and grid; a frame whose flux size
    // classes still need timing samples runs uncaptured
    std::vector<uint64_t> est_sub(sh.max_sub);
    bool learn = false;
    for (uint32_t s = 0; s < sh.max_sub; s++) {
        uint64_t est = 0, tot = 0;
        for (uint32_t m = 0; m < nm; m++)
            if (sub[m] > s) {
                est += std::min(b->members[m]->active_hint, b->members[m]->nt16);
                tot += b->members[m]->nt16;
            }
        sh.g_sub.push_back(grid_for(est, tot, 1));
        est_sub[s] = est;
        learn = learn || b->flux_tuner.learning(est);
    }
    const bool graph = b->use_graphs && !learn;
    for (uint32_t s = 0; s < sh.max_sub; s++) {
        int bk = 0;
        bool timed = false;
        const int shape = b->flux_tuner.choose(est_sub[s], !graph, bk, timed);
        sh.g_flux.push_back((uint32_t)shape);
        sh.f_bucket.push_back(bk);
        sh.f_timed.push_back(timed ? 1 : 0);
        sh.f_items.push_back((double)est_sub[s]);
    }
    for (uint32_t i = 0; i < sh.max_imp; i++) {
        uint64_t est = 0, tot = 0;
        uint32_t cnt = 0;
        for (uint32_t m = 0; m < nm; m++)
            if (imp[m] > i) {
                est += std::min(b->members[m]->active_hint, b->members[m]->nt16);
                tot += b->members[m]->nt16;
                cnt++;
            }
        sh.g_imp.push_back(grid_for(est, tot, 1));
        sh.m_imp.push_back(cnt);
    }
    sh.g_all = grid_for(hint_all, b->n16, 1);
    sh.g16 = std::min<uint32_t>((b->n16 + 3u) / 4u, 8192u);
    sh.g64 = std::min<uint32_t>(b->n64, 8192u);
    sh.g_cure = std::min<uint32_t>(b->n16, kGridCap);
    // descriptors: each member in its four offset-buffer parities
    for (uint32_t m = 0; m < nm; m++) {
        cz_film* f = b->members[m];
        const bool d0 = f->delta_cur_a;
        const uint32_t flips[CZ_VARIANTS] = {0u, 1u, (sub[m] - 1u) & 1u, sub[m] & 1u};
        for (uint32_t v = 0; v < CZ_VARIANTS; v++) {
            f->delta_cur_a = (flips[v] & 1u) ? !d0 : d0;
            FilmDev F = f->dev();
            F.cg_part = b->part;
            F.cg_part2 = b->part2;
            F.frame_out = b->frame_out + 16u * m;
            b->desc_h[m * CZ_VARIANTS + v] = F;
        }
        f->delta_cur_a = d0;
    }
    HIP_TRY(hipMemcpyAsync(b->desc, b->desc_h, (size_t)nm * CZ_VARIANTS * sizeof(FilmDev), hipMemcpyHostToDevice,
                           b->stream));
    HIP_TRY(hipMemcpyAsync(b->plan, b->plan_h, (size_t)b->max_members * 4 * 4, hipMemcpyHostToDevice, b->stream));
    if (graph) {
        std::vector<uint32_t> key = sh.key();
        hipGraphExec_t exec = nullptr;
        for (auto& g : b->graphs)
            if (g.first == key) exec = g.second;
        bool seen = exec != nullptr;
        for (const auto& o : b->graph_seen) seen = seen || o == key;
        if (!seen) {
            // a shape is captured the second time it comes: one-off shapes
            // (membership or substep changes) run direct
            if (b->graph_seen.size() >= 64) b->graph_seen.erase(b->graph_seen.begin());
            b->graph_seen.push_back(key);
            rc = record(b, sh);
            if (rc) return rc;
        } else if (!exec) {
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = record(b, sh);
            hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (rc) return rc;
            if (e != hipSuccess) return set_err(CZ_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (e != hipSuccess) return set_err(CZ_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
            if (b->graphs.size() >= 32) {
                hipGraphExecDestroy(b->graphs.front().second);
                b->graphs.erase(b->graphs.begin());
            }
            b->graphs.push_back({key, exec});
            b->graph_captures++;
        }
        if (exec) {
            HIP_TRY(hipGraphLaunch(exec, b->stream));
            b->graph_replays++;
        }
    } else {
        rc = record(b, sh);
        if (rc) return rc;
    }
    for (uint32_t m = 0; m < nm; m++) {
        b->members[m]->one_launch_now = false;
        for (uint32_t& v : b->members[m]->shape_subs_now) v = 0u;
        frame_done(b->members[m], sub[m], false);
    }
    b->stats_pending = true;
    b->frames++;
    return CZ_OK;
}

int cz_batch_sync(cz_batch* b) {
    if (!b) return set_err(CZ_EINVAL, "null batch");
    return poll(b);
}

}  // extern "C"
