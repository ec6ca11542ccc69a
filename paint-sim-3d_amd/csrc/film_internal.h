// film_internal.h - the film handle and the host helpers shared by the
// single-canvas entry points (film_host.hip) and the batched ones
// (batch_host.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "film_kernels.h"
#include "shape_tuner.h"

using namespace cz;

// Message of the last error on this thread (cz_last_error); returns code.
int cz_set_err(int code, const char* fmt, ...);
#define set_err cz_set_err

// hipGetLastError returns the last failed runtime call of the calling
// thread, whoever made it (the caller's own HIP or torch calls included).
// An entry point that checks its launches with it first drops what an
// earlier call left, so that the check reports its own launches only.
inline void drop_stale_error() { (void)hipGetLastError(); }

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) return set_err(CZ_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

struct GraphKeyStore {
    unsigned char bytes[14 * 4 + sizeof(cz_params)];
};

namespace czi {

constexpr uint32_t kMaxTileBlocks = 2048u;  // resident 256-thread blocks on 256 CUs
// one-launch substeps (k_substeps): one block per expected active tile (a
// power of two) up to kPersistBlocks.  The automatic choice takes it for
// active sets of at most kPersistTiles tiles; 0 = never: on gfx950 a grid
// barrier (agent-scope release, atomic arrival, acquire) costs about 9 us
// against 3.3 us for a dependent launch inside a hipGraph, so the three
// launches per substep are faster at every active-set size measured.
constexpr uint32_t kPersistTiles = 0u, kPersistBlocks = 256u;
// deposit binning on the device from this many deposits in a frame
constexpr uint32_t kDevBinMin = 1024u;
// Launch shapes of a substep, picked online per active-set size by the
// substep tuner (all give the same bits):
//   0 lap + plug, staged flux (window per tile), limiter + apply;
//   1 lap + plug, face-parallel flux (a block per tile and face), limiter + apply;
//   2 lap + plug, flux + limiter + apply in one launch;
//   3 as 0 with 1024 threads per tile (flux one face per thread);
//   4 as 2 with 1024 threads per tile;
//   5 the whole substep in one launch (k_substep_fused; single films);
//   6 lap + plug and limiter + apply with 1024 threads per tile around the
//     face-parallel flux of 1 (single films).
// 256-thread blocks hold 4 waves per tile; a few hundred active tiles leave
// the GPU at 1-2 waves per SIMD, where the wide shapes take each window phase
// in one pass with 16 waves per tile.  The face-parallel flux already spreads
// a tile over four blocks, so it beats the wide flux at every size; with the
// wide lap and apply beside it, shape 6 is the fastest below a few hundred
// active tiles and shape 1 above (2048^2 blob canvas: 19.9 against 23.2 us per
// substep for shape 4 at 100-130 tiles, 62.8 against 110 for shape 3 at 1500).
constexpr int kSubstepShapes = 7;
// Frames of more active tiles than this run direct, not as graph replays:
// their launches take long enough on the device that direct submissions keep
// ahead of it, so a replay saves nothing, while a capture costs 0.5-0.8 ms of
// host time (2048^2 blob canvas, 80-1500 active tiles: replayed and direct
// frames within 2 %; at 0-20 tiles a replay is 20 % faster)
constexpr uint64_t kGraphMaxTiles = 64;
constexpr int kBatchShapes = 5;  // a batch runs shapes 0-4

struct ProfSlot {
    std::string label;
    hipEvent_t a, b;
};

struct TunerSlot {
    ShapeTuner* t;
    int bucket, shape;
    hipEvent_t a, b;
    double items;  // work items of the timed launch (active tiles)
};

}  // namespace czi

struct cz_film {
    uint32_t W = 0, H = 0, n = 0, tw = 0, th = 0, nt16 = 0, t64x = 0, t64y = 0, pw = 0, ph = 0;
    float dx = 0.0f;
    cz_settings s;
    cz_overrides ov;
    bool ov_from_env = true;
    cz_params P;
    float h_max = 0.0f, mu_min = 1.0f, h_max_pending = 0.0f;
    float s_max = 0.0f;  // the surface maximum a painting file carries (src/film/mod.rs:567)
    uint32_t implicit_steps = 1, substeps_last = 1, implicit_last = 0, frame = 0, active_tiles = 0;
    uint32_t active_hint = 0xffffffffu;  // grid sizing: last folded active count
    // flux pass shape: 0 = one block per tile over an LDS window, 1 = one
    // block per (tile, face) reading through the caches
    czi::FilmTuner flux_tuner;
    std::vector<czi::TunerSlot> tuner_pending;
    bool capturing = false;
    // substep launches: -1 chosen by the active-set size, 0 three launches per
    // substep, 1 one launch for all substeps
    int substep_mode = -1;
    bool one_launch_now = false;  // set by record_frame for the frame it records
    bool sync_fault = false;  // a grid barrier of k_substeps timed out
    uint32_t persist_grid() const {
        const uint32_t act = std::min(active_hint, nt16);
        // the slump passes run between the substeps, outside the one launch
        if (s.slump || substep_mode == 0 || (substep_mode < 0 && (czi::kPersistTiles == 0u || act > czi::kPersistTiles)))
            return 0u;
        uint32_t p2 = 16u;  // power of two: few distinct grids, few distinct graphs
        while (p2 < act) p2 <<= 1;
        return std::max(1u, std::min(std::min(p2, czi::kPersistBlocks), nt16));
    }
    double total_deposited = 0.0, weave_volume = 0.0;
    float cg_report[4] = {0, 0, 0, 0};
    bool stats_pending = false;
    uint64_t cnt_act_cells = 0, cnt_act_cell_substeps = 0, cnt_frames = 0, cnt_act_cell_implicit = 0;
    uint64_t cnt_one_frames = 0, cnt_one_cells = 0, cnt_one_cell_substeps = 0;
    bool one_launch_last = false;  // the frame being folded ran its substeps as one launch
    uint64_t cnt_shape_cell_substeps[czi::kSubstepShapes] = {};
    uint32_t shape_subs_now[czi::kSubstepShapes] = {}, shape_subs_last[czi::kSubstepShapes] = {};  // substeps per shape
    int device = 0;
    hipStream_t stream = nullptr;
    void* batch = nullptr;  // the cz_batch this film is a member of
    uint64_t bytes = 0;     // device memory held
    uint32_t cap_n = 0, cap_nt16 = 0, cap_nt64 = 0, cap_np = 0;  // sizes the buffers were made for

    std::vector<void*> allocs;
    // device buffers
    float *h = nullptr, *m = nullptr;
    float2* rheo = nullptr;
    float4 *strata_h = nullptr, *strata_cure = nullptr, *strata_h2 = nullptr, *strata_cure2 = nullptr;
    uint4 *strata_pig = nullptr, *strata_pig2 = nullptr;
    uint32_t *sub_refl = nullptr, *lay = nullptr;
    float2 *pres = nullptr, *fnorm = nullptr;
    float4 *flux = nullptr, *acc = nullptr, *acc_s = nullptr, *alpha_s = nullptr, *cg_m = nullptr;
    float *alpha = nullptr, *tmp = nullptr, *h_frame = nullptr, *h_sub = nullptr;
    uint32_t *delta_a = nullptr, *delta_b = nullptr;
    bool delta_cur_a = true;
    float *cg_d = nullptr, *cg_x = nullptr, *cg_r = nullptr, *cg_p = nullptr, *cg_q = nullptr, *cg_t = nullptr,
          *cg_t2 = nullptr;
    double *cg_part = nullptr, *cg_part2 = nullptr;
    float* cg_s = nullptr;
    uint32_t *tile_flags = nullptr, *tile_list = nullptr, *tile_moved = nullptr, *active_count = nullptr;
    uint32_t *cure_list = nullptr, *cure_count = nullptr;
    uint32_t* grid_sync = nullptr;
    uint8_t* tile_bits = nullptr;
    uint32_t* compact_cnt = nullptr;
    int32_t* dep_rect = nullptr;   // [CZ_MAX_DEPOSITS][4] device binning: 64-tile rectangle per deposit
    uint32_t* dep_cnt = nullptr;   // [tiles64] device binning: counts (zero between frames)
    uint32_t* dep_n = nullptr;     // [1] device binning: the frame's deposit count
    uint32_t* dep_n_h = nullptr;   // pinned staging of it
    float* cmf = nullptr;          // [16][3] device copy of cmf_h
    uint2 *color = nullptr, *detail = nullptr;  // shade outputs, allocated when shading is turned on
    bool shade_on = false;         // run the shade pass at the end of every frame
    // undo journal (src/film/mod.rs:357-443): the ring and its per-tile arrays
    // exist from the first nonzero generation
    uint32_t j_gen = 0, j_cap = 0;
    uint64_t j_budget = 384ull << 20;
    uint32_t *j_meta = nullptr, *j_data = nullptr;   // [cap][4], [cap][53][256]
    uint32_t *j_mark = nullptr, *j_touch = nullptr, *j_best = nullptr;  // [cap_nt16] each
    uint8_t* j_bits = nullptr;                       // [cap_nt16] tiles saved this frame
    uint32_t* j_cnt = nullptr;                       // [blocks of 256 tiles] their counts
    uint2* j_list = nullptr;                         // [cap_nt16] (tile, slot) to save
    uint32_t* j_ctl = nullptr;                       // [4] head, records to save, generation, tiles marked
    uint32_t* j_ctl_h = nullptr;                     // pinned staging of j_ctl
    uint32_t* j_touch_h = nullptr;                   // pinned staging of the touches
    std::vector<uint32_t> j_touch_host;              // Journal.touch
    bool j_touched = false;
    // tool contacts and the knife (src/film/mod.rs:1030-1173), allocated at first use
    bool tools_ready = false;
    uint32_t contact_count = 0;                      // contacts for the next frame's brush pass
    cz_contact* contacts = nullptr;                  // [CZ_MAX_CONTACTS]
    uint2* ctile_range = nullptr;                    // [tiles64] first item, count
    uint32_t* ctile_items = nullptr;                 // [CZ_MAX_CONTACTS * 16]
    uint32_t* reservoir = nullptr;                   // [CZ_MAX_BRISTLES][17]
    void* knife_dev = nullptr;                       // the knife and its cell box (tools.hip)
    cz_scrap* scraps = nullptr;                      // [CZ_MAX_SCRAPS]
    uint32_t* kcnt = nullptr;                        // [knife blocks] scraps per block, then offsets
    uint32_t* kctl = nullptr;                        // [4] scrap count, frame base, frame total
    uint32_t kcap = 0;                               // knife blocks kcnt holds
    bool knife_active = false;
    uint32_t knife_grid[2] = {0, 0};                 // blocks of the knife passes (0 = nothing to do)
    cz_contact* contacts_h = nullptr;                // pinned staging
    uint2* crange_h = nullptr;
    uint32_t* citems_h = nullptr;
    void* knife_h = nullptr;
    int bin_mode = -1;             // deposit binning: -1 by count, 0 host, 1 device
    bool bin_on_device = false;    // the uploaded deposits are binned in the frame
    float2* stats_out = nullptr;
    float* smax_out = nullptr;
    float* n_cell = nullptr;         // [n] per-cell power-law index (settings.n_per_cell)
    float* lmu_cell = nullptr;       // [n] ln max(rheo.x, 1e-4), written with the rheology
    uint4* mass_ws = nullptr;        // [cz_mass_sums_workspace(cap_n) / 16] partials of the mass report
    cz_mass_totals* mass_out = nullptr;  // [1] its result
    float* pool_out = nullptr;
    float* tsum_a = nullptr;
    float* pigtab = nullptr;
    cz_deposit* deposits = nullptr;
    uint32_t *tile_start = nullptr, *tile_count = nullptr, *tile_items = nullptr;
    float* frame_out = nullptr;
    // host tables and pinned staging
    float pigtab_h[CZ_NPIG * CZ_PIG_STRIDE];
    float cmf_h[CZ_NBAND * 3];
    float* frame_out_h = nullptr;
    cz_deposit* dep_h = nullptr;
    uint32_t *tstart_h = nullptr, *tcount_h = nullptr, *titems_h = nullptr;
    uint32_t deposit_count = 0;
    // profiling
    bool use_graphs = true;
    std::vector<std::pair<struct GraphKeyStore, hipGraphExec_t>> graphs;
    std::vector<struct GraphKeyStore> graph_seen;  // shapes met once, not yet captured
    uint64_t graph_captures = 0, graph_replays = 0;
    bool prof_on = false;
    std::string prof_only;
    std::vector<czi::ProfSlot> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, std::pair<double, uint64_t>> prof_acc;

    uint32_t* delta_cur() const { return delta_cur_a ? delta_a : delta_b; }
    uint32_t* delta_other() const { return delta_cur_a ? delta_b : delta_a; }

    FilmDev dev() const {
        FilmDev F;
        memset(&F, 0, sizeof(F));
        F.P = P;
        F.W = W; F.H = H; F.n = n; F.tw = tw; F.th = th; F.nt16 = nt16; F.t64x = t64x; F.t64y = t64y;
        F.pw = pw; F.ph = ph;
        F.imp_steps = s.implicit_hydro ? implicit_steps : 0u;
        F.h = h; F.m = m; F.rheo = rheo; F.strata_h = strata_h; F.strata_pig = strata_pig;
        F.strata_cure = strata_cure; F.sub_refl = sub_refl; F.lay = lay;
        F.pres = pres; F.flux = flux; F.acc = acc; F.alpha = alpha; F.h_sub = h_sub;
        F.delta = delta_cur(); F.delta_out = delta_other();
        F.tmp = tmp; F.acc_s = acc_s; F.alpha_s = alpha_s; F.strata_h2 = strata_h2; F.strata_pig2 = strata_pig2;
        F.strata_cure2 = strata_cure2; F.h_frame = h_frame; F.fnorm = fnorm; F.cg_m = cg_m;
        F.cg_d = cg_d; F.cg_x = cg_x; F.cg_r = cg_r; F.cg_p = cg_p; F.cg_q = cg_q; F.cg_t = cg_t; F.cg_t2 = cg_t2;
        F.cg_part = cg_part; F.cg_part2 = cg_part2; F.cg_s = cg_s;
        F.tile_flags = tile_flags; F.tile_list = tile_list; F.tile_moved = tile_moved; F.active_count = active_count;
        F.cure_list = cure_list; F.cure_count = cure_count;
        F.n_cell = n_cell;
        F.lmu_cell = lmu_cell;
        F.stats_out = stats_out; F.smax_out = smax_out; F.pool_out = pool_out; F.tsum_a = tsum_a; F.grid_sync = grid_sync; F.tile_bits = tile_bits; F.compact_cnt = compact_cnt; F.dep_rect = dep_rect; F.dep_cnt = dep_cnt; F.dep_n = dep_n; F.cmf = cmf; F.color = color; F.detail = detail; F.pigtab = pigtab;
        F.deposits = deposits; F.tile_start = tile_start; F.tile_count = tile_count; F.tile_items = tile_items;
        F.frame_out = frame_out;
        derive_consts(F);
        return F;
    }

    // Grid of the tile kernels: the kernels walk the device-side list with a
    // grid-stride loop, so any size is correct; size it from the last known
    // active count (4 work items per tile for the face-parallel flux).
    uint32_t tile_grid(uint32_t per_tile = 1) const {
        uint64_t est = (uint64_t)std::max(std::min(active_hint, nt16), 1u) * per_tile;
        est = est + est / 4u + 16u;
        uint64_t p2 = 64;  // power of two: few distinct grids, few distinct graphs
        while (p2 < est) p2 <<= 1;
        est = p2;
        uint64_t cap = std::min<uint64_t>((uint64_t)nt16 * per_tile, (uint64_t)czi::kMaxTileBlocks * per_tile);
        return (uint32_t)std::max<uint64_t>(1u, std::min(est, cap));
    }

    hipEvent_t take_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};


namespace czi {
// update_params (src/film/mod.rs:869-917) on the film's host state.
void update_params(cz_film* f, float frame_dt, uint32_t substeps);
// Fold one frame's scalars (k_frame_stats layout) into the film's host state
// (poll_stats, src/film/mod.rs:935-973).
void fold_frame_out(cz_film* f, const float* o);
// Wait for the film's stream and fold its event timings.
int wait_idle(cz_film* f);
// The zeroing of write_state (src/film/mod.rs:1303-1320) after new state
// parts: solver buffers cleared, every tile marked moved; waits for it.
int reset_solver_state(cz_film* f);
// Host-side effects of an encoded frame of `substeps` substeps.
void frame_done(cz_film* f, uint32_t substeps, bool parity_done);
// Film::new on the current device; init = false leaves the state parts zero
// (for a caller that writes or seeds them next).
int film_create(uint32_t w, uint32_t h, const cz_settings* s, const float* pigtab, const float* cmf, bool init,
                cz_film** out);
// Re-dimension a film within the capacity it was created with, as a new
// Film with settings s (host state fresh; the caller writes the state).
int film_reshape(cz_film* f, uint32_t w, uint32_t h, const cz_settings* s);
// Everything a canvas carries from one frame to the next, on the host: the
// state parts, the solver buffers later frames read where they did not write
// them (both nose offset buffers, front normals, pressures, plug means, the
// frame heights, CG mobilities and directions, flux sums, tile marks, cell
// indices) and the host scalars that choose the next frame's substeps.  A
// film restored from it (after film_reshape to the same size and settings)
// steps on exactly as the film it was taken from.
struct FilmSnapshot {
    uint32_t W = 0, H = 0;
    cz_settings s;
    std::vector<uint8_t> data;
    float h_max = 0.0f, mu_min = 1.0f, h_max_pending = 0.0f, s_max = 0.0f;
    uint32_t implicit_steps = 1, substeps_last = 1, implicit_last = 0, frame = 0, active_tiles = 0, active_hint = 0;
    bool delta_cur_a = true;
    double total_deposited = 0.0, weave_volume = 0.0;
};
int film_snapshot(cz_film* f, FilmSnapshot& out);
uint64_t film_snapshot_bytes(cz_film* f);
int film_restore(cz_film* f, const FilmSnapshot& snap);
// Fold the tuner timings of finished frames (events already complete).
void fold_tuners(std::vector<TunerSlot>& pending, std::vector<hipEvent_t>& pool);
// Take a film out of its batch (before it is destroyed).
void batch_detach(cz_film* f);
// The journal pass at the start of a frame (journal.wgsl:79-98): three
// launches on the film's stream; only while the generation is nonzero.
void journal_record(cz_film* f);
// Journal::flush (src/film/mod.rs:405-412): the touches since the last frame
// to the device, before a frame is enqueued.
int journal_flush(cz_film* f);
// touch_rect (src/film/mod.rs:1246-1266) on the host touch array.
void journal_touch_rect(cz_film* f, float minx, float miny, float maxx, float maxy);
// A reshaped film starts a fresh journal (generation 0, empty ring).
int journal_reset(cz_film* f);
// Pinned staging the journal allocated (freed with the film).
void journal_free_host(cz_film* f);
// The brush pass (brush.wgsl) before the deposits and the knife passes
// (knife.wgsl) after them, on the film's stream.
void tools_record_brush(cz_film* f);
void tools_record_knife(cz_film* f);
// Pinned staging of the tool passes (freed with the film).
void tools_free_host(cz_film* f);
// A reshaped film drops its contacts, knife and unread outputs.
int tools_reset(cz_film* f);
}  // namespace czi
