/*
 * canvaz_film.h - C ABI of the MI355X (gfx950) thin-film paint step.
 *
 * This is the drop-in boundary that replaces the wgpu dispatch of the
 * reference's `Film` (src/film/mod.rs:445-1462 over src/gpu/mod.rs) on the
 * headless `--auto` path.  Every entry point names the reference method it
 * stands for.  Rules that hold for the whole ABI:
 *   - status returns are int: 0 = ok, negative = error (CZ_E*); the message of
 *     the last error on the calling thread is cz_last_error().
 *   - no exceptions, panics or callbacks cross the ABI.
 *   - a cz_film owns all of its device memory; host arrays are borrowed for the
 *     duration of the call only.
 *   - one cz_film = one HIP stream; calls on one handle are serialised by the
 *     caller, distinct handles may be driven from different threads.
 *   - state crossing the ABI is row-major, idx = j*W + i
 *     (shaders/common.wgsl:140-142), in the reference's per-cell layout.
 */
#ifndef CANVAZ_FILM_H
#define CANVAZ_FILM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CZ_NPIG 12          /* shaders/common.wgsl:47 */
#define CZ_NBAND 16         /* shaders/common.wgsl:48 */
#define CZ_PIG_STRIDE 40    /* shaders/common.wgsl:49 */
#define CZ_STRATA 4         /* shaders/common.wgsl:46 */
#define CZ_TILE64 64        /* shaders/common.wgsl:50 */
#define CZ_MAX_DEPOSITS 65536               /* src/film/mod.rs:584 */
#define CZ_MAX_TILE_ITEMS (65536 * 16)      /* src/film/mod.rs:661,1208 */

#define CZ_OK 0
#define CZ_EINVAL (-1)      /* bad argument */
#define CZ_ESIZE (-2)       /* size mismatch (write_state returns false) */
#define CZ_EHIP (-3)        /* HIP runtime error */
#define CZ_ENOMEM (-4)      /* device allocation failed / over budget */
#define CZ_EUNSUP (-5)      /* feature outside the built path */

/* Uniform parameter block, byte-identical to `Params`
 * (src/film/mod.rs:5-23, shaders/common.wgsl:1-17): 176 bytes. */
typedef struct cz_params {
    uint32_t size[2];
    uint32_t tiles[2];
    float dx;
    float dt;
    float inv_dx;
    float frame_dt;
    float g[4];
    float phys[4];
    float ein[4];
    float cure[4];
    float misc[4];
    uint32_t flags[4];
    float front[4];
    float front2[4];
    float front3[4];
} cz_params;

/* One deposit segment, byte-identical to `Deposit`
 * (src/film/mod.rs:25-39, shaders/common.wgsl:19-34): 96 bytes. */
typedef struct cz_deposit {
    float p0[2];
    float p1[2];
    float n0[2];
    float n1[2];
    float r0;
    float r1;
    float h0;
    float vol;
    uint32_t kind;
    uint32_t pad[3];
    uint32_t pig[8];
} cz_deposit;

#define CZ_DEPOSIT_RECYCLED 1u   /* src/film/mod.rs:41 */
#define CZ_DEPOSIT_CAP_START 2u  /* src/film/mod.rs:42 */
#define CZ_DEPOSIT_CAP_END 4u    /* src/film/mod.rs:43 */
#define CZ_DEPOSIT_METHOD_SHIFT 8u
enum { CZ_METHOD_THREAD = 1, CZ_METHOD_DROP = 2, CZ_METHOD_POUR = 3,
       CZ_METHOD_BRUSH = 4, CZ_METHOD_KNIFE = 5, CZ_METHOD_ROLLER = 6 };

#define CZ_MAX_CONTACTS 65536     /* src/film/mod.rs:235 */
#define CZ_MAX_SCRAPS (1u << 18)  /* src/film/mod.rs:236 */
#define CZ_MAX_BRISTLES 512       /* src/film/mod.rs:237 */

/* One tool contact segment, byte-identical to `Contact`
 * (src/film/mod.rs:239-248, shaders/brush.wgsl:1-8): 32 bytes.  pad is the
 * contact kind: 0 bristle, 1 roller, 2 sponge, 3 body. */
typedef struct cz_contact {
    float p0[2];
    float p1[2];
    float r;
    float zb;
    uint32_t bristle;
    uint32_t pad;
} cz_contact;
enum { CZ_CONTACT_BRISTLE = 0, CZ_CONTACT_ROLLER = 1, CZ_CONTACT_SPONGE = 2, CZ_CONTACT_BODY = 3 };

/* The palette knife of one frame, byte-identical to `KnifeParams`
 * (src/film/mod.rs:250-270, shaders/knife.wgsl:6-23): 384 bytes. */
typedef struct cz_knife {
    float a[2];
    float b[2];
    float dir[2];
    float sweep;
    float height;
    float half_width;
    uint32_t active;
    uint32_t frame;
    uint32_t pad1;
    uint32_t bands;
    uint32_t pad2, pad3, pad4;
    uint32_t pig[8][8];   /* per band: 16 f16 pigment fractions */
    float fill[8];
    float smear[8];
} cz_knife;

/* One knife scrap, byte-identical to `Scrap` (src/film/mod.rs:272-279,
 * shaders/knife.wgsl:25-32): 48 bytes.  vol > 0: paint taken off the cell;
 * vol < 0: paint laid from the knife's load. */
typedef struct cz_scrap {
    uint32_t cell;
    float vol;
    uint32_t frame;
    uint32_t pad1;
    uint32_t pig[8];
} cz_scrap;

/* `FilmSettings` (src/film/mod.rs:281-313); bools as uint32 0/1.
 * cz_settings_default() fills the reference defaults (mod.rs:315-350). */
typedef struct cz_settings {
    float width_m;
    float height_m;
    float rho;
    float sigma;
    float n_index;
    float m_cap;
    float contact_angle_deg;
    float hysteresis_deg;
    float h_star;
    uint32_t raw_canvas;
    float weave_um;
    float cure_rate;
    float tau_cure;
    float lnmu_cure;
    float v_cap;
    float tilt_deg;
    float tilt_dir_deg;
    float gravity;
    float medium_mu;
    uint32_t debug_view;
    float h_min;
    float m_cap_hydro;
    uint32_t max_substeps;
    uint32_t min_substeps;
    uint32_t implicit_iters;
    uint32_t implicit_hydro;
    /* 1 = the slump pass after each substep's nose (shed, slump, commit_tmp;
     * film.wgsl:563-801, 866-876), off by default as in the reference
     * (src/film/mod.rs:345).  A film with it on runs alone: cz_batch_add
     * refuses it with CZ_EUNSUP. */
    uint32_t slump;
    uint32_t capillary;
    uint32_t capillary_implicit;
    /* Not in the reference (whose power-law index is the film's n_index,
     * film.wgsl:95): 1 = per-cell power-law index.  Each stratum's pigment
     * slot 15 (unused by the reference) holds its index in f16, 0 standing
     * for n_index; a cell's index is the thickness-weighted mean over its
     * strata (formed with the rheology by the cure pass), a face's the
     * height-weighted mean of its two cells' (as for the yield stress).
     * update_params sets Params.flags[2] bit 0 (CZ_FLAG_CELL_N) from it.
     * 0 (default) = the reference's global index. */
    uint32_t n_per_cell;
    /* `canvas_rgb: Option<[f32; 3]>` (src/film/mod.rs:293): with
     * has_canvas_rgb = 1 the substrate reflects canvas_rgb, the colour the
     * canvas shows (sRGB, 0..1), through color::canvas_albedo and
     * reflectance_of (src/color/mod.rs:29-100); 0 (default) = the raw or the
     * white canvas by raw_canvas. */
    uint32_t has_canvas_rgb;
    float canvas_rgb[3];
} cz_settings;

#define CZ_FLAG_CELL_N 1u   /* Params.flags[2]: per-cell power-law index (n_per_cell) */

/* The environment toggles read by `update_params` every frame
 * (src/film/mod.rs:870-915).  cz_overrides_from_env() reads PAINT_NO_IMPLICIT,
 * PAINT_NO_CAP, PAINT_CAPCOND, PAINT_FCAP, PAINT_NO_HOLD, PAINT_KSHAPE,
 * PAINT_NO_TRANSL, PAINT_NO_SHAPE, PAINT_KLINE, PAINT_HYST. */
typedef struct cz_overrides {
    uint32_t no_implicit;
    uint32_t no_cap;
    float capcond;      /* default 80 */
    float fcap;         /* default 2.5 */
    uint32_t no_hold;
    float kshape;       /* default 1 */
    uint32_t no_transl;
    uint32_t no_shape;
    float kline;        /* default 0 */
    uint32_t has_hyst;
    float hyst_deg;
} cz_overrides;

/* Frame statistics: the `poll_stats` readback (src/film/mod.rs:935-973) plus
 * the per-frame solver counts.  h_max / mu_min are the CFL inputs of the next
 * frame (deterministic lag of one frame). */
typedef struct cz_stats {
    float h_max;
    float mu_min;
    uint32_t active_tiles;
    uint32_t substeps_last;
    uint32_t implicit_steps;
    uint32_t frame;
    float cg_report[4];      /* cg_s[1]: (rz0, drift, total, rz_final) */
    double total_deposited;  /* src/film/mod.rs:1188 */
    double weave_volume;     /* src/film/mod.rs:689 */
    /* running sums over the frames folded so far (since create or
     * cz_film_reset_counters): sum of active cells, of active cells times
     * that frame's substeps, and of active cells times its implicit steps */
    uint64_t act_cells;
    uint64_t act_cell_substeps;
    uint64_t frames_counted;
    uint64_t act_cell_implicit;
    /* the same sums over the frames whose substeps ran as one launch */
    uint64_t one_launch_frames;
    uint64_t one_launch_cells;
    uint64_t one_launch_cell_substeps;
    /* active cells times substeps run with each substep shape (0 window
     * flux, 1 face-parallel flux, 2 flux + apply in one launch, 3 and 4 as 0
     * and 2 with 1024 threads per tile, 5 the whole substep in one launch,
     * 6 the flux of 1 between the lap/plug and apply of 3; see
     * cz_film_flux_shape) */
    uint64_t flux_shape_cell_substeps[7];
} cz_stats;

/* The mass report of `compute_stats` (src/main.rs:1679-1819), sums in FP64. */
typedef struct cz_mass_report {
    double wet_ml;
    double cured_ml;
    double deposited_ml;
    double err_pct;
    float hmax;
    uint32_t pad;
    uint64_t wet_cells;
    uint64_t nan_cells;
} cz_mass_report;

/* The FP64 sums behind the mass report (compute_stats, src/main.rs:1708-1720):
 * over the cells whose height is finite, the sums of h and m, the wet count
 * and the maxima of h and of h + m (film.s_max); the cells with a
 * non-finite height are counted and skipped.  48 bytes. */
typedef struct cz_mass_totals {
    double sum_h;
    double sum_m;
    uint64_t wet_cells;
    uint64_t nan_cells;
    float h_max;
    float s_max;
    uint32_t pad[2];
} cz_mass_totals;

/* Indices of the eight state parts of read_state/write_state, in the order of
 * `state_buffers` (src/film/mod.rs:1277-1279). */
enum {
    CZ_PART_H = 0,          /* f32 [n]            */
    CZ_PART_M = 1,          /* f32 [n]            */
    CZ_PART_RHEO = 2,       /* f32 [n][2] (mu, tau_y) */
    CZ_PART_STRATA_H = 3,   /* f32 [n][4]         */
    CZ_PART_STRATA_PIG = 4, /* u32 [n][4][8] (16 f16 per stratum) */
    CZ_PART_STRATA_CURE = 5,/* f32 [n][4]         */
    CZ_PART_SUB_REFL = 6,   /* u32 [n][8] (16 f16 bands) */
    CZ_PART_LAY = 7,        /* u32 [n]            */
    CZ_NPARTS = 8
};

/* Solver buffers that persist across frames and that the step reads at
 * cells outside the active tiles; exposed for exact checkpointing and tests. */
enum {
    CZ_BUF_DELTA = 16,      /* f32 bits [n]: nose offset, shaders/film.wgsl:570-580 */
    CZ_BUF_FNORM = 17,      /* f32 [n][2] */
    CZ_BUF_PRES = 18,       /* f32 [n][2] */
    CZ_BUF_H_FRAME = 19,    /* f32 [n] */
    CZ_BUF_CG_M = 20,       /* f32 [n][4] */
    CZ_BUF_CG_P = 21,       /* f32 [n] */
    CZ_BUF_TILE_MOVED = 22, /* u32 [tiles16] */
    CZ_BUF_TILE_FLAGS = 23, /* u32 [tiles16] */
    CZ_BUF_TMP = 24,        /* f32 [n] */
    CZ_BUF_FLUX = 25,       /* f32 [n][4] */
    CZ_BUF_ACC = 26,        /* f32 [n][4] */
    CZ_BUF_ALPHA = 27,      /* f32 [n] */
    CZ_BUF_STATS = 28,      /* f32 [tiles64][2] */
    CZ_BUF_POOL = 29,       /* f32 [2][pw*ph] */
    CZ_BUF_TILE_LIST = 30,  /* u32 [tiles16] (ascending) */
    CZ_BUF_ACC_S = 31,      /* f32 [n][4][4] */
    CZ_BUF_ALPHA_S = 32,    /* f32 [n][4] */
    CZ_BUF_DEP_START = 33,  /* u32 [tiles64]: first item of each 64-tile (upload_deposits) */
    CZ_BUF_DEP_COUNT = 34,  /* u32 [tiles64]: its item count */
    CZ_BUF_DEP_ITEMS = 35,  /* u32 [CZ_MAX_TILE_ITEMS]: deposit indices, per tile ascending */
    CZ_BUF_COLOR = 36,      /* f16 [n][4]: shade color_out (linear rgb, roughness); with cz_film_shade_output */
    CZ_BUF_DETAIL = 37,     /* f16 [n][4]: shade detail_out (method + turn, cure, yield stress, thickness) */
    CZ_BUF_JMETA = 38,      /* u32 [journal cap][4]: slot (tile16, generation, sequence, valid) */
    CZ_BUF_JMARK = 39,      /* u32 [tiles16]: generation each tile was last saved in */
    CZ_BUF_N_CELL = 41,     /* f32 [n]: per-cell power-law index (settings.n_per_cell; derived by the cure pass) */
    CZ_BUF_JDATA = 40       /* u32 [journal cap][13568]: per slot the parts h, m, rheo, strata_h, strata_pig,
                               strata_cure, sub_refl, lay, each as [256 cells (tile rows)][its words] */
};

typedef struct cz_film cz_film;

/* ---- host-side helpers of the reference (no device needed) ------------- */

/* FilmSettings::default (src/film/mod.rs:315-350). */
void cz_settings_default(cz_settings* s);
/* Defaults of the environment toggles, then the values found in the env. */
void cz_overrides_default(cz_overrides* o);
void cz_overrides_from_env(cz_overrides* o);
/* pigment_table (src/color/mod.rs:243-258): refl[12][16], scatter/tint/
 * viscosity/yield[12] -> out[12*40]. */
void cz_pigment_table(const float* refl, const float* scatter, const float* tint,
                      const float* viscosity, const float* yield_stress, float* out);
/* default_pigments (src/color/mod.rs:190-224) as the five arrays above. */
void cz_default_pigments(float* refl, float* scatter, float* tint, float* viscosity,
                         float* yield_stress);
/* cmf_table (src/color/mod.rs:150-171) -> out[16*3]. */
void cz_cmf_table(float* out);
/* weave (src/film/mod.rs:212-233) -> out[w*h]. */
void cz_weave(uint32_t w, uint32_t h, float dx, float amp, float* out);
/* substrate_refl (src/film/mod.rs:188-210) without a canvas colour -> out[n*8]. */
void cz_substrate_refl(uint64_t n, uint32_t raw, uint32_t* out);
/* color::reflectance_of(color::canvas_albedo(rgb), cmf) (src/color/mod.rs:
 * 29-100): the 16-band reflectance of a canvas shown in sRGB colour rgb,
 * cmf = the cz_cmf_table layout.  Powers and exponentials are evaluated in
 * double and rounded once. */
void cz_canvas_reflectance(const float rgb[3], const float* cmf, float out[16]);
/* The substrate of one cell for settings s (the raw or white canvas, or
 * s->canvas_rgb's reflectance when s->has_canvas_rgb): 8 words of two f16
 * bands each, as substrate_refl packs them. */
void cz_substrate_words(const cz_settings* s, const float* cmf, uint32_t out[8]);
/* Deposit::ribbon / capsule / stroke / laid_by (src/film/mod.rs:62-98). */
void cz_deposit_ribbon(const float p0[2], const float p1[2], const float n0[2], const float n1[2],
                       float r0, float r1, uint32_t caps, float volume, const uint32_t pig[8],
                       float dx, cz_deposit* out);
void cz_deposit_capsule(const float p0[2], const float p1[2], float r, float volume,
                        const uint32_t pig[8], float dx, cz_deposit* out);
void cz_deposit_stroke(const float p0[2], const float p1[2], float r, float volume,
                       const uint32_t pig[8], float dx, cz_deposit* out);
void cz_deposit_laid_by(cz_deposit* d, uint32_t method);
/* Deposit::sampled_volume / height_at / inside_fraction (mod.rs:113-150). */
float cz_deposit_sampled_volume(const cz_deposit* d, float lattice);
float cz_deposit_height_at(const cz_deposit* d, float x, float y);
float cz_deposit_inside_fraction(const cz_deposit* d, float w, float h);
/* PaintMix::normalized().packed() (src/color/mod.rs:348-371) of frac[16]. */
void cz_paint_pack(const float* frac16, int normalize, uint32_t out[8]);
/* The lagged per-frame stats the host side of the step keeps
 * (Film.h_max / mu_min / h_max_pending, src/film/mod.rs:855). */
typedef struct cz_host_state {
    float h_max;
    float mu_min;
    float h_max_pending;
    uint32_t pad;
} cz_host_state;
/* gravity_vec, update_params, choose_substeps (src/film/mod.rs:861-933) as
 * pure host functions of (settings, toggles, grid, stats). */
void cz_gravity_vec_host(const cz_settings* s, float out[3]);
void cz_update_params_host(const cz_settings* s, const cz_overrides* o, uint32_t w, uint32_t h, float frame_dt,
                           uint32_t substeps, const cz_host_state* hs, cz_params* out, uint32_t* implicit_steps);
uint32_t cz_choose_substeps_host(const cz_settings* s, uint32_t w, float frame_dt, const cz_host_state* hs);

/* ---- the film (Film, src/film/mod.rs:445-1462) ------------------------- */

/* Film::new (mod.rs:579-859) on the current HIP device.  pigtab: 12*40 f32,
 * cmf: 16*3 f32 (NULL -> reference defaults).  Initial state: h = 0,
 * m = weave, rheo = (1,0), sub_refl = raw canvas, params from
 * update_params(1/60, 1). */
int cz_film_create(uint32_t w, uint32_t h, const cz_settings* s, const float* pigtab,
                   const float* cmf, cz_film** out);
void cz_film_destroy(cz_film* f);
int cz_film_dims(const cz_film* f, uint32_t* w, uint32_t* h, float* dx);
/* Replace the settings used by the next update_params (FilmSettings is pub). */
int cz_film_set_settings(cz_film* f, const cz_settings* s);
int cz_film_get_settings(const cz_film* f, cz_settings* s);
/* Environment toggles used by update_params (default: read from the env at
 * create time and at every cz_film_step). */
int cz_film_set_overrides(cz_film* f, const cz_overrides* o);
/* gravity_vec (mod.rs:861-867). */
int cz_film_gravity_vec(const cz_film* f, float out[3]);
/* update_params (mod.rs:869-917). */
int cz_film_update_params(cz_film* f, float frame_dt, uint32_t substeps);
/* choose_substeps (mod.rs:919-933). */
uint32_t cz_film_choose_substeps(const cz_film* f, float frame_dt);
int cz_film_get_params(const cz_film* f, cz_params* p);
int cz_film_set_params(cz_film* f, const cz_params* p);
/* upload_deposits (mod.rs:1175-1224): culls, accounts total_deposited, bins
 * into 64-tiles; *sent = deposits taken (the rest are held by the caller). */
int cz_film_upload_deposits(cz_film* f, const cz_deposit* deps, uint32_t n, uint32_t* sent);
/* encode_frame (mod.rs:1322-1462) enqueued on the film's stream with the
 * current params; no host sync.  The deposits of the last upload are
 * consumed by this frame. */
int cz_film_encode_frame(cz_film* f, uint32_t substeps);
/* encode_sim (src/main.rs:1875-1909): poll the previous frame's stats (lag 1,
 * deterministic), upload deposits, choose_substeps, update_params,
 * encode_frame, request stats.  *sent as in upload_deposits (may be NULL). */
int cz_film_step(cz_film* f, const cz_deposit* deps, uint32_t n, float frame_dt, uint32_t* sent);
/* Wait for the film's stream and fold the last frame's stats (poll_stats). */
int cz_film_sync(cz_film* f);
int cz_film_stats(cz_film* f, cz_stats* st);
/* pool8 output (mod.rs:956-963): pw*ph surface max then pw*ph relief max. */
int cz_film_pool(cz_film* f, float* out, uint64_t n_floats);
/* compute_stats mass report (src/main.rs:1679-1819); the sums run on the
 * device in FP64 (cz_mass_sums over the film's h and m). */
int cz_film_mass_report(cz_film* f, cz_mass_report* r);

/* The summing loop of compute_stats (src/main.rs:1708-1720), which the
 * reference runs on the host after reading h and m back, as two device
 * launches over device arrays h, m of n cells: per 65536-cell chunk a
 * partial record into `workspace`, then a fixed-order fold into *out
 * (device memory, 16-byte aligned), enqueued on `stream` (hipStream_t,
 * NULL = default) of the current device.  The bits of the result depend
 * only on the inputs.  workspace: at least cz_mass_sums_workspace(n)
 * bytes, 16-byte aligned (NULL allowed when n = 0, as are h and m). */
uint64_t cz_mass_sums_workspace(uint64_t n);
int cz_mass_sums(const float* h, const float* m, uint64_t n, void* workspace, uint64_t workspace_bytes,
                 cz_mass_totals* out, void* stream);
/* No counterpart in the reference.  cz_mass_sums for callers that cannot
 * allocate (code under HIP graph capture, serving loops that forbid
 * hipMalloc): the partials live in a static device array of the library
 * (192 KiB per GPU, held until the process ends), so the call allocates
 * nothing; the caller passes h, m and out as above.  At most 268,435,456
 * cells (4096 chunks) per call, else CZ_EINVAL.  All calls on one GPU share
 * the array: calls that overlap on a GPU race. */
int cz_mass_sums_noalloc(const float* h, const float* m, uint64_t n, cz_mass_totals* out, void* stream);
/* state_buffers sizes, read_state, write_state (mod.rs:1277-1320). */
int cz_film_state_sizes(const cz_film* f, uint64_t sizes[CZ_NPARTS]);
int cz_film_read_state(cz_film* f, void* const parts[CZ_NPARTS]);
int cz_film_write_state(cz_film* f, const void* const parts[CZ_NPARTS],
                        const uint64_t sizes[CZ_NPARTS]);
/* Any state part or CZ_BUF_* solver buffer, host copies (tests/checkpoints). */
int cz_film_buffer_size(const cz_film* f, int which, uint64_t* bytes);
int cz_film_read_buffer(cz_film* f, int which, void* host, uint64_t bytes);
int cz_film_write_buffer(cz_film* f, int which, const void* host, uint64_t bytes);
/* Device pointer of a buffer (zero-copy interop, e.g. torch tensors). */
int cz_film_device_ptr(cz_film* f, int which, void** ptr, uint64_t* bytes);
/* dry_now (mod.rs:1230-1244) and clear (mod.rs:1009-1028). */
int cz_film_dry_now(cz_film* f);
int cz_film_clear(cz_film* f);
/* The HIP stream the film's work is enqueued on (hipStream_t). */
void* cz_film_stream(cz_film* f);
/* Replay each frame of cz_film_encode_frame from a hipGraph captured once
 * per frame shape (substeps, implicit steps, deposit pass, offset-buffer
 * parity, params, grid sizes, flux shape). 1 = on (default), 0 = direct
 * launches.  Frames whose flux size class is still being timed by the shape
 * tuner, frames of more than 64 active tiles (whose launches run long enough
 * that a replay saves nothing) and profiled frames run direct; a shape is
 * captured the second time it comes. */
int cz_film_use_graphs(cz_film* f, int on);
/* How the substeps of a frame are launched: 0 = three launches per substep
 * (lap/plug, flux, limiter + apply), 1 = one launch for all of them with
 * grid-wide barriers between the passes (k_substeps), -1 = chosen per frame
 * (default; currently always the three launches, measured faster on gfx950).
 * Results are identical; no counterpart in the reference, whose encoder
 * records the passes of every substep. */
int cz_film_substep_launch(cz_film* f, int mode);
/* The display pass (shade_full, shaders/shade.wgsl:29-83, src/film/mod.rs:
 * 1449-1456) at the end of every frame: 1 = on, 0 = off (default; the
 * reference runs it every frame for its renderer).  Its two outputs, the
 * color and detail textures as f16 x 4 per cell (round to nearest even),
 * are read with CZ_BUF_COLOR and CZ_BUF_DETAIL; they are allocated when
 * shading is first turned on. */
int cz_film_shade_output(cz_film* f, int on);
/* Where upload_deposits' tile binning (src/film/mod.rs:1175-1224) runs: 0 on
 * the host (as the reference), 1 on the device (three kernels in the frame:
 * rectangles and counts, an ordered scan, an ordered fill), -1 = by deposit
 * count (default: the device from 1024 deposits in a frame).  The bins are
 * identical either way; the host keeps the volume and height accounting. */
int cz_film_bin_on_device(cz_film* f, int mode);
/* ---- Tool contacts and the palette knife (shaders/brush.wgsl,
 * shaders/knife.wgsl, src/film/mod.rs:1030-1173) ----------------------------
 * upload_contacts (1044-1084): the contacts for the next frame's brush pass
 * (before the deposits), binned per 64-cell tile on the host; they touch the
 * journal.  Consumed by that frame.  n > CZ_MAX_CONTACTS keeps the first. */
int cz_film_upload_contacts(cz_film* f, const cz_contact* cs, uint32_t n);
/* set_knife (1030-1042): the knife for the frames that follow (after the
 * deposits), NULL or active == 0 = none; an active knife touches the journal
 * over its reach. */
int cz_film_set_knife(cz_film* f, const cz_knife* k);
/* request_reservoir + poll_reservoir (1086-1133): the paint the brush pass
 * scraped since the last call, per bristle [CZ_MAX_BRISTLES][17] u32
 * (volume, then 16 pigment volumes, in units of 1e-15 m^3); the device sums
 * are cleared.  Waits for the stream. */
int cz_film_take_reservoir(cz_film* f, uint32_t* out);
/* request_scraps + poll_scraps (1135-1173): the knife's scraps since the
 * last call, in ascending cell order, taken paint before laid paint of a
 * cell, frame by frame (the reference appends them with an atomic counter in
 * whatever order cells run).  *count = scraps recorded (at most
 * CZ_MAX_SCRAPS are kept); min(*count, cap) are copied to out.  The device
 * list is cleared.  Waits for the stream. */
int cz_film_take_scraps(cz_film* f, cz_scrap* out, uint32_t cap, uint32_t* count);

/* ---- Undo journal (src/film/mod.rs:357-443, 1246-1275; shaders/journal.wgsl)
 * While the generation is nonzero, the first pass of every frame saves each
 * 16-cell tile that the frame's deposits or cz_film_touch_rect touched, or
 * that lies next to a tile active in the last frame, once per generation,
 * into a ring of `cap` tile records (53 words per cell: h, m, rheo,
 * strata_h, strata_pig, strata_cure, sub_refl, lay).  The tiles of one frame
 * take ring sequence numbers in ascending tile order (the reference's
 * atomic order is unspecified).  A film stepped by a batch cannot journal. */
/* Journal.gen (src/session.rs:139-143); 0 = off (the default).  The ring
 * is allocated at the first nonzero generation. */
int cz_film_journal_gen(cz_film* f, uint32_t gen);
/* PAINT_UNDO_MB (src/film/mod.rs:374-376): the ring's budget in bytes, set
 * before the first nonzero generation; cap = max(bytes / 54272, 64) tiles
 * (default 384 MiB: 7419 tiles). */
int cz_film_journal_budget(cz_film* f, uint64_t bytes);
/* Journal::head (src/film/mod.rs:414-424), the sequence number of the next
 * record (waits for the stream), and the ring's capacity in tiles (0 before
 * the ring exists). */
int cz_film_journal_head(cz_film* f, uint32_t* head, uint32_t* cap);
/* Film::touch_rect (src/film/mod.rs:1246-1266): mark the 16-cell tiles of a
 * rectangle in metres for the next frame's journal pass (ignored while the
 * generation is 0 or for a non-finite rectangle). */
int cz_film_touch_rect(cz_film* f, float minx, float miny, float maxx, float maxy);
/* Film::restore_journal (src/film/mod.rs:1272-1275, journal.wgsl:100-123):
 * every tile with a live record of generation >= target goes back to its
 * oldest such record and is marked moved. */
int cz_film_restore_journal(cz_film* f, uint32_t target);

/* Frame graphs captured and frames replayed from them so far. */
int cz_film_graph_counts(const cz_film* f, uint64_t* captures, uint64_t* replays);
/* Launch shape of a substep (lap + plug, flux, limiter + apply): 0 = flux
 * with one block per tile over an LDS window of the tile and a 4-cell ring,
 * 1 = flux with one block per (tile, face) reading through the caches, 2 =
 * flux and apply in one launch (each block computes the fluxes of its tile
 * and a 2-cell ring itself; for small active sets), 3 and 4 = shapes 0 and 2
 * with 1024 threads per tile (every window phase in one pass, the flux one
 * face per thread; for a few hundred active tiles), 5 = the whole substep in
 * one launch (each block recomputes the nose, pressures and plug means of a
 * 7-cell ring; heights alternate between two buffers over the frame's
 * substeps, and a frame whose parity would make its first substep read and
 * write the same buffer runs that substep as shape 4), 6 = the face-parallel
 * flux of shape 1 between the 1024-thread lap/plug and limiter/apply of
 * shape 3 (for a few hundred active tiles), -1 = chosen online
 * (default): per active-set size the shapes are timed with HIP events on live
 * substeps (they give bit-identical results) and the fastest is used. */
int cz_film_flux_shape(cz_film* f, int shape);
/* Average per-kernel device time over the last frames (HIP events on the
 * film's stream), in microseconds; label = kernel family name. */
int cz_film_profile(cz_film* f, int on);
int cz_film_profile_read(cz_film* f, const char* label, double* avg_us, uint64_t* launches);
/* Restrict the event timing of cz_film_profile to one kernel family
 * (NULL or "" = all). */
int cz_film_profile_only(cz_film* f, const char* label);
/* Zero the running sums of cz_stats. */
int cz_film_reset_counters(cz_film* f);

/* ---- Seeded canvases generated on the device -----------------------------
 * The plain workload's canvas (no counterpart in the reference, which
 * starts from an empty film): a sum of nblobs Gaussian blobs (centre
 * uniform, sigma uniform in cells, peak log-uniform, truncated at 3 sigma,
 * zero below 1e-9 m); every wet cell one stratum of Dirichlet(alpha)
 * pigments with log-uniform yield stress and viscosity in slots 13/14;
 * relief the canvas weave (mod.rs:212-233), substrate the raw canvas
 * (mod.rs:188-210).  With n_hi > 0 every wet cell also carries its own
 * power-law index, uniform in [n_lo, n_hi], in pigment slot 15 (read only
 * when the film's settings set n_per_cell).  Counter-based randomness: the
 * result depends only on the scene and the canvas size. */
typedef struct cz_scene {
    uint64_t seed;
    uint32_t nblobs;          /* <= 64 */
    float sigma_lo, sigma_hi; /* cells */
    float peak_lo, peak_hi;   /* m */
    float tau_lo, tau_hi;     /* Pa */
    float mu_lo, mu_hi;       /* Pa s */
    float alpha;
    float weave_um;
    float n_lo, n_hi;         /* per-cell power-law index; n_hi = 0: none */
} cz_scene;
void cz_scene_default(cz_scene* s);
/* Replace the film's state with the scene (as write_state does). */
int cz_film_seed_blobs(cz_film* f, const cz_scene* s);
/* Re-dimension a film within the sizes its buffers were created for, as a
 * new Film::new of w x h with settings s (fresh host state); the state
 * parts are then written or seeded by the caller.  CZ_ESIZE if it does not
 * fit.  Lets a pool of films serve canvases of varying size. */
int cz_film_reshape(cz_film* f, uint32_t w, uint32_t h, const cz_settings* s);
/* The same cells computed on the host, rows [j0, j1): h, m, rheo (2 per
 * cell) and stratum 0's packed pigments (8 per cell). */
int cz_scene_cells_host(uint32_t w, uint32_t h, float dx, const cz_scene* s, uint32_t j0, uint32_t j1,
                        float* h_out, float* m_out, float* rheo_out, uint32_t* pig0_out);

/* ---- Many canvases stepped together ------------------------------------
 * The reference steps one Film per App (encode_sim, src/main.rs:1875-1909);
 * a cz_batch steps many films, each exactly as cz_film_step would step it
 * alone without deposits (its own params, substeps, implicit steps and
 * lag-1 stats), with one launch per kernel phase over the union of the
 * members' tiles.  While a film is a member, the calls that enqueue work on
 * it (step, encode_frame, upload_deposits, write_state, write_buffer,
 * dry_now, clear) return CZ_EINVAL; its reads are valid after
 * cz_batch_sync.  Destroying a member removes it from its batch. */
typedef struct cz_batch cz_batch;
int cz_batch_create(uint32_t max_members, cz_batch** out);   /* max_members <= 1024 */
void cz_batch_destroy(cz_batch* b);
/* Join / leave between frames; a joining film's own pending work is
 * finished first.  Members must share settings.implicit_iters; a film with
 * settings.slump on is refused (CZ_EUNSUP). */
int cz_batch_add(cz_batch* b, cz_film* f);
int cz_batch_remove(cz_batch* b, cz_film* f);
uint32_t cz_batch_size(const cz_batch* b);
/* One frame of every member, enqueued on the batch's stream. */
int cz_batch_step(cz_batch* b, float frame_dt);
/* Wait for the last frame and fold its stats into every member. */
int cz_batch_sync(cz_batch* b);
/* Replay each frame shape from a captured hipGraph (1 = on, the default;
 * frames still timing flux shapes run direct). */
int cz_batch_use_graphs(cz_batch* b, int on);
/* As cz_film_graph_counts, for the batch's frames. */
int cz_batch_graph_counts(const cz_batch* b, uint64_t* captures, uint64_t* replays);
/* As cz_film_flux_shape, for the batched flux launches. */
int cz_batch_flux_shape(cz_batch* b, int shape);
void* cz_batch_stream(cz_batch* b);

/* ---- Serving ---------------------------------------------------------------
 * A loop that admits step requests (a seeded canvas and a frame count) as
 * they arrive, under a device-memory budget and a batch-size limit, steps all
 * admitted canvases together (cz_batch; each canvas exactly as alone) and
 * retires each when its frames are done.  With max_wait_s = 0 admission is
 * first come, first served, and canvases stay resident from admission to
 * completion.  With max_wait_s > 0 the waiting canvas with the least work
 * left (cells x frames) goes first, and any that fits is admitted past one
 * that does not; a request or evicted canvas waiting longer than max_wait_s
 * goes before all others and, when it does not fit, running canvases that
 * have run for at least max_wait_s since their admission and have more work
 * left than it are evicted to host memory (most work left first, within
 * host_budget) until it does.  An evicted canvas keeps all
 * it carries between frames (film_snapshot) and resumes exactly where it
 * stopped.  No counterpart in the reference (one canvas per App); its
 * per-canvas step is encode_sim (src/main.rs:1875-1909). */
typedef struct cz_request {
    uint64_t id;
    uint32_t w, h;
    uint32_t steps;          /* frames */
    uint32_t pad;
    double arrival_s;        /* after the start of cz_server_run; non-decreasing */
    cz_settings settings;    /* width_m = w * cell size */
    cz_scene scene;          /* initial state (cz_film_seed_blobs) */
} cz_request;
typedef struct cz_result {
    uint64_t id;
    double arrival_s, admit_s, done_s;  /* done: its last frame finished on the device */
    uint64_t cell_steps;
    int32_t status;          /* 0 done, else the CZ_E* code it failed with */
    uint32_t pad;
} cz_result;
typedef struct cz_server_config {
    uint64_t hbm_budget;     /* bytes of films held at once (active and pooled) */
    uint32_t max_members;    /* canvases stepped together, <= 1024 */
    float frame_dt;
    int32_t keep_films;      /* 1: keep finished canvases for cz_server_film (outside the budget) */
    float max_wait_s;        /* 0: first come, first served; > 0: least work first, see above */
    uint64_t host_budget;    /* bytes of evicted canvas state held in host memory at once */
} cz_server_config;
typedef struct cz_server_stats {
    double wall_s;
    uint64_t frames, admitted, cell_steps, peak_members, peak_bytes;
    uint64_t films_created, films_reused;
    uint64_t evictions, resumes;   /* canvases moved to host memory and back */
    uint64_t peak_host_bytes;      /* evicted canvas state held at once */
} cz_server_stats;
typedef struct cz_server cz_server;
void cz_server_config_default(cz_server_config* c);
int cz_server_create(const cz_server_config* c, cz_server** out);
void cz_server_destroy(cz_server* s);
/* Serve the requests in real time (arrival_s measured from the call) until
 * every one is done; res[k] belongs to reqs[k]. */
int cz_server_run(cz_server* s, const cz_request* reqs, uint32_t n, cz_result* res, cz_server_stats* st);
/* A finished canvas kept with keep_films (owned by the server), or NULL. */
cz_film* cz_server_film(cz_server* s, uint64_t id);

/* Painting files (.canvaz; src/session.rs:59-99,248-317).  save_painting:
 * the 44-byte header ("PSIM", version 1, w, h, tilt_deg, tilt_dir_deg,
 * cure_rate, h_max, s_max, total_deposited) and pack() of the eight
 * read_state parts, each cut into 32 MiB chunks compressed as LZ4 blocks
 * with the uncompressed size prepended (lz4_flex::compress_prepend_size).
 * open_painting for a film of the file's size: write_state, then the
 * header's settings and scalars.  The file is read whole; a size mismatch
 * is CZ_ESIZE. */
int cz_film_save_canvaz(cz_film* f, const char* path);
int cz_film_load_canvaz(cz_film* f, const char* path);
/* The LZ4 framing on its own: one block of src with its size prepended
 * (lz4_flex::compress_prepend_size), its inverse
 * (lz4_flex::decompress_size_prepended), and the worst-case output size of
 * the former. */
int cz_lz4_compress(const void* src, uint64_t n, void* dst, uint64_t cap, uint64_t* out_len);
int cz_lz4_decompress(const void* src, uint64_t n, void* dst, uint64_t cap, uint64_t* out_len);
uint64_t cz_lz4_bound(uint64_t n);

/* Device self-check with no counterpart in the reference: the kernels pack
 * and unpack f16 slots (pigments, substrate bands) with the hardware
 * conversions; this counts, over every non-NaN f32 and f16 bit pattern, the
 * inputs where they differ from the software round-to-nearest-even forms
 * the host restatement uses (both counts are 0 on a conforming device). */
int cz_f16_selftest(uint64_t* bad_f2h, uint64_t* bad_h2f);
/* No counterpart in the reference: the bandwidth probe the measurements use
 * as the device's delivered HBM ceiling.  mode 0 copies `bytes` from src to
 * dst, mode 1 fills dst; grid-stride 16-byte non-temporal accesses, enqueued
 * on `stream` (hipStream_t, NULL = default).  bytes and both pointers must be
 * multiples of 16. */
int cz_stream_probe(void* dst, const void* src, uint64_t bytes, int mode, void* stream);

const char* cz_last_error(void);
const char* cz_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CANVAZ_FILM_H */
