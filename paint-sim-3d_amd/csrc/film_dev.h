// film_dev.h - device view of one canvas and the per-cell device functions
// of the thin-film step (gfx950).  Formulas follow shaders/film.wgsl,
// strata.wgsl, deposit.wgsl, cure.wgsl and km.wgsl of the reference, cited
// per function; the evaluation order of every float expression is the WGSL
// source order (the library is built with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/canvaz_film.h"
#include "czmath.h"

namespace cz {

constexpr float DRAINED_FILM = 1e-9f;
constexpr uint32_t WAKE_BAND = 8u;
constexpr int PLUG_RADIUS = 3;
constexpr float YIELD_MENISCUS = 1e-3f;
constexpr uint32_t MOVED = 64u;
constexpr uint32_t NEEDS_CURE = 128u;  // k_mark -> compaction only, never stored
constexpr float W_AXIS = 0.6666667f;
constexpr float W_DIAG_FLUX = 0.23570226f;
constexpr float W_DIAG_LAP = 0.16666667f;
constexpr float MOVE_EPS = 1e-6f;
constexpr float FRAME_EPS = 2e-6f;
constexpr float COX_LOG = 132.0f;
constexpr float OIL_SHRINK = 0.04f;
constexpr int NPIG = 12;
constexpr int NBAND = 16;
constexpr int PIG_STRIDE = 40;
constexpr int TILE64 = 64;

// Everything a kernel needs about one canvas, passed by value as a kernel
// argument.  Layout in HBM is the reference's row-major SoA/AoS per buffer.
struct FilmDev {
    cz_params P;
    uint32_t W, H, n, tw, th, nt16, t64x, t64y, pw, ph;
    uint32_t imp_steps;   // implicit steps of the frame being encoded
    // state
    float* h;
    float* m;
    float2* rheo;
    float4* strata_h;
    uint4* strata_pig;   // [n][4 strata][2] uint4 = 8 u32 per stratum
    float4* strata_cure;
    uint32_t* sub_refl;  // [n][8]
    uint32_t* lay;
    // persistent solver buffers
    float2* pres;
    float4* flux;
    float4* acc;
    float* alpha;
    float* h_sub;         // heights at the start of the substep's apply (written by the lap pass)
    uint32_t* delta;      // current nose offsets (read)
    uint32_t* delta_out;  // written by nose
    float* tmp;
    float4* acc_s;        // [n][4]
    float4* alpha_s;
    float4* strata_h2;
    uint4* strata_pig2;
    float4* strata_cure2;
    float* h_frame;
    float2* fnorm;
    float4* cg_m;
    float *cg_d, *cg_x, *cg_r, *cg_p, *cg_q, *cg_t, *cg_t2;
    double* cg_part;
    double* cg_part2;
    float* cg_s;          // [8]
    uint32_t* tile_flags;
    uint32_t* tile_list;
    uint32_t* tile_moved;
    uint32_t* active_count;
    uint32_t* cure_list;      // tiles with cells for the wet cure branch
    uint32_t* cure_count;
    float2* stats_out;
    float* smax_out;       // [tiles64] largest finite surface of the 64-tile's pool cells
    float* pool_out;
    float* tsum_a;        // strata sums, fill pass input
    const float* pigtab;
    // deposits
    const cz_deposit* deposits;
    uint32_t* tile_start;   // per 64-tile: first item (written by the binning, read by k_deposit)
    uint32_t* tile_count;
    uint32_t* tile_items;
    float* frame_out;     // [16] per-frame scalars read by the host
    uint32_t* grid_sync;  // [4] barrier of the one-launch substeps: count, generation, timeout
    uint8_t* tile_bits;   // [nt16] compaction: 1 active, 2 cure
    uint32_t* compact_cnt; // [2 per 256 tiles] compaction: active and cure counts per block
    int32_t* dep_rect;     // [CZ_MAX_DEPOSITS][4] device binning: 64-tile rectangle per deposit (-1: none)
    uint32_t* dep_cnt;     // [tiles64] device binning: deposits per tile (zero between frames)
    uint32_t* dep_n;       // [1] device binning: the frame's deposit count
    const float* cmf;      // [16][3] CIE 1931 x D65 per band pair (shade)
    uint2* color;          // [n] shade output: rgb + roughness, f16 x 4
    uint2* detail;         // [n] shade output: method + turn, cure, yield stress, thickness, f16 x 4
    float* n_cell;         // [n] per-cell power-law index (with CZ_FLAG_CELL_N)
    // [n] ln max(rheo.x, 1e-4), the log viscosity the faces and the plug
    // read (film.wgsl:324, 68): written wherever the rheology is, by the
    // cure pass for every cell at the start of each frame before any
    // reader, and copied with it by apply's inheritance
    float* lmu_cell;
    // Constants of the face terms that depend on P alone, computed when the
    // descriptor is built (derive_consts below) with the IEEE operations the
    // face code would otherwise repeat per face: the same bits.
    float k_hold1;         // hold_level(1) (film.wgsl:154-156)
    float k_theta;         // atan(tan theta), the static contact angle (film.wgsl:238)
    float k_rim_reach;     // 3 / front.y: reach of the rim translation (film.wgsl:273)
    float k_gc_max;        // sigma / meniscus^2, yield_cap's bound (film.wgsl:46)
    float k_gb_plug;       // rho |g_xy|, the plug's tilt gradient (film.wgsl:62)
    float k_lg_plug;       // ln(rho |g_xy|)
    float k_gt_norm[4];    // |g_tilt| of the owned faces +x, +y, +x+y, +x-y (plug_mobility, film.wgsl:52)
};

// film.wgsl:154-156, on the parameter block (host and device)
CZ_HD inline float hold_level_p(const cz_params& P, float d) {
    return P.front[0] * (1.0f - cz_expf(-(cz_max(d, 1.0f) - 0.5f) * P.front[1]));
}

// The tilt gradient of owned face k (film.wgsl:387-439): +x, +y, +x+y, +x-y.
CZ_HD inline void face_tilt(const cz_params& P, int k, float& tx, float& ty) {
    const float gx = P.phys[0] * P.g[0], gy = P.phys[0] * P.g[1];
    const float s2 = 0.70710678f;
    if (k == 0) { tx = gx; ty = gy; }
    else if (k == 1) { tx = gy; ty = gx; }
    else if (k == 2) { tx = (gx + gy) * s2; ty = (gx - gy) * s2; }
    else { tx = (gx - gy) * s2; ty = (gx + gy) * s2; }
}

constexpr float YIELD_MENISCUS_SQ = YIELD_MENISCUS * YIELD_MENISCUS;

// Fill the k_* constants of a descriptor from its P.
inline void derive_consts(FilmDev& F) {
    const cz_params& P = F.P;
    F.k_hold1 = hold_level_p(P, 1.0f);
    F.k_theta = cz_atanf(P.ein[0]);
    F.k_rim_reach = 3.0f / P.front[1];
    F.k_gc_max = P.phys[1] / YIELD_MENISCUS_SQ;
    F.k_gb_plug = P.phys[0] * cz_len2(P.g[0], P.g[1]);
    F.k_lg_plug = cz_logf(F.k_gb_plug);
    for (int k = 0; k < 4; k++) {
        float tx, ty;
        face_tilt(P, k, tx, ty);
        F.k_gt_norm[k] = cz_len2(tx, ty);
    }
}

__device__ __forceinline__ uint32_t idx(const FilmDev& F, uint32_t i, uint32_t j) { return j * F.W + i; }

// Per-cell power-law index, an extension of the reference (whose index is the
// film's P.phys.z, film.wgsl:95), on when Params.flags[2] has CZ_FLAG_CELL_N:
// each stratum's pigment slot 15, unused by the reference, holds its index
// (0 = the film's); a cell's index is the thickness-weighted mean over its
// strata, formed by the cure pass with the rheology (F.n_cell), and a face
// takes the height-weighted mean of its two cells' indices, as it takes the
// yield stress (film.wgsl:326-329).  Off, every index is P.phys.z.
__device__ __forceinline__ bool cell_n_on(const FilmDev& F) { return (F.P.flags[2] & CZ_FLAG_CELL_N) != 0u; }
// t: stratum thicknesses, n15: their slot-15 values
__device__ __forceinline__ float strata_n(const FilmDev& F, const float* t, const float* n15) {
    float tot = 0.0f, acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (t[k] <= 0.0f) continue;
        const float nk = (n15[k] > 0.0f) ? n15[k] : F.P.phys[2];
        tot = tot + t[k];
        acc = acc + t[k] * nk;
    }
    return (tot > 0.0f) ? acc / tot : F.P.phys[2];
}

// Mirrored coordinates (shaders/common.wgsl:152-166: -1 - i below 0,
// 2W - 1 - i from W, then clamped), for coordinates at most one cell outside
// the canvas, which is every caller's case (-1 <= i <= W): the mirror of -1
// is 0 and of W is W - 1, i.e. the clamp.
__device__ __forceinline__ uint32_t refl1_i(const FilmDev& F, int i) { return (uint32_t)cz_clamp_i(i, 0, (int)F.W - 1); }
__device__ __forceinline__ uint32_t refl1_j(const FilmDev& F, int j) { return (uint32_t)cz_clamp_i(j, 0, (int)F.H - 1); }

// common.wgsl:127-130
__device__ __forceinline__ bool cell_active(const FilmDev& F, uint32_t i, uint32_t j) {
    return (F.tile_flags[(i >> 4) + (j >> 4) * F.tw] & 32u) != 0u;
}

struct Pig16 { float v[16]; };

__device__ __forceinline__ void unpack8(const uint4 a, const uint4 b, float* o) {
    o[0] = cz_unpack_lo(a.x); o[1] = cz_unpack_hi(a.x); o[2] = cz_unpack_lo(a.y); o[3] = cz_unpack_hi(a.y);
    o[4] = cz_unpack_lo(a.z); o[5] = cz_unpack_hi(a.z); o[6] = cz_unpack_lo(a.w); o[7] = cz_unpack_hi(a.w);
    o[8] = cz_unpack_lo(b.x); o[9] = cz_unpack_hi(b.x); o[10] = cz_unpack_lo(b.y); o[11] = cz_unpack_hi(b.y);
    o[12] = cz_unpack_lo(b.z); o[13] = cz_unpack_hi(b.z); o[14] = cz_unpack_lo(b.w); o[15] = cz_unpack_hi(b.w);
}
__device__ __forceinline__ void pack8(const float* o, uint4& a, uint4& b) {
    a.x = cz_pack2(o[0], o[1]); a.y = cz_pack2(o[2], o[3]); a.z = cz_pack2(o[4], o[5]); a.w = cz_pack2(o[6], o[7]);
    b.x = cz_pack2(o[8], o[9]); b.y = cz_pack2(o[10], o[11]); b.z = cz_pack2(o[12], o[13]); b.w = cz_pack2(o[14], o[15]);
}

// common.wgsl:190-195
__device__ __forceinline__ float disjoining(const FilmDev& F, float hv) {
    float hs = F.P.cure[3];
    float a = F.P.ein[2];
    float x = cz_max(1.0f - hv / hs, 0.0f);
    return a * x * x;
}

// film.wgsl:570-589
__device__ __forceinline__ uint32_t pack_off(int dx, int dy) {
    uint32_t ux = (uint32_t)cz_clamp_i(dx + 128, 0, 255);
    uint32_t uy = (uint32_t)cz_clamp_i(dy + 128, 0, 255);
    return (ux << 8u) | uy | 0x10000u;
}
__device__ __forceinline__ void unpack_off(uint32_t bits, int& ox, int& oy) {
    if ((bits & 0x10000u) == 0u) { ox = 64; oy = 64; return; }
    ox = (int)((bits >> 8u) & 0xffu) - 128;
    oy = (int)(bits & 0xffu) - 128;
}
__device__ __forceinline__ float nose_dist_b(uint32_t bits) {
    int ox, oy;
    unpack_off(bits, ox, oy);
    return cz_len2((float)ox, (float)oy);
}
__device__ __forceinline__ float nose_dist(const FilmDev& F, uint32_t c) { return nose_dist_b(F.delta[c]); }

// a / (a^2 + b^2) for a, b in -2..2 (0 at the centre): the weights of the
// front normal's sum (film.wgsl:160-180), folded at compile time with the
// same IEEE division the shader performs.
struct FnQuot {
    float v[5][5];
    constexpr FnQuot() : v() {
        for (int a = -2; a <= 2; a++)
            for (int b = -2; b <= 2; b++)
                v[a + 2][b + 2] = (a == 0 && b == 0) ? 0.0f : (float)a / (float)(a * a + b * b);
    }
};
__device__ constexpr FnQuot FN_QUOT{};

// film.wgsl:154-156
__device__ __forceinline__ float hold_level(const FilmDev& F, float d) { return hold_level_p(F.P, d); }

// film.wgsl:44-49
__device__ __forceinline__ float2 yield_cap(const FilmDev& F, float2 gc, float ty) {
    float gcl = cz_len2(gc.x, gc.y);
    float gc_max = F.k_gc_max;
    if (ty <= 0.0f || gcl <= gc_max) return gc;
    return make_float2(gc.x * (gc_max / gcl), gc.y * (gc_max / gcl));
}

// The viscosity argument of hb_flux: its logarithm, and the value when the
// caller has it (has_mu; else the Newtonian branch forms exp(lmu)).
struct Visc {
    float mu;
    float lmu;
    bool has_mu;
};

// the rheology of cell c and its log viscosity
__device__ __forceinline__ void set_rheo(const FilmDev& F, uint32_t c, float2 r) {
    F.rheo[c] = r;
    F.lmu_cell[c] = cz_logf(cz_max(r.x, 1e-4f));
}

// film.wgsl:94-107.  The power-law velocity scale n/(n+1) (g/mu)^(1/n)
// y^((n+1)/n) is evaluated as n/(n+1) exp((ln g - ln mu + (n+1) ln y) / n),
// one exponential and three logarithms at most where the shader's two pow
// calls (exp(e ln x) each, by the WGSL definition) take two of each and two
// more divisions.  lg: ln g when the caller has it (lg_known); mu: the
// viscosity's logarithm (the power-law branch) and value (the Newtonian).
__device__ __forceinline__ float hb_flux(float n, float gmag, bool lg_known, float lg, float hf, Visc mu, float ty,
                                         float& oy, float& oup) {
    oy = 0.0f;
    oup = 0.0f;
    if (gmag < 1e-9f || hf <= 0.0f) return 0.0f;
    float y = cz_max(hf - ty / gmag, 0.0f);
    if (y <= 0.0f) return 0.0f;
    float up;
    if (cz_abs(n - 1.0f) < 1e-4f) {
        const float mv = mu.has_mu ? mu.mu : cz_expf(mu.lmu);
        up = 0.5f * gmag / mv * y * y;
    } else {
        const float lgv = lg_known ? lg : cz_logf(gmag);
        up = n / (n + 1.0f) * cz_expf((lgv - mu.lmu + (n + 1.0f) * cz_logf(y)) / n);
    }
    float q = up * (hf - n * y / (2.0f * n + 1.0f));
    oy = y;
    oup = up;
    return q;
}

// film.wgsl:109-126
__device__ __forceinline__ float band_flux(float n, float up, float y, float za, float zb) {
    float ta = 0.0f, tb = 0.0f;
    if (y > 0.0f) {
        float ra = cz_max(y - za, 0.0f) / y;
        float rb = cz_max(y - zb, 0.0f) / y;
        if (cz_abs(n - 1.0f) < 1e-4f) {
            ta = ra * ra * ra;
            tb = rb * rb * rb;
        } else {
            float e = (2.0f * n + 1.0f) / n;
            ta = cz_powf(ra, e);
            tb = cz_powf(rb, e);
        }
    }
    return up * ((zb - za) - n / (2.0f * n + 1.0f) * y * (ta - tb));
}

// The face terms below (film.wgsl:134-372, 387-439, 895-920) read the
// per-cell fields through an accessor, by cell coordinates inside the canvas:
// GlobalRead goes to HBM; the flux kernel's window reader serves its tile and
// halo from LDS and only the far steps of cell_excess (h_far, delta_far) from
// HBM.  The arithmetic is the same whichever accessor is used.
struct GlobalRead {
    // far reads are plain loads: cell_excess may issue its three at once
    static constexpr bool kFarCheap = true;
    const FilmDev& F;
    __device__ __forceinline__ uint32_t at(int i, int j) const { return (uint32_t)j * F.W + (uint32_t)i; }
    __device__ __forceinline__ float h(int i, int j) const { return F.h[at(i, j)]; }
    __device__ __forceinline__ float2 rheo(int i, int j) const { return F.rheo[at(i, j)]; }
    __device__ __forceinline__ uint32_t delta(int i, int j) const { return F.delta[at(i, j)]; }
    __device__ __forceinline__ float2 fnorm(int i, int j) const { return F.fnorm[at(i, j)]; }
    __device__ __forceinline__ float2 pres(int i, int j) const { return F.pres[at(i, j)]; }
    __device__ __forceinline__ float tmp(int i, int j) const { return F.tmp[at(i, j)]; }
    __device__ __forceinline__ float ncell(int i, int j) const { return F.n_cell[at(i, j)]; }
    __device__ __forceinline__ bool active(int i, int j) const { return cell_active(F, (uint32_t)i, (uint32_t)j); }
    __device__ __forceinline__ float h_far(int i, int j) const { return h(i, j); }
    __device__ __forceinline__ uint32_t delta_far(int i, int j) const { return delta(i, j); }
    // ln of the clamped viscosity, front distance, hold level at d = 1 and
    // the static contact angle: evaluated on demand here, staged by the
    // flux kernel's window reader
    __device__ __forceinline__ float lmu(int i, int j) const { return F.lmu_cell[at(i, j)]; }
    __device__ __forceinline__ float fdist(int i, int j) const {
        return h(i, j) <= 0.0f ? 0.0f : nose_dist_b(delta(i, j));
    }
    __device__ __forceinline__ float hold1() const { return F.k_hold1; }
    __device__ __forceinline__ float theta() const { return F.k_theta; }
};

// film.wgsl:134-145
template <class A>
__device__ __forceinline__ uint32_t feeders(const FilmDev& F, const A& a, int i, int j) {
    float hc = a.h(i, j);
    float thr = 2.0f * hc + F.P.ein[3];
    uint32_t n = 0;
    if (i > 0 && a.h(i - 1, j) > thr) n++;
    if (i + 1 < (int)F.W && a.h(i + 1, j) > thr) n++;
    if (j > 0 && a.h(i, j - 1) > thr) n++;
    if (j + 1 < (int)F.H && a.h(i, j + 1) > thr) n++;
    return n;
}

// film.wgsl:160-180
template <class A>
__device__ __forceinline__ float2 front_normal(const FilmDev& F, const A& a, int i, int j) {
    float dx = 0.0f, dy = 0.0f;
    for (int dj = -2; dj <= 2; dj++) {
        for (int di = -2; di <= 2; di++) {
            if (di == 0 && dj == 0) continue;
            int x = i + di, y = j + dj;
            if (x < 0 || y < 0 || x >= (int)F.W || y >= (int)F.H) continue;
            if (di * di + dj * dj > 5) continue;
            if (a.h(x, y) <= F.P.ein[3]) {
                dx = dx + FN_QUOT.v[di + 2][dj + 2];
                dy = dy + FN_QUOT.v[dj + 2][di + 2];
            }
        }
    }
    float l = cz_len2(dx, dy);
    if (l < 1e-6f) return make_float2(0.0f, 0.0f);
    return make_float2(dx / l, dy / l);
}

// film.wgsl:182-188
__device__ __forceinline__ float wet_gain(float2 n) {
    float s2 = 0.70710678f;
    float w = W_AXIS * (cz_max(n.x, 0.0f) + cz_max(-n.x, 0.0f) + cz_max(n.y, 0.0f) + cz_max(-n.y, 0.0f));
    w = w + W_DIAG_FLUX * (cz_max(s2 * (n.x + n.y), 0.0f) + cz_max(-s2 * (n.x + n.y), 0.0f) +
                           cz_max(s2 * (n.x - n.y), 0.0f) + cz_max(-s2 * (n.x - n.y), 0.0f));
    float m = cz_max(cz_max(cz_abs(n.x), cz_abs(n.y)), 1e-3f);
    return 1.0f / (m * cz_max(w, 1e-3f));
}

// film.wgsl:203-222
template <class A>
__device__ __forceinline__ float front_shape(const FilmDev& F, const A& a, int i, int j, float2 n) {
    float w = 0.0f, w0 = 0.0f;
    for (int dj = -3; dj <= 3; dj++) {
        for (int di = -3; di <= 3; di++) {
            if (di == 0 && dj == 0) continue;
            if (di * di + dj * dj > 9) continue;
            float ox = (float)di, oy = (float)dj;
            if (ox * n.x + oy * n.y < -0.5f) w0 = w0 + 1.0f;
            int x = i + di, y = j + dj;
            if (x < 0 || y < 0 || x >= (int)F.W || y >= (int)F.H) continue;
            if (a.h(x, y) > F.P.ein[3]) w = w + 1.0f;
        }
    }
    if (w0 < 1.0f || (((uint32_t)F.P.front2[3]) & 2u) != 0u) return 1.0f;
    return cz_powf(cz_clamp(w / w0, 0.0f, 2.5f), F.P.front2[1]);
}

// film.wgsl:224-229
template <class A>
__device__ __forceinline__ float wet_threshold(const FilmDev& F, const A& a, int i, int j) {
    float ty = a.rheo(i, j).y;
    float need = cz_max(ty / (F.P.phys[0] * cz_max(a.h(i, j), 1e-4f)) - cz_len2(F.P.g[0], F.P.g[1]), 0.0f);
    float ycap = need / cz_max(F.P.g[2], 1e-3f) * F.P.dx;
    return cz_max(a.hold1(), ycap);
}

// film.wgsl:231-234
template <class A>
__device__ __forceinline__ float front_dist(const A& a, int i, int j) {
    return a.fdist(i, j);
}

// film.wgsl:236-244
template <class A>
__device__ __forceinline__ float cox_speed(const FilmDev& F, const A& a, float excess, float mu) {
    if (excess <= 0.0f) return 0.0f;
    float theta = a.theta();
    float span = cz_max(F.P.front[0] - a.hold1(), 1e-6f);
    float th = theta + (3.14159265f - theta) * cz_clamp(excess / span, 0.0f, 1.0f);
    float adv = cz_min(theta + F.P.front2[0], 3.1f);
    if (th <= adv) return 0.0f;
    return F.P.phys[1] * (th * th * th - adv * adv * adv) / (COX_LOG * cz_max(mu, 1e-4f));
}

// film.wgsl:246-265
template <class A>
__device__ __forceinline__ float cell_excess(const FilmDev& F, const A& a, int ci, int cj) {
    float e = a.h(ci, cj) - hold_level(F, front_dist(a, ci, cj));
    int ox, oy;
    unpack_off(a.delta(ci, cj), ox, oy);
    float ol = cz_len2((float)ox, (float)oy);
    if (ol > 0.5f && ol < 60.0f) {
        float inv = -1.0f / ol;
        float dirx = (float)ox * inv, diry = (float)oy * inv;
        float bx = 0.5f * cz_sign((float)(-ox)), by = 0.5f * cz_sign((float)(-oy));
        if constexpr (A::kFarCheap) {
            // the three steps' heights and offsets read together (a step
            // beyond the canvas reads the cell itself and is not used); the
            // walk then stops where the step-by-step walk stops
            float hs[3];
            uint32_t ds[3];
            bool ok[3];
#pragma unroll
            for (uint32_t k = 0; k < 3u; k++) {
                float reach = (float)(4u << k);
                int sx = (int)(dirx * reach + bx);
                int sy = (int)(diry * reach + by);
                int ri = ci + sx, rj = cj + sy;
                ok[k] = !(ri < 0 || rj < 0 || ri >= (int)F.W || rj >= (int)F.H);
                if (!ok[k]) { ri = ci; rj = cj; }
                hs[k] = a.h_far(ri, rj);
                ds[k] = a.delta_far(ri, rj);
            }
#pragma unroll
            for (uint32_t k = 0; k < 3u; k++) {
                if (!ok[k] || hs[k] <= 0.0f) break;
                e = cz_max(e, hs[k] - hold_level(F, nose_dist_b(ds[k])));
            }
        } else {
            for (uint32_t k = 0; k < 3u; k++) {
                float reach = (float)(4u << k);
                int sx = (int)(dirx * reach + bx);
                int sy = (int)(diry * reach + by);
                int ri = ci + sx, rj = cj + sy;
                if (ri < 0 || rj < 0 || ri >= (int)F.W || rj >= (int)F.H) break;
                float hr = a.h_far(ri, rj);
                if (hr <= 0.0f) break;
                e = cz_max(e, hr - hold_level(F, nose_dist_b(a.delta_far(ri, rj))));
            }
        }
    }
    return e;
}

// film.wgsl:267-295; returns the translation flux, sign of the normal in tsign
template <class A>
__device__ __forceinline__ float rim_translation(const FilmDev& F, const A& a, int ia, int ja, int ib, int jb, float hf,
                                        float2 face_dir, float& tsign) {
    tsign = 0.0f;
    if ((((uint32_t)F.P.front2[3]) & 1u) != 0u) return 0.0f;
    float da = front_dist(a, ia, ja);
    float db = front_dist(a, ib, jb);
    if (cz_max(da, db) > F.k_rim_reach) return 0.0f;
    float ha = a.h(ia, ja), hb = a.h(ib, jb);
    float nx = 0.0f, ny = 0.0f;
    if (da > db && ha > 0.0f) {
        float2 v = a.fnorm(ia, ja);
        nx = v.x; ny = v.y;
    } else if (db > da && hb > 0.0f) {
        float2 v = a.fnorm(ib, jb);
        nx = v.x; ny = v.y;
    } else {
        if (ha > 0.0f) { float2 v = a.fnorm(ia, ja); nx = nx + v.x; ny = ny + v.y; }
        if (hb > 0.0f) { float2 v = a.fnorm(ib, jb); nx = nx + v.x; ny = ny + v.y; }
    }
    float nl = cz_len2(nx, ny);
    if (nl < 1e-6f) return 0.0f;
    nx = nx / nl;
    ny = ny / nl;
    float s = nx * face_dir.x + ny * face_dir.y;
    if (cz_abs(s) < 1e-4f) return 0.0f;
    int si = ia, sj = ja, di = ib, dj = jb;
    if (s < 0.0f) { si = ib; sj = jb; di = ia; dj = ja; }
    if (a.h(si, sj) <= 0.0f) return 0.0f;
    float e = cell_excess(F, a, si, sj);
    if (e <= 0.0f) return 0.0f;
    float e_out = -a.hold1();
    if (a.h(di, dj) > 0.0f) e_out = cell_excess(F, a, di, dj);
    float room = cz_clamp(1.0f - cz_max(e_out, 0.0f) / e, 0.0f, 1.0f);
    float u = cox_speed(F, a, e, a.rheo(si, sj).x);
    tsign = cz_sign(s);
    return u * hf * room * cz_mix(1.0f, cz_abs(s), F.P.front2[0]);
}

struct FaceFlux { float q, y, up, n; };

// film.wgsl:297-372: face between cell a = (ia, ja) and cell b = (ib, jb)
// gtn: |gtilt|, the face's entry of F.k_gt_norm
template <class A>
__device__ __forceinline__ FaceFlux face(const FilmDev& F, const A& acc, int ia, int ja, int ib, int jb, float2 gcap,
                                float2 ghyd, float2 gtilt, float gtn, bool diag, float2 face_dir) {
    FaceFlux out = {0.0f, 0.0f, 0.0f, F.P.phys[2]};
    float ha = acc.h(ia, ja);
    float hb = acc.h(ib, jb);
    float hf = 0.5f * (ha + hb);
    if (hf <= 0.0f) return out;
    float shape = 1.0f;
    bool wetting = false;
    float proj = 1.0f;
    if (cz_min(ha, hb) <= 0.0f) {
        const int wi = (ha > 0.0f) ? ia : ib, wj = (ha > 0.0f) ? ja : jb;
        const int di = (ha > 0.0f) ? ib : ia, dj = (ha > 0.0f) ? jb : ja;
        if (acc.h(wi, wj) <= wet_threshold(F, acc, wi, wj)) return out;
        wetting = true;
        float2 n = acc.fnorm(wi, wj);
        if (n.x * n.x + n.y * n.y < 0.5f) n = front_normal(F, acc, wi, wj);
        float toward = (ha > 0.0f) ? 1.0f : -1.0f;
        proj = cz_max(n.x * (face_dir.x * toward) + n.y * (face_dir.y * toward), 0.0f);
        if (proj <= 0.0f) return out;
        shape = front_shape(F, acc, di, dj, n) * wet_gain(n);
        if (shape <= 0.0f) return out;
    }
    float2 ra = acc.rheo(ia, ja);
    float2 rb = acc.rheo(ib, jb);
    float wa = ha / (ha + hb);
    float wb = 1.0f - wa;
    // the face viscosity exp(wa ln mu_a + wb ln mu_b), kept as its logarithm
    const Visc mu{0.0f, wa * acc.lmu(ia, ja) + wb * acc.lmu(ib, jb), false};
    float ty = wa * ra.y + wb * rb.y;
    const float nf = cell_n_on(F) ? wa * acc.ncell(ia, ja) + wb * acc.ncell(ib, jb) : F.P.phys[2];
    float thin = cz_min(ha, hb) / hf;
    float fcap = thin * thin * thin;
    float2 yc = yield_cap(F, make_float2(gcap.x * fcap, gcap.y * fcap), ty);
    float2 g = make_float2(yc.x + ghyd.x + gtilt.x, yc.y + ghyd.y + gtilt.y);
    float gmag = cz_len2(g.x, g.y);
    float ydep = (ty <= 0.0f) ? hf : ((g.x >= 0.0f) ? ha : hb);
    float qy_y, qy_up;
    float qy_x = hb_flux(nf, gmag, false, 0.0f, ydep, mu, ty, qy_y, qy_up);
    float meff = (qy_x > 0.0f) ? qy_x / cz_max(gmag, 1e-9f) : 0.0f;
    float mtilt = meff;
    if (ty > 0.0f) {
        // plug_mobility (film.wgsl:51-56)
        const float gb = gtn;
        if (gb < 1e-6f) {
            mtilt = 0.0f;
        } else {
            const int ui = (gtilt.x >= 0.0f) ? ia : ib, uj = (gtilt.x >= 0.0f) ? ja : jb;
            mtilt = acc.tmp(ui, uj) * acc.h(ui, uj) / gb;
        }
    }
    if (meff <= 0.0f && mtilt <= 0.0f) return out;
    float fcap2 = cz_pow_fcap(thin, F.P.front3[3]);
    float qn = mtilt * gtilt.x + cz_min(meff, F.P.misc[3]) * ghyd.x + cz_min(meff, F.P.phys[3]) * (gcap.x * fcap2);
    float trs;
    float trx = rim_translation(F, acc, ia, ja, ib, jb, hf, face_dir, trs);
    if (wetting) {
        float hwet = (ha > 0.0f) ? ha : hb;
        float want = trx * 2.0f * hwet / cz_max(hf, 1e-9f);
        qn = trs * cz_max(qn * trs * proj, want) * shape;
        float vcap = 2.0f;
        if (cz_abs(qn) > vcap * hf) qn = cz_sign(qn) * vcap * hf;
    } else {
        if (trx > 0.0f) qn = trs * cz_max(qn * trs, trx);
        float vcap = 2.0f;
        if (cz_abs(qn) > vcap * hf) qn = cz_sign(qn) * vcap * hf;
        float lo = cz_min(ha, hb);
        float hi = cz_max(ha, hb);
        if (lo < 0.5f * hi) {
            if (diag) {
                qn = 0.0f;
            } else {
                const int ti = (hb > ha) ? ia : ib, tj = (hb > ha) ? ja : jb;
                qn = qn / cz_feed_pow(feeders(F, acc, ti, tj));
            }
        }
    }
    out.q = qn;
    out.y = qy_y;
    out.up = qy_up;
    out.n = nf;
    return out;
}

// pressure_at / pressure_wet: (i, j) at most one cell outside the canvas, as
// every face's neighbours are
template <class A>
__device__ __forceinline__ float2 pressure_at(const FilmDev& F, const A& a, int i, int j) {
    return a.pres((int)refl1_i(F, i), (int)refl1_j(F, j));
}

// film.wgsl:9-14
template <class A>
__device__ __forceinline__ float2 pressure_wet(const FilmDev& F, const A& a, int i, int j, float2 fallback) {
    const int ri = (int)refl1_i(F, i), rj = (int)refl1_j(F, j);
    float2 p = a.pres(ri, rj);
    if (a.h(ri, rj) > 0.0f) return p;
    return make_float2(fallback.x, p.y);
}

struct FaceGeom {
    bool has;
    bool diag;
    int ib, jb;   // the other cell of the face
    float2 gcap, ghyd, gtilt, dir;
    float gtn;    // |gtilt|
};

// Face geometry of the k-th owned face (+x, +y, +x+y, +x-y) of cell (x, y):
// pressure gradients of film.wgsl:387-439 / 940-978.
template <class A>
__device__ __forceinline__ FaceGeom face_geom(const FilmDev& F, const A& a, int i, int j, int k, bool open_check, float2 p00) {
    FaceGeom fg;
    fg.has = false;
    fg.diag = false;
    fg.ib = i;
    fg.jb = j;
    fg.gtn = 0.0f;
    const int W_ = (int)F.W, H_ = (int)F.H;
    float2 gxy = make_float2(F.P.phys[0] * F.P.g[0], F.P.phys[0] * F.P.g[1]);
    const float s2 = 0.70710678f;
    const float idx_ = F.P.inv_dx;
    if (k == 0) {
        if (!(i + 1 < W_ && a.active(i + 1, j))) return fg;
        fg.ib = i + 1;
        float2 p10 = pressure_at(F, a, i + 1, j);
        float gmx = -(p10.x - p00.x) * idx_, gmy = -(p10.y - p00.y) * idx_;
        float2 pa = pressure_wet(F, a, i, j + 1, p00), pb = pressure_wet(F, a, i + 1, j + 1, p10);
        float2 pc = pressure_wet(F, a, i, j - 1, p00), pd = pressure_wet(F, a, i + 1, j - 1, p10);
        float ptx = pa.x + pb.x - pc.x - pd.x, pty = pa.y + pb.y - pc.y - pd.y;
        fg.gcap = make_float2(gmx, -ptx * 0.25f * idx_);
        fg.ghyd = make_float2(gmy, -pty * 0.25f * idx_);
        fg.gtilt = gxy;
        fg.dir = make_float2(1.0f, 0.0f);
        fg.gtn = F.k_gt_norm[0];
        fg.has = true;
    } else if (k == 1) {
        if (!(j + 1 < H_ && a.active(i, j + 1))) return fg;
        fg.jb = j + 1;
        float2 p01 = pressure_at(F, a, i, j + 1);
        float gmx = -(p01.x - p00.x) * idx_, gmy = -(p01.y - p00.y) * idx_;
        float2 pa = pressure_wet(F, a, i + 1, j, p00), pb = pressure_wet(F, a, i + 1, j + 1, p01);
        float2 pc = pressure_wet(F, a, i - 1, j, p00), pd = pressure_wet(F, a, i - 1, j + 1, p01);
        float ptx = pa.x + pb.x - pc.x - pd.x, pty = pa.y + pb.y - pc.y - pd.y;
        fg.gcap = make_float2(gmx, -ptx * 0.25f * idx_);
        fg.ghyd = make_float2(gmy, -pty * 0.25f * idx_);
        fg.gtilt = make_float2(gxy.y, gxy.x);
        fg.dir = make_float2(0.0f, 1.0f);
        fg.gtn = F.k_gt_norm[1];
        fg.has = true;
    } else if (k == 2) {
        if (!(i + 1 < W_ && j + 1 < H_ && a.active(i + 1, j + 1))) return fg;
        fg.ib = i + 1;
        fg.jb = j + 1;
        if (open_check) {
            float hs = F.P.ein[3];
            bool open = cz_min(a.h(i, j), a.h(i + 1, j + 1)) > 0.0f || a.h(i + 1, j) > hs || a.h(i, j + 1) > hs;
            if (!open) return fg;
        }
        float2 p11 = pressure_at(F, a, i + 1, j + 1);
        float2 pa = pressure_wet(F, a, i + 1, j, p00);
        float2 pb = pressure_wet(F, a, i, j + 1, p11);
        fg.gcap = make_float2(-(p11.x - p00.x) * idx_ * s2, -(pa.x - pb.x) * idx_ * s2);
        fg.ghyd = make_float2(-(p11.y - p00.y) * idx_ * s2, -(pa.y - pb.y) * idx_ * s2);
        fg.gtilt = make_float2((gxy.x + gxy.y) * s2, (gxy.x - gxy.y) * s2);
        fg.dir = make_float2(s2, s2);
        fg.gtn = F.k_gt_norm[2];
        fg.diag = true;
        fg.has = true;
    } else {
        if (!(i + 1 < W_ && j > 0 && a.active(i + 1, j - 1))) return fg;
        fg.ib = i + 1;
        fg.jb = j - 1;
        if (open_check) {
            float hs = F.P.ein[3];
            bool open = cz_min(a.h(i, j), a.h(i + 1, j - 1)) > 0.0f || a.h(i + 1, j) > hs || a.h(i, j - 1) > hs;
            if (!open) return fg;
        }
        float2 p1m = pressure_at(F, a, i + 1, j - 1);
        float2 pa = pressure_wet(F, a, i + 1, j, p00);
        float2 pb = pressure_wet(F, a, i, j - 1, p1m);
        fg.gcap = make_float2(-(p1m.x - p00.x) * idx_ * s2, -(pa.x - pb.x) * idx_ * s2);
        fg.ghyd = make_float2(-(p1m.y - p00.y) * idx_ * s2, -(pa.y - pb.y) * idx_ * s2);
        fg.gtilt = make_float2((gxy.x - gxy.y) * s2, (gxy.x + gxy.y) * s2);
        fg.dir = make_float2(s2, -s2);
        fg.gtn = F.k_gt_norm[3];
        fg.diag = true;
        fg.has = true;
    }
    return fg;
}

// ---------------------------------------------------------------------------
// the 4-deep pigment stack (strata.wgsl:1-155)

struct Stack {
    float th[4];
    float cu[4];
    float pg[4][16];
};

__device__ inline void stack_load(const FilmDev& F, uint32_t c, Stack& s) {
    float4 t = F.strata_h[c];
    float4 u = F.strata_cure[c];
    s.th[0] = t.x; s.th[1] = t.y; s.th[2] = t.z; s.th[3] = t.w;
    s.cu[0] = u.x; s.cu[1] = u.y; s.cu[2] = u.z; s.cu[3] = u.w;
#pragma unroll
    for (int k = 0; k < 4; k++) unpack8(F.strata_pig[(c * 4u + k) * 2u], F.strata_pig[(c * 4u + k) * 2u + 1u], s.pg[k]);
}
__device__ inline void stack_store_to(float4* sh, float4* scu, uint4* spg, uint32_t c, const Stack& s) {
    sh[c] = make_float4(s.th[0], s.th[1], s.th[2], s.th[3]);
    scu[c] = make_float4(s.cu[0], s.cu[1], s.cu[2], s.cu[3]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint4 a, b;
        pack8(s.pg[k], a, b);
        spg[(c * 4u + k) * 2u] = a;
        spg[(c * 4u + k) * 2u + 1u] = b;
    }
}

__device__ __forceinline__ float pig_dist(const float* a, const float* b) {
    float d = 0.0f;
#pragma unroll
    for (int i = 0; i < 13; i++) d = d + cz_abs(a[i] - b[i]);
    return d;
}

// strata.wgsl:45-59 (o may alias a); one reciprocal of the weight and
// sixteen products (within 1.5 ULP of the quotient, inside WGSL's 2.5 ULP
// for the shader's divisions)
__device__ __forceinline__ void pig_mix(const float* a, float wa, const float* b, float wb, float* o) {
    float a14 = a[14], b14 = b[14];
    float w = cz_max(wa + wb, 1e-20f);
    float rw = 1.0f / w;
#pragma unroll
    for (int i = 0; i < 16; i++) o[i] = (a[i] * wa + b[i] * wb) * rw;
    float va = (a14 > 0.0f) ? wa : 0.0f;
    float vb = (b14 > 0.0f) ? wb : 0.0f;
    if (va + vb > 0.0f) o[14] = (a14 * va + b14 * vb) / (va + vb);
    else o[14] = 0.0f;
}

// The stack operations below index the stack arrays with compile-time
// indices only (loops fully unrolled, runtime positions turned into selects),
// so that a Stack lives in registers: a runtime index would place the whole
// struct in scratch memory.
__device__ __forceinline__ uint32_t stack_count(const Stack& s) {
    uint32_t n = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++)
        if (s.th[k] > 0.0f) n = k + 1u;
    return n;
}

__device__ inline void copy16(float* d, const float* s) {
#pragma unroll
    for (int i = 0; i < 16; i++) d[i] = s[i];
}

// strata.wgsl:69-83
template <uint32_t A>
__device__ __forceinline__ void stack_merge_at(Stack& s) {
    constexpr uint32_t B = A + 1u;
    float ta = s.th[A], tb = s.th[B];
    pig_mix(s.pg[A], ta, s.pg[B], tb, s.pg[A]);
    s.cu[A] = (s.cu[A] * ta + s.cu[B] * tb) / cz_max(ta + tb, 1e-20f);
    s.th[A] = ta + tb;
#pragma unroll
    for (uint32_t k = B; k < 3u; k++) {
        s.th[k] = s.th[k + 1u];
        s.cu[k] = s.cu[k + 1u];
        copy16(s.pg[k], s.pg[k + 1u]);
    }
    s.th[3] = 0.0f;
    s.cu[3] = 0.0f;
}
__device__ __forceinline__ void stack_merge(Stack& s, uint32_t a) {
    if (a == 0u) stack_merge_at<0>(s);
    else if (a == 1u) stack_merge_at<1>(s);
    else stack_merge_at<2>(s);
}

// strata.wgsl:85-103
__device__ inline void stack_compact(const FilmDev& F, Stack& s) {
#pragma unroll
    for (int sweep = 0; sweep < 4; sweep++) {
#pragma unroll
        for (uint32_t k = 0; k < 3u; k++) {
            if (s.th[k] <= 0.0f && s.th[k + 1u] > 0.0f) {
                s.th[k] = s.th[k + 1u];
                s.cu[k] = s.cu[k + 1u];
                copy16(s.pg[k], s.pg[k + 1u]);
                s.th[k + 1u] = 0.0f;
                s.cu[k + 1u] = 0.0f;
            }
        }
    }
    float tiny = F.P.misc[0];
#pragma unroll
    for (uint32_t k = 0; k < 3u; k++) {
        if (s.th[k] > 0.0f && s.th[k] < tiny && s.th[k + 1u] > 0.0f) stack_merge(s, k);
    }
}

// strata.wgsl:105-131
__device__ inline void stack_push_top(const FilmDev& F, Stack& s, float vol, const float* pig, float cure) {
    if (vol <= 0.0f) return;
    uint32_t n = stack_count(s);
    bool done = false;
#pragma unroll
    for (uint32_t t = 0; t < 4u; t++) {  // the top stratum t = n - 1
        if (t + 1u != n) continue;
        bool close = pig_dist(s.pg[t], pig) < F.P.misc[1] || s.th[t] < F.P.misc[0];
        if (close) {
            pig_mix(s.pg[t], s.th[t], pig, vol, s.pg[t]);
            s.cu[t] = (s.cu[t] * s.th[t] + cure * vol) / (s.th[t] + vol);
            s.th[t] = s.th[t] + vol;
            done = true;
        }
    }
    if (done) return;
    if (n == 4u) {
        uint32_t best = 0;
        float bd = 1e30f;
#pragma unroll
        for (uint32_t k = 0; k < 3u; k++) {
            float d = pig_dist(s.pg[k], s.pg[k + 1u]);
            if (d < bd) { bd = d; best = k; }
        }
        stack_merge(s, best);
    }
    const uint32_t top = stack_count(s);
#pragma unroll
    for (uint32_t t = 0; t < 4u; t++) {
        if (t != top) continue;
        s.th[t] = vol;
        copy16(s.pg[t], pig);
        s.cu[t] = cure;
    }
}

// strata.wgsl:133-155
__device__ inline void stack_add_band(const FilmDev& F, Stack& s, float zlo, float zhi, float vol, const float* pig,
                                      float cure) {
    if (vol <= 0.0f) return;
    float span = cz_max(zhi - zlo, 1e-12f);
    float z0 = 0.0f;
    float above = vol;
    bool ended = false;
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++) {
        float tk = s.th[k];
        if (ended || tk <= 0.0f) {
            ended = true;
            continue;
        }
        float z1 = z0 + tk;
        float ov = cz_max(cz_min(zhi, z1) - cz_max(zlo, z0), 0.0f);
        float part = vol * ov / span;
        if (part > 0.0f) {
            pig_mix(s.pg[k], tk, pig, part, s.pg[k]);
            s.cu[k] = (s.cu[k] * tk + cure * part) / (tk + part);
            s.th[k] = tk + part;
            above = above - part;
        }
        z0 = z1;
    }
    if (above > 1e-14f) stack_push_top(F, s, above, pig, cure);
}

// ---------------------------------------------------------------------------
// km.wgsl:1-33

__device__ __forceinline__ void km_layer(float k, float s, float t, float& R, float& T) {
    R = 0.0f;
    T = 1.0f;
    if (k + s < 1e-6f || t <= 0.0f) return;
    float bs = sqrtf(cz_max(k * (k + 2.0f * s), 0.0f));
    float arg = cz_min(bs * t, 20.0f);
    float e = cz_expf(arg);
    float sh = 0.5f * (e - 1.0f / e);
    float ch = 0.5f * (e + 1.0f / e);
    float den = (k + s) * sh + bs * ch;
    if (den < 1e-12f) return;
    R = s * sh / den;
    T = bs / den;
}
__device__ __forceinline__ float composite(float rtx, float rty, float rb) {
    return rtx + rty * rty * rb / cz_max(1.0f - rtx * rb, 1e-4f);
}
__device__ __forceinline__ void layer_ks(const float* pigtab, const float* pig, uint32_t band, float& ko, float& so) {
    float k = 0.0f, s = 0.0f, amount = 0.0f, weight = 0.0f;
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)NPIG; i++) {
        float c = pig[i];
        if (c > 0.0f) {
            float w = c * pigtab[i * PIG_STRIDE + 34u];
            k = k + w * pigtab[i * PIG_STRIDE + band];
            s = s + w * pigtab[i * PIG_STRIDE + NBAND + band];
            amount = amount + c;
            weight = weight + w;
        }
    }
    float sc = amount / cz_max(weight, 1e-9f);
    ko = k * sc;
    so = s * sc;
}

}  // namespace cz
