// tools.hip - the tool passes of the film step (gfx950): the brush scrape
// over uploaded contact segments (shaders/brush.wgsl) with its per-bristle
// reservoir, and the palette knife (shaders/knife.wgsl) with its scrap list;
// src/film/mod.rs:1030-1173.
//
// Scrap order: the reference appends scraps with an atomic counter, in
// whatever order cells run.  Here the knife runs as three launches over the
// cells it can reach (a box around the blade's swept parallelogram, 256
// cells of one row per block, blocks in row order): a dry pass counts each
// block's scraps, one block turns the counts into offsets, and the wet pass
// recomputes each cell, writes it and places its scraps.  The list is then
// in ascending cell order, the paint a cell gives up before the paint it
// receives, frame after frame; past CZ_MAX_SCRAPS the later ones are dropped.
#include "film_internal.h"

namespace {

constexpr float RES_UNIT = 1e15f;
constexpr float LEAVE_FILM = 8e-5f;
constexpr float BRISTLE_SCRAPE = 0.25f;
constexpr float SPONGE_FILM = 1e-5f;
constexpr uint32_t KNIFE_BANDS = 8u;
constexpr float KNIFE_BLADE_LEN = 0.07f;
constexpr float KNIFE_SLOPE = 0.25f;
constexpr float KNIFE_FLAT = 0.01f;
constexpr float KNIFE_JOURNAL_REACH = 0.08f;
constexpr int KNIFE_BOX_MARGIN = 8;  // cells around the exact box (rounding of t and u)

struct KnifeDev {
    cz_knife k;
    uint32_t x0, y0, bw, bh;  // the cells the knife passes visit
};

// f32 -> u32, saturating (NaN gives 0)
__device__ __forceinline__ uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

// brush.wgsl:23-31
__device__ __forceinline__ float seg_dist(float px, float py, const cz_contact& k) {
    const float abx = k.p1[0] - k.p0[0], aby = k.p1[1] - k.p0[1];
    const float l2 = abx * abx + aby * aby;
    float t = 0.0f;
    if (l2 > 0.0f) t = cz_clamp(((px - k.p0[0]) * abx + (py - k.p0[1]) * aby) / l2, 0.0f, 1.0f);
    return cz_len2(px - (k.p0[0] + abx * t), py - (k.p0[1] + aby * t));
}

// lay_pack(method, normalize(v + (1e-9, 0))) (common.wgsl:81-84)
__device__ __forceinline__ uint32_t lay_dir(uint32_t method, float sx, float sy) {
    const float vx = sx + 1e-9f, vy = sy + 0.0f;
    const float vl = cz_len2(vx, vy);
    const float turn = cz_atan2f(vy / vl, vx / vl) / 6.2831853f + 0.5f;
    return (method & 0xffu) | ((uint32_t)(cz_clamp(turn, 0.0f, 1.0f) * 65535.0f) << 8u);
}

// Paint off the top of the stack, strata 3 down to 0, until `excess` is
// taken (brush.wgsl:93-107, knife.wgsl:89-103); its pigment mix in pig.
__device__ __forceinline__ float strip_top(Stack& s, float excess, float* pig) {
    float removed = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) pig[i] = 0.0f;
#pragma unroll
    for (int k = 3; k >= 0; k--) {
        if (excess > 0.0f) {
            const float th = s.th[k];
            if (th > 0.0f) {
                const float take = cz_min(th, excess);
                pig_mix(pig, removed, s.pg[k], take, pig);
                removed += take;
                excess -= take;
                s.th[k] = th - take;
            }
        }
    }
    return removed;
}

constexpr uint32_t kResWords = (uint32_t)CZ_MAX_BRISTLES * 17u;

// brush_scrape (brush.wgsl:33-122): over the cells of 64-tiles with
// contacts; the nearest-cap contact sets the cap, the laying direction and
// the bristle whose reservoir receives the paint taken.  The block's cells
// add their amounts into an LDS copy of the reservoir first; the block then
// adds the words it touched to the device reservoir (integer sums, so the
// totals do not depend on the order).
__device__ __forceinline__ void brush_cell(const FilmDev& F, const cz_contact* contacts, uint2 r,
                                           const uint32_t* items, uint32_t* s_res, uint32_t i, uint32_t j);

__global__ __launch_bounds__(256) void k_brush(FilmDev F, const cz_contact* contacts, const uint2* range,
                                               const uint32_t* items, uint32_t* reservoir) {
    __shared__ uint32_t s_res[kResWords];
    const uint32_t i = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t j = blockIdx.y * 16u + (threadIdx.x >> 4);
    // a 16x16 block lies inside one 64-tile: the same contact range for all
    // of its threads (i, j may pass the edge; the 64-tile of the block's
    // first cell is the one)
    const uint32_t bi = blockIdx.x * 16u, bj = blockIdx.y * 16u;
    const uint2 r = range[(bi / TILE64) + (bj / TILE64) * F.P.tiles[0]];
    if (r.y == 0u) return;
    for (uint32_t q = threadIdx.x; q < kResWords; q += 256u) s_res[q] = 0u;
    __syncthreads();
    if (i < F.W && j < F.H) brush_cell(F, contacts, r, items, s_res, i, j);
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < kResWords; q += 256u)
        if (s_res[q] != 0u) atomicAdd(reservoir + q, s_res[q]);
}

__device__ __forceinline__ void brush_cell(const FilmDev& F, const cz_contact* contacts, uint2 r,
                                           const uint32_t* items, uint32_t* s_res, uint32_t i, uint32_t j) {
    const uint32_t c = idx(F, i, j);
    const float hv = F.h[c];
    if (hv <= SPONGE_FILM) return;
    const float px = ((float)i + 0.5f) * F.P.dx, py = ((float)j + 0.5f) * F.P.dx;
    float cap = 1e9f;
    uint32_t who = 0xffffffffu, method = CZ_METHOD_BRUSH;
    float sx = 1.0f, sy = 0.0f, tx = 0.0f, ty = 0.0f;
    bool swept = false;
    for (uint32_t q = 0; q < r.y; q++) {
        const cz_contact k = contacts[items[r.x + q]];
        if (!(seg_dist(px, py, k) < k.r)) continue;
        if (k.pad == 0u) {
            swept = true;
            tx = tx + (k.p1[0] - k.p0[0]);
            ty = ty + (k.p1[1] - k.p0[1]);
        }
        float cc = cz_max(k.zb - F.m[c], LEAVE_FILM);
        if (k.pad == CZ_CONTACT_ROLLER) cc = cz_max(k.zb, LEAVE_FILM);
        if (k.pad == CZ_CONTACT_SPONGE) cc = cz_max(k.zb, SPONGE_FILM);
        if (cc < cap) {
            cap = cc;
            who = k.bristle;
            if (k.pad == CZ_CONTACT_ROLLER) {
                method = CZ_METHOD_ROLLER;
                sx = k.p1[1] - k.p0[1];
                sy = k.p0[0] - k.p1[0];
            } else {
                method = k.pad == CZ_CONTACT_SPONGE ? CZ_METHOD_ROLLER : CZ_METHOD_BRUSH;
                sx = k.p1[0] - k.p0[0];
                sy = k.p1[1] - k.p0[1];
            }
        }
    }
    if (who == 0xffffffffu) return;
    if (swept && method == CZ_METHOD_BRUSH) {
        sx = tx;
        sy = ty;
    }
    F.lay[c] = lay_dir(method, sx, sy);
    if (hv <= cap) return;
    Stack s;
    stack_load(F, c, s);
    float excess = hv - cap;
    if (method == CZ_METHOD_BRUSH) excess *= BRISTLE_SCRAPE;
    float pig[16];
    const float removed = strip_top(s, excess, pig);
    if (removed <= 0.0f) return;
    F.h[c] = hv - removed;
    F.tile_moved[(i >> 4) + (j >> 4) * F.tw] = 1u;
    stack_compact(F, s);
    stack_store_to(F.strata_h, F.strata_cure, F.strata_pig, c, s);
    if (who >= (uint32_t)CZ_MAX_BRISTLES) return;  // no reservoir for it
    const float vol = removed * F.P.dx * F.P.dx;
    uint32_t* res = s_res + who * 17u;
    atomicAdd(res, f2u_sat(vol * RES_UNIT + 0.5f));
#pragma unroll
    for (int b = 0; b < 16; b++) {
        const float sc = b >= 13 ? 0.01f : 1.0f;
        atomicAdd(res + 1 + b, f2u_sat(cz_max(vol * pig[b] * sc, 0.0f) * RES_UNIT + 0.5f));
    }
}

// knife.wgsl:34-40
__device__ __forceinline__ float cell_hash(uint32_t x, uint32_t y) {
    uint32_t h = (x * 0x8da6b343u) ^ (y * 0xd8163841u);
    h = (h ^ (h >> 15u)) * 0x2c1b3c6du;
    h = (h ^ (h >> 12u)) * 0x297a2d39u;
    h = h ^ (h >> 15u);
    return (float)(h & 0xffffu) / 65535.0f;
}

// knife (knife.wgsl:56-128) on one cell.  Returns its scraps as bits (1:
// paint taken, 2: paint laid) with their volumes and the taken mix; WRITE
// stores the cell, otherwise nothing is written.
template <bool WRITE>
__device__ __forceinline__ uint32_t knife_cell(const FilmDev& F, const cz_knife& K, uint32_t i, uint32_t j,
                                               float* pig, float& removed, float& add) {
    const uint32_t c = idx(F, i, j);
    const float dxm = F.P.dx;
    const float px = ((float)i + 0.5f) * dxm, py = ((float)j + 0.5f) * dxm;
    const float abx = K.b[0] - K.a[0], aby = K.b[1] - K.a[1];
    const float mvx = K.dir[0] * K.sweep, mvy = K.dir[1] * K.sweep;
    const float det = abx * mvy - aby * mvx;
    if (cz_abs(det) < 1e-12f) return 0u;
    const float qx = px - K.a[0], qy = py - K.a[1];
    const float t = (qx * mvy - qy * mvx) / det;
    const float u = -(abx * qy - aby * qx) / det;
    if (t < 0.0f || t > 1.0f) return 0u;
    const float behind = u * K.sweep;
    if (behind < -dxm || behind > KNIFE_BLADE_LEN) return 0u;
    const bool at_edge = u <= 1.0f;
    uint32_t nb = K.bands < KNIFE_BANDS ? K.bands : KNIFE_BANDS;
    if (nb < 1u) nb = 1u;
    uint32_t bi = f2u_sat(t * (float)nb);
    if (bi > nb - 1u) bi = nb - 1u;
    const float ratio = at_edge ? K.fill[bi] : 0.0f;
    const bool covered = cell_hash(i / 2u, j / 2u) < cz_powf(ratio, 0.6f);
    const float level_b = covered ? K.height * cz_min(1.0f, 1.5f * ratio) : 0.0f;
    const float smear_b = at_edge ? K.smear[bi] : 0.0f;
    const float hv = F.h[c];
    const float cap = cz_max(K.height + cz_max(behind - KNIFE_FLAT, 0.0f) * KNIFE_SLOPE - F.m[c], 0.0f);
    const float level = cz_min(level_b, cap);
    const float keep = cz_min(hv, cap);
    const float swap = keep * smear_b;
    if (hv <= cap && hv >= level && swap <= 0.0f) return 0u;
    Stack s;
    stack_load(F, c, s);
    removed = strip_top(s, (hv - cap) + swap, pig);
    float hn = hv - removed;
    add = cz_max(cz_max(level, keep) - hn, 0.0f);
    if (WRITE) {
        if (add > 0.0f) {
            float band_pig[16];
            const uint32_t* bp = K.pig[bi];
            unpack8(make_uint4(bp[0], bp[1], bp[2], bp[3]), make_uint4(bp[4], bp[5], bp[6], bp[7]), band_pig);
            stack_push_top(F, s, add, band_pig, 0.0f);
            hn += add;
        }
        F.h[c] = hn;
        F.lay[c] = lay_dir(CZ_METHOD_KNIFE, mvx, mvy);
        F.tile_moved[(i >> 4) + (j >> 4) * F.tw] = 1u;
        stack_compact(F, s);
        stack_store_to(F.strata_h, F.strata_cure, F.strata_pig, c, s);
    }
    return (removed > 0.0f ? 1u : 0u) | (add > 0.0f ? 2u : 0u);
}

// Dry pass: scraps per block (256 cells of one box row).  Block 0 also
// keeps the list length the frame starts from.
__global__ __launch_bounds__(256) void k_knife_count(FilmDev F, const KnifeDev* KD, uint32_t* kcnt, uint32_t* kctl) {
    const uint32_t gi = blockIdx.x * 256u + threadIdx.x;
    uint32_t e = 0u;
    if (gi < KD->bw) {
        float pig[16], removed = 0.0f, add = 0.0f;
        e = knife_cell<false>(F, KD->k, KD->x0 + gi, KD->y0 + blockIdx.y, pig, removed, add);
    }
    const uint32_t n = (uint32_t)__syncthreads_count((e & 1u) ? 1 : 0) + (uint32_t)__syncthreads_count((e & 2u) ? 1 : 0);
    if (threadIdx.x == 0u) kcnt[blockIdx.y * gridDim.x + blockIdx.x] = n;
    if (blockIdx.x == 0u && blockIdx.y == 0u && threadIdx.x == 0u) kctl[1] = kctl[0];
}

// The per-block counts to exclusive offsets, in place; the total to kctl[2].
__global__ __launch_bounds__(1024) void k_knife_scan(uint32_t* kcnt, uint32_t nblocks, uint32_t* kctl) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0u) s_carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < nblocks; base += 1024u) {
        const uint32_t q = base + threadIdx.x;
        const uint32_t v = q < nblocks ? kcnt[q] : 0u;
        uint32_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off) incl += o;
        }
        if (lane == 63u) s_w[wave] = incl;
        __syncthreads();
        uint32_t before = s_carry;
        for (uint32_t w = 0; w < wave; w++) before += s_w[w];
        if (q < nblocks) kcnt[q] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023u) s_carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0u) kctl[2] = s_carry;
}

// Wet pass: each cell again, written, its scraps placed after the block's
// offset in thread order.
__global__ __launch_bounds__(256) void k_knife_apply(FilmDev F, const KnifeDev* KD, const uint32_t* koff,
                                                     uint32_t* kctl, cz_scrap* scraps) {
    __shared__ uint32_t s_w[4];
    const uint32_t gi = blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t i = KD->x0 + gi, j = KD->y0 + blockIdx.y;
    uint32_t e = 0u;
    float pig[16], removed = 0.0f, add = 0.0f;
    if (gi < KD->bw) e = knife_cell<true>(F, KD->k, i, j, pig, removed, add);
    const unsigned long long br = __ballot((e & 1u) != 0u), ba = __ballot((e & 2u) != 0u);
    const unsigned long long lmask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    if (lane == 0u) s_w[wave] = (uint32_t)(__popcll(br) + __popcll(ba));
    __syncthreads();
    uint32_t slot = kctl[1] + koff[blockIdx.y * gridDim.x + blockIdx.x] + (uint32_t)(__popcll(br & lmask) + __popcll(ba & lmask));
    for (uint32_t w = 0; w < wave; w++) slot += s_w[w];
    const uint32_t c = idx(F, i, j);
    const float dxm = F.P.dx;
    if (e & 1u) {
        if (slot < CZ_MAX_SCRAPS) {
            uint4* r = reinterpret_cast<uint4*>(scraps + slot);
            uint4 p0, p1;
            pack8(pig, p0, p1);
            r[0] = make_uint4(c, __float_as_uint(removed * dxm * dxm), KD->k.frame, 0u);
            r[1] = p0;
            r[2] = p1;
        }
        slot++;
    }
    if ((e & 2u) && slot < CZ_MAX_SCRAPS) {
        uint4* r = reinterpret_cast<uint4*>(scraps + slot);
        const float none[16] = {};
        uint4 p0, p1;
        pack8(none, p0, p1);
        r[0] = make_uint4(c, __float_as_uint(-add * dxm * dxm), KD->k.frame, 0u);
        r[1] = p0;
        r[2] = p1;
    }
    // no block of this launch reads kctl[0]
    if (blockIdx.x == 0u && blockIdx.y == 0u && threadIdx.x == 0u) kctl[0] = kctl[1] + kctl[2];
}

template <typename T>
int talloc(cz_film* f, T** p, size_t count) {
    void* v = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess) return set_err(CZ_ENOMEM, "tools: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    f->allocs.push_back(v);
    f->bytes += bytes;
    *p = (T*)v;
    HIP_TRY(hipMemsetAsync(v, 0, bytes, f->stream));
    return CZ_OK;
}

int tools_alloc(cz_film* f) {
    if (f->tools_ready) return CZ_OK;
    // knife blocks: ceil(W / 256) * H <= n / 256 + H, and H <= 16 * tiles16
    const uint32_t kcap = f->cap_n / 256u + 16u * f->cap_nt16 + 1u;
    int rc = CZ_OK;
    KnifeDev* kd = nullptr;
    if (!rc && !f->contacts) rc = talloc(f, &f->contacts, CZ_MAX_CONTACTS);
    if (!rc && !f->ctile_range) rc = talloc(f, &f->ctile_range, f->cap_nt64);
    if (!rc && !f->ctile_items) rc = talloc(f, &f->ctile_items, (size_t)CZ_MAX_CONTACTS * 16u);
    if (!rc && !f->reservoir) rc = talloc(f, &f->reservoir, (size_t)CZ_MAX_BRISTLES * 17u);
    if (!rc && !f->knife_dev) {
        rc = talloc(f, &kd, 1);
        if (!rc) f->knife_dev = kd;  // kept at once, so a retry after a later failure reuses it
    }
    if (!rc && !f->scraps) rc = talloc(f, &f->scraps, CZ_MAX_SCRAPS);
    if (!rc && !f->kcnt) rc = talloc(f, &f->kcnt, kcap);
    if (!rc && !f->kctl) rc = talloc(f, &f->kctl, 4);
    if (!rc) {
        hipError_t e = hipSuccess;
        if (!f->contacts_h) e = hipHostMalloc((void**)&f->contacts_h, CZ_MAX_CONTACTS * sizeof(cz_contact), hipHostMallocDefault);
        if (e == hipSuccess && !f->crange_h) e = hipHostMalloc((void**)&f->crange_h, f->cap_nt64 * sizeof(uint2), hipHostMallocDefault);
        if (e == hipSuccess && !f->citems_h) e = hipHostMalloc((void**)&f->citems_h, (size_t)CZ_MAX_CONTACTS * 16u * 4u, hipHostMallocDefault);
        if (e == hipSuccess && !f->knife_h) e = hipHostMalloc(&f->knife_h, sizeof(KnifeDev), hipHostMallocDefault);
        if (e != hipSuccess) rc = set_err(CZ_ENOMEM, "tools: hipHostMalloc: %s", hipGetErrorString(e));
    }
    if (rc) return rc;  // what was allocated stays with the film and is freed with it
    f->kcap = kcap;
    f->tools_ready = true;
    return czi::wait_idle(f);
}

// The cells the knife can reach: a + t*(b - a) - behind*dir for t in [0, 1]
// and behind in [-dx, KNIFE_BLADE_LEN] (knife.wgsl:66-72: q = t*(b - a) -
// u*mv with mv = dir*sweep and behind = u*sweep, so behind runs against the
// sweep), as a box widened by a margin for the rounding of t and u.  A
// nearly degenerate blade/sweep pair (where that rounding is unbounded)
// takes the whole canvas.  false: no cell can pass.
bool knife_box(const cz_film* f, const cz_knife& K, uint32_t box[4]) {
    const float abx = K.b[0] - K.a[0], aby = K.b[1] - K.a[1];
    const float mvx = K.dir[0] * K.sweep, mvy = K.dir[1] * K.sweep;
    const float det = abx * mvy - aby * mvx;  // as the kernel computes it
    if (!(fabsf(det) >= 1e-12f)) return false;
    const double dx = f->dx;
    const double lab = std::hypot((double)abx, (double)aby), lmv = std::hypot((double)mvx, (double)mvy);
    double lo_x = 1e300, hi_x = -1e300, lo_y = 1e300, hi_y = -1e300;
    for (int ti = 0; ti < 2; ti++)
        for (int bi = 0; bi < 2; bi++) {
            const double b = bi ? (double)KNIFE_BLADE_LEN : -dx;
            const double x = (double)K.a[0] + ti * (double)abx - b * (double)K.dir[0];
            const double y = (double)K.a[1] + ti * (double)aby - b * (double)K.dir[1];
            lo_x = std::min(lo_x, x); hi_x = std::max(hi_x, x);
            lo_y = std::min(lo_y, y); hi_y = std::max(hi_y, y);
        }
    const bool sane = std::isfinite(lo_x) && std::isfinite(hi_x) && std::isfinite(lo_y) && std::isfinite(hi_y) &&
                      std::fabs((double)det) >= 1e-2 * lab * lmv;
    double x0 = 0.0, x1 = (double)f->W - 1.0, y0 = 0.0, y1 = (double)f->H - 1.0;
    if (sane) {
        x0 = std::max(x0, std::floor(lo_x / dx) - KNIFE_BOX_MARGIN);
        x1 = std::min(x1, std::ceil(hi_x / dx) + KNIFE_BOX_MARGIN);
        y0 = std::max(y0, std::floor(lo_y / dx) - KNIFE_BOX_MARGIN);
        y1 = std::min(y1, std::ceil(hi_y / dx) + KNIFE_BOX_MARGIN);
    }
    if (x1 < x0 || y1 < y0) return false;
    box[0] = (uint32_t)x0;
    box[1] = (uint32_t)y0;
    box[2] = (uint32_t)(x1 - x0) + 1u;
    box[3] = (uint32_t)(y1 - y0) + 1u;
    return true;
}

}  // namespace

namespace czi {

void tools_record_brush(cz_film* f) {
    const dim3 gcell((f->W + 15u) / 16u, (f->H + 15u) / 16u);
    hipLaunchKernelGGL(k_brush, gcell, dim3(256), 0, f->stream, f->dev(), f->contacts, f->ctile_range, f->ctile_items,
                       f->reservoir);
}

void tools_record_knife(cz_film* f) {
    const FilmDev F = f->dev();
    const KnifeDev* kd = (const KnifeDev*)f->knife_dev;
    const dim3 g(f->knife_grid[0], f->knife_grid[1]);
    hipLaunchKernelGGL(k_knife_count, g, dim3(256), 0, f->stream, F, kd, f->kcnt, f->kctl);
    hipLaunchKernelGGL(k_knife_scan, dim3(1), dim3(1024), 0, f->stream, f->kcnt, g.x * g.y, f->kctl);
    hipLaunchKernelGGL(k_knife_apply, g, dim3(256), 0, f->stream, F, kd, f->kcnt, f->kctl, f->scraps);
}

void tools_free_host(cz_film* f) {
    if (f->contacts_h) hipHostFree(f->contacts_h);
    if (f->crange_h) hipHostFree(f->crange_h);
    if (f->citems_h) hipHostFree(f->citems_h);
    if (f->knife_h) hipHostFree(f->knife_h);
    f->contacts_h = nullptr;
    f->crange_h = nullptr;
    f->citems_h = nullptr;
    f->knife_h = nullptr;
}

int tools_reset(cz_film* f) {
    f->contact_count = 0u;
    f->knife_active = false;
    f->knife_grid[0] = f->knife_grid[1] = 0u;
    if (!f->tools_ready) return CZ_OK;
    HIP_TRY(hipMemsetAsync(f->reservoir, 0, (size_t)CZ_MAX_BRISTLES * 17u * 4u, f->stream));
    HIP_TRY(hipMemsetAsync(f->kctl, 0, 16u, f->stream));
    return CZ_OK;
}

}  // namespace czi

using namespace czi;

extern "C" {

// upload_contacts (src/film/mod.rs:1044-1084)
int cz_film_upload_contacts(cz_film* f, const cz_contact* cs, uint32_t n) {
    if (!f || (n > 0 && !cs)) return set_err(CZ_EINVAL, "upload_contacts: bad arguments");
    if (f->batch) return set_err(CZ_EINVAL, "upload_contacts: the film is stepped by a batch (cz_batch_remove it first)");
    const uint32_t count = std::min<uint32_t>(n, CZ_MAX_CONTACTS);
    f->contact_count = 0u;
    if (count == 0u) return CZ_OK;
    int rc = tools_alloc(f);
    if (rc) return rc;
    rc = wait_idle(f);  // the pinned staging may still feed the last frame's copies
    if (rc) return rc;
    const uint32_t nt = f->t64x * f->t64y;
    const float tile_m = 64.0f * f->dx;
    std::vector<uint32_t> cnt(nt, 0u);
    std::vector<int32_t> rect((size_t)count * 4, -1);
    for (uint32_t ci = 0; ci < count; ci++) {
        const cz_contact& c = cs[ci];
        const float minx = fminf(c.p0[0], c.p1[0]) - c.r, maxx = fmaxf(c.p0[0], c.p1[0]) + c.r;
        const float miny = fminf(c.p0[1], c.p1[1]) - c.r, maxy = fmaxf(c.p0[1], c.p1[1]) + c.r;
        journal_touch_rect(f, minx, miny, maxx, maxy);
        if (maxx < 0.0f || maxy < 0.0f || minx > f->s.width_m || miny > f->s.height_m ||
            !(std::isfinite(minx) && std::isfinite(miny)))
            continue;
        const float mx = (float)(f->t64x - 1u), my = (float)(f->t64y - 1u);
        // ((v / tile_m).floor().max(0.0) as i64).min(n - 1): the cast saturates
        rect[4 * ci] = (int32_t)fminf(fmaxf(floorf(minx / tile_m), 0.0f), mx);
        rect[4 * ci + 1] = (int32_t)fminf(fmaxf(floorf(maxx / tile_m), 0.0f), mx);
        rect[4 * ci + 2] = (int32_t)fminf(fmaxf(floorf(miny / tile_m), 0.0f), my);
        rect[4 * ci + 3] = (int32_t)fminf(fmaxf(floorf(maxy / tile_m), 0.0f), my);
        for (int32_t ty = rect[4 * ci + 2]; ty <= rect[4 * ci + 3]; ty++)
            for (int32_t tx = rect[4 * ci]; tx <= rect[4 * ci + 1]; tx++) cnt[(size_t)ty * f->t64x + (size_t)tx]++;
    }
    const uint32_t cap = CZ_MAX_CONTACTS * 16u;
    uint32_t items = 0;
    for (uint32_t t = 0; t < nt; t++) {
        const uint32_t take = std::min(cnt[t], cap - items);
        f->crange_h[t] = make_uint2(items, take);
        items += take;
    }
    std::fill(cnt.begin(), cnt.end(), 0u);
    for (uint32_t ci = 0; ci < count; ci++) {
        if (rect[4 * ci] < 0) continue;
        for (int32_t ty = rect[4 * ci + 2]; ty <= rect[4 * ci + 3]; ty++)
            for (int32_t tx = rect[4 * ci]; tx <= rect[4 * ci + 1]; tx++) {
                const size_t t = (size_t)ty * f->t64x + (size_t)tx;
                if (cnt[t] < f->crange_h[t].y) f->citems_h[f->crange_h[t].x + cnt[t]++] = ci;
            }
    }
    memcpy(f->contacts_h, cs, (size_t)count * sizeof(cz_contact));
    HIP_TRY(hipMemcpyAsync(f->contacts, f->contacts_h, (size_t)count * sizeof(cz_contact), hipMemcpyHostToDevice, f->stream));
    HIP_TRY(hipMemcpyAsync(f->ctile_range, f->crange_h, (size_t)nt * sizeof(uint2), hipMemcpyHostToDevice, f->stream));
    if (items > 0)
        HIP_TRY(hipMemcpyAsync(f->ctile_items, f->citems_h, (size_t)items * 4u, hipMemcpyHostToDevice, f->stream));
    f->contact_count = count;
    return CZ_OK;
}

// set_knife (src/film/mod.rs:1030-1042)
int cz_film_set_knife(cz_film* f, const cz_knife* k) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    if (f->batch) return set_err(CZ_EINVAL, "set_knife: the film is stepped by a batch (cz_batch_remove it first)");
    f->knife_active = false;
    f->knife_grid[0] = f->knife_grid[1] = 0u;
    if (!k || k->active == 0u) return CZ_OK;
    int rc = tools_alloc(f);
    if (rc) return rc;
    const float reach = KNIFE_JOURNAL_REACH + k->sweep;
    journal_touch_rect(f, fminf(k->a[0], k->b[0]) - reach, fminf(k->a[1], k->b[1]) - reach,
                       fmaxf(k->a[0], k->b[0]) + reach, fmaxf(k->a[1], k->b[1]) + reach);
    uint32_t box[4];
    f->knife_active = true;
    if (!knife_box(f, *k, box)) return CZ_OK;  // active, but no cell can pass its tests
    rc = wait_idle(f);  // the pinned staging may still feed the last copy
    if (rc) return rc;
    KnifeDev* kh = (KnifeDev*)f->knife_h;
    kh->k = *k;
    kh->x0 = box[0];
    kh->y0 = box[1];
    kh->bw = box[2];
    kh->bh = box[3];
    const uint32_t gx = (box[2] + 255u) / 256u, gy = box[3];
    if ((uint64_t)gx * gy > f->kcap) {
        f->knife_active = false;
        return set_err(CZ_EINVAL, "set_knife: %u x %u blocks exceed %u", gx, gy, f->kcap);
    }
    HIP_TRY(hipMemcpyAsync(f->knife_dev, kh, sizeof(KnifeDev), hipMemcpyHostToDevice, f->stream));
    f->knife_grid[0] = gx;
    f->knife_grid[1] = gy;
    return CZ_OK;
}

int cz_film_take_reservoir(cz_film* f, uint32_t* out) {
    if (!f || !out) return set_err(CZ_EINVAL, "null argument");
    const size_t bytes = (size_t)CZ_MAX_BRISTLES * 17u * 4u;
    if (!f->tools_ready) {
        memset(out, 0, bytes);
        return CZ_OK;
    }
    HIP_TRY(hipMemcpyAsync(out, f->reservoir, bytes, hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipMemsetAsync(f->reservoir, 0, bytes, f->stream));
    return wait_idle(f);
}

int cz_film_take_scraps(cz_film* f, cz_scrap* out, uint32_t cap, uint32_t* count) {
    if (!f || !count || (cap > 0 && !out)) return set_err(CZ_EINVAL, "null argument");
    *count = 0u;
    if (!f->tools_ready) return CZ_OK;
    uint32_t n = 0;
    HIP_TRY(hipMemcpyAsync(&n, f->kctl, 4, hipMemcpyDeviceToHost, f->stream));
    int rc = wait_idle(f);
    if (rc) return rc;
    n = std::min<uint32_t>(n, CZ_MAX_SCRAPS);
    const uint32_t take = std::min(n, cap);
    if (take > 0) HIP_TRY(hipMemcpyAsync(out, f->scraps, (size_t)take * sizeof(cz_scrap), hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipMemsetAsync(f->kctl, 0, 4, f->stream));
    rc = wait_idle(f);
    if (rc) return rc;
    *count = n;
    return CZ_OK;
}

}  // extern "C"
