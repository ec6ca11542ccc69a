// film_kernels.hip - gfx950 kernels of the thin-film paint step.
//
// One kernel per reference dispatch (or a fused group), cited per kernel.
// Kernels over the active tiles walk the ordered tile list with a grid-stride
// loop: a block of 256 threads (4 wave64) takes one 16x16 tile per
// iteration, so the launch size never depends on the device-side count and
// no host round trip is needed inside a frame.
//
// Where the reference reads, inside one dispatch, a neighbour value that the
// same dispatch writes, these kernels read the value from before the
// dispatch: nose writes a second offset buffer, the strata fill passes carry
// the stack sums in ping-pong buffers, and cg_finish leaves the new height in
// a scratch buffer until the rescale pass.
#include <type_traits>

#include "film_kernels.h"

namespace cz {

// Binds, per item of the launch: _a (item slot), F (the item's canvas) and
// _t (its tile).
// Which list items a block takes.  Workgroups are dealt to the 8 XCDs in
// turn (block b on XCD b % 8), each XCD with its own L2.  For the tile lists
// (active, cure, batched), the list (ascending tiles) is cut into 8
// contiguous chunks, one per XCD, so neighbouring tiles share an L2 and the
// halos of their LDS windows are fetched once.  Passes over every tile of a
// canvas, whose per-tile work follows where the paint is, keep the plain
// stride, which spreads that work over all XCDs.  Tiles are independent
// within a launch, so any assignment gives the same results; a grid that is
// not a multiple of 8 keeps the plain stride.
struct XcdRange {
    uint32_t first, end, step;
};
template <bool CHUNKS>
__device__ __forceinline__ XcdRange xcd_range(uint32_t n) {
    constexpr uint32_t NX = 8u;
    if (!CHUNKS || gridDim.x < NX || (gridDim.x & (NX - 1u)) != 0u) return {blockIdx.x, n, gridDim.x};
    const uint32_t chunk = (n + NX - 1u) / NX;
    const uint32_t lo = (blockIdx.x % NX) * chunk;
    return {lo + blockIdx.x / NX, min(n, lo + chunk), gridDim.x / NX};
}

#define CZ_FOR_TILES(WS)                                                             \
    const uint32_t _nitems = (WS).count();                                           \
    const uint32_t _lx = threadIdx.x & 15u, _ly = threadIdx.x >> 4;                  \
    const XcdRange _rg = xcd_range<std::decay_t<decltype(WS)>::kXcdChunks>(_nitems); \
    for (uint32_t _a = _rg.first; _a < _rg.end; _a += _rg.step)                      \
        if (const uint32_t _item = (WS).item(_a); true)                              \
            if (const FilmDev& F = (WS).film(_item); true)                           \
                if (const uint32_t _t = (WS).tile(F, _item); true)

#define CZ_TILE_CELL(F)                                                              \
    const uint32_t x = (_t % (F).tw) * 16u + _lx, y = (_t / (F).tw) * 16u + _ly;     \
    const bool in_grid = x < (F).W && y < (F).H;                                     \
    const uint32_t c = in_grid ? idx((F), x, y) : 0u;

// ---------------------------------------------------------------------------
// deposit.wgsl:6-29
__device__ inline float ribbon_height(float px, float py, const cz_deposit& d) {
    float abx = d.p1[0] - d.p0[0], aby = d.p1[1] - d.p0[1];
    float l2 = abx * abx + aby * aby;
    if (l2 <= 0.0f) {
        float xx = cz_len2(px - d.p0[0], py - d.p0[1]) / d.r0;
        return (xx < 1.0f) ? d.h0 * (1.0f - xx * xx) : 0.0f;
    }
    float l = sqrtf(l2);
    float dirx = abx / l, diry = aby / l;
    float q0x = px - d.p0[0], q0y = py - d.p0[1];
    float q1x = px - d.p1[0], q1y = py - d.p1[1];
    bool cap_start = (d.kind & CZ_DEPOSIT_CAP_START) != 0u;
    bool cap_end = (d.kind & CZ_DEPOSIT_CAP_END) != 0u;
    if (!cap_start && q0x * d.n0[0] + q0y * d.n0[1] < 0.0f) return 0.0f;
    if (!cap_end && q1x * d.n1[0] + q1y * d.n1[1] >= 0.0f) return 0.0f;
    float u = q0x * dirx + q0y * diry;
    if (u < -d.r0 || u > l + d.r1) return 0.0f;
    float dist = cz_abs(q0x * diry - q0y * dirx);
    if (cap_start && u < 0.0f) dist = cz_len2(q0x, q0y);
    if (cap_end && u > l) dist = cz_len2(q1x, q1y);
    float rt = cz_mix(d.r0, d.r1, cz_clamp(u / l, 0.0f, 1.0f));
    float xx = dist / rt;
    return (xx < 1.0f) ? d.h0 * (rt / d.r0) * (1.0f - xx * xx) : 0.0f;
}

// upload_deposits' tile binning (src/film/mod.rs:1175-1224) on the device,
// with the host's float arithmetic for the rectangles.  k_dep_rects: one
// thread per deposit, its 64-tile rectangle (or none) and the tile counts;
// k_dep_scan: the tiles' first items in tile order under the item cap, and
// the counts cleared for the next frame; k_dep_fill: one wave per tile scans
// the deposits in index order and appends the ones covering it (ballot
// prefix), so every tile's items are ascending as the host writes them.
__global__ __launch_bounds__(256) void k_dep_rects(FilmDev F, float width_m, float height_m, float tile_m) {
    const uint32_t count = *F.dep_n;
    for (uint32_t di = blockIdx.x * 256u + threadIdx.x; di < count; di += gridDim.x * 256u) {
    const cz_deposit& d = F.deposits[di];
    int32_t r[4] = {-1, -1, -1, -1};
    const float rb = fmaxf(d.r0, d.r1);
    const bool inside = fmaxf(d.p0[0], d.p1[0]) + rb > 0.0f && fmaxf(d.p0[1], d.p1[1]) + rb > 0.0f &&
                        fminf(d.p0[0], d.p1[0]) - rb < width_m && fminf(d.p0[1], d.p1[1]) - rb < height_m;
    if (inside) {
        const float minx = fminf(d.p0[0], d.p1[0]) - rb, maxx = fmaxf(d.p0[0], d.p1[0]) + rb;
        const float miny = fminf(d.p0[1], d.p1[1]) - rb, maxy = fmaxf(d.p0[1], d.p1[1]) + rb;
        const int64_t mx = (int64_t)F.t64x - 1, my = (int64_t)F.t64y - 1;
        const int64_t tx0 = (int64_t)fminf(fmaxf(floorf(minx / tile_m), 0.0f), (float)mx);
        const int64_t tx1 = (int64_t)fminf(fmaxf(floorf(maxx / tile_m), 0.0f), (float)mx);
        const int64_t ty0 = (int64_t)fminf(fmaxf(floorf(miny / tile_m), 0.0f), (float)my);
        const int64_t ty1 = (int64_t)fminf(fmaxf(floorf(maxy / tile_m), 0.0f), (float)my);
        if (!(maxx < 0.0f || maxy < 0.0f)) {
            r[0] = (int32_t)tx0; r[1] = (int32_t)tx1; r[2] = (int32_t)ty0; r[3] = (int32_t)ty1;
            for (int64_t ty = ty0; ty <= ty1; ty++)
                for (int64_t tx = tx0; tx <= tx1; tx++) atomicAdd(&F.dep_cnt[(uint32_t)ty * F.t64x + (uint32_t)tx], 1u);
        }
    }
    *(int4*)(F.dep_rect + 4u * di) = make_int4(r[0], r[1], r[2], r[3]);
    }
}

__global__ __launch_bounds__(1024) void k_dep_scan(FilmDev F) {
    __shared__ uint64_t s_sum[1024];
    const uint32_t nt = F.t64x * F.t64y;
    const uint32_t run = (nt + 1023u) / 1024u;
    const uint32_t t0 = threadIdx.x * run, t1 = t0 + run < nt ? t0 + run : nt;
    uint64_t own = 0u;
    for (uint32_t t = t0; t < t1; t++) own += F.dep_cnt[t];
    s_sum[threadIdx.x] = own;
    __syncthreads();
    for (uint32_t off = 1u; off < 1024u; off <<= 1) {
        const uint64_t v = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
        __syncthreads();
        s_sum[threadIdx.x] += v;
        __syncthreads();
    }
    // items before this thread's run, then the host's take = min(count, cap - items)
    uint64_t items = s_sum[threadIdx.x] - own;
    const uint64_t cap = CZ_MAX_TILE_ITEMS;
    for (uint32_t t = t0; t < t1; t++) {
        const uint64_t c = F.dep_cnt[t];
        const uint64_t start = items < cap ? items : cap;
        const uint64_t end = items + c < cap ? items + c : cap;
        F.tile_start[t] = (uint32_t)start;
        F.tile_count[t] = (uint32_t)(end - start);
        F.dep_cnt[t] = 0u;
        items += c;
    }
}

__global__ __launch_bounds__(256) void k_dep_fill(FilmDev F) {
    const uint32_t count = *F.dep_n;
    const uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (t >= F.t64x * F.t64y) return;
    const uint32_t want = F.tile_count[t], base = F.tile_start[t];
    if (want == 0u) return;
    const int32_t tx = (int32_t)(t % F.t64x), ty = (int32_t)(t / F.t64x);
    const unsigned long long lmask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    uint32_t filled = 0u;
    // 8 x 64 rectangles per round, all loads issued before the first test
    constexpr uint32_t U = 8u;
    for (uint32_t d0 = 0; d0 < count && filled < want; d0 += 64u * U) {
        int4 r[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t di = d0 + 64u * u + lane;
            r[u] = di < count ? *(const int4*)(F.dep_rect + 4u * di) : make_int4(-1, -1, -1, -1);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t di = d0 + 64u * u + lane;
            const bool hit = r[u].x >= 0 && tx >= r[u].x && tx <= r[u].y && ty >= r[u].z && ty <= r[u].w;
            const unsigned long long b = __ballot(hit);
            const uint32_t pos = filled + (uint32_t)__popcll(b & lmask);
            if (hit && pos < want) F.tile_items[base + pos] = di;
            filled += (uint32_t)__popcll(b);
        }
    }
}

// deposit.wgsl:31-75 (full grid, 16x16 blocks)
__global__ __launch_bounds__(256) void k_deposit(FilmDev F) {
    uint32_t i = blockIdx.x * 16u + (threadIdx.x & 15u);
    uint32_t j = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (i >= F.W || j >= F.H) return;
    uint32_t tile = (i / TILE64) + (j / TILE64) * F.P.tiles[0];
    uint32_t n = F.tile_count[tile];
    if (n == 0u) return;
    uint32_t start = F.tile_start[tile];
    uint32_t c = idx(F, i, j);
    float px = ((float)i + 0.5f) * F.P.dx, py = ((float)j + 0.5f) * F.P.dx;
    Stack s;
    stack_load(F, c, s);
    float added = 0.0f, top_add = 0.0f;
    uint32_t top_lay = 0;
    for (uint32_t q = 0; q < n; q++) {
        const cz_deposit d = F.deposits[F.tile_items[start + q]];
        float abx = d.p1[0] - d.p0[0], aby = d.p1[1] - d.p0[1];
        float along = cz_clamp(((px - d.p0[0]) * abx + (py - d.p0[1]) * aby) / cz_max(abx * abx + aby * aby, 1e-30f), 0.0f, 1.0f);
        float reach = 1.415f * cz_max(d.r0, d.r1) + 0.6f * F.P.dx;
        float offx = px - d.p0[0] - along * abx, offy = py - d.p0[1] - along * aby;
        if (offx * offx + offy * offy > reach * reach) continue;
        float add = 0.0f;
        for (uint32_t sy = 0; sy < 4u; sy++) {
            for (uint32_t sx = 0; sx < 4u; sx++) {
                float psx = px + ((float)sx - 1.5f) * (0.25f * F.P.dx);
                float psy = py + ((float)sy - 1.5f) * (0.25f * F.P.dx);
                add = add + 0.0625f * ribbon_height(psx, psy, d);
            }
        }
        if (add > 0.0f) {
            float pig[16];
            unpack8(make_uint4(d.pig[0], d.pig[1], d.pig[2], d.pig[3]), make_uint4(d.pig[4], d.pig[5], d.pig[6], d.pig[7]), pig);
            stack_push_top(F, s, add, pig, 0.0f);
            added = added + add;
            if (add > top_add) {
                top_add = add;
                float vx = d.p1[0] - d.p0[0] + 1e-9f, vy = d.p1[1] - d.p0[1] + 0.0f;
                float vl = cz_len2(vx, vy);
                float turn = cz_atan2f(vy / vl, vx / vl) / 6.2831853f + 0.5f;
                uint32_t method = d.kind >> CZ_DEPOSIT_METHOD_SHIFT;
                top_lay = (method & 0xffu) | ((uint32_t)(cz_clamp(turn, 0.0f, 1.0f) * 65535.0f) << 8u);
            }
        }
    }
    if (added > 0.0f) {
        F.h[c] = F.h[c] + added;
        F.lay[c] = top_lay;
        F.tile_moved[(i >> 4) + (j >> 4) * F.tw] = 1u;
        stack_compact(F, s);
        stack_store_to(F.strata_h, F.strata_cure, F.strata_pig, c, s);
    }
}

// film.wgsl:1393-1425 as an ordered compaction (one block of 1024 threads
// per canvas): the list is ascending in tile index instead of atomic append
// order.  The same pass lists the tiles with cells for the cure pass (the
// NEEDS_CURE bit of k_mark, which it strips from the flags).  The flag bytes of a band of tile rows and its one-row halo are
// staged in LDS first, so every wake test reads the flags from before the
// pass and the neighbour reads stay on chip.
constexpr uint32_t COMPACT_LDS = 65536u;

template <class CS>
__global__ __launch_bounds__(1024) void k_compact_tiles(CS cs) {
    const FilmDev& F = cs.film(blockIdx.x);
    __shared__ uint8_t sf[COMPACT_LDS];
    __shared__ uint32_t wsum[16], wsum2[16];
    __shared__ uint32_t base, base2;
    const uint32_t tw = F.tw, th = F.th;
    const uint32_t band = COMPACT_LDS / tw - 2u;  // tile rows per band (tw <= 21845 at creation)
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) base = base2 = 0u;
    for (uint32_t ty0 = 0; ty0 < th; ty0 += band) {
        const uint32_t ty1 = ty0 + band < th ? ty0 + band : th;
        const uint32_t ry0 = ty0 > 0u ? ty0 - 1u : 0u, ry1 = ty1 + 1u < th ? ty1 + 1u : th;
        const uint32_t nst = (ry1 - ry0) * tw;
        const uint32_t* src = F.tile_flags + ry0 * tw;
        __syncthreads();
        for (uint32_t q0 = 0; q0 < nst; q0 += 8u * 1024u) {
            uint32_t v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint32_t q = q0 + (uint32_t)k * 1024u + threadIdx.x;
                v[k] = q < nst ? src[q] : 0u;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint32_t q = q0 + (uint32_t)k * 1024u + threadIdx.x;
                if (q < nst) sf[q] = (uint8_t)v[k];
            }
        }
        __syncthreads();
        for (uint32_t chunk = ty0 * tw; chunk < ty1 * tw; chunk += 1024u) {
            const uint32_t t = chunk + threadIdx.x;
            uint32_t active = 0u, fl_t = 0u;
            if (t < ty1 * tw) {
                const int tx = (int)(t % tw), ty = (int)(t / tw);
                fl_t = sf[t - ry0 * tw];
                uint32_t anyv = fl_t & MOVED;
                for (int dy = -1; dy <= 1; dy++) {
                    for (int dxi = -1; dxi <= 1; dxi++) {
                        if (dxi == 0 && dy == 0) continue;
                        const int xx = tx + dxi, yy = ty + dy;
                        if (xx >= 0 && yy >= 0 && xx < (int)tw && yy < (int)th) {
                            const uint32_t fl = sf[((uint32_t)yy - ry0) * tw + (uint32_t)xx];
                            bool need = true;
                            if (dxi == 1 && (fl & 2u) == 0u) need = false;
                            if (dxi == -1 && (fl & 4u) == 0u) need = false;
                            if (dy == 1 && (fl & 8u) == 0u) need = false;
                            if (dy == -1 && (fl & 16u) == 0u) need = false;
                            if (need) anyv |= fl & MOVED;
                        }
                    }
                }
                active = anyv != 0u ? 1u : 0u;
            }
            // block-wide exclusive scans of `active` and of the cure bit
            const uint32_t cure = (fl_t & NEEDS_CURE) != 0u ? 1u : 0u;
            const unsigned long long lmask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
            const unsigned long long bal = __ballot(active), bal2 = __ballot(cure);
            const uint32_t below = (uint32_t)__popcll(bal & lmask), below2 = (uint32_t)__popcll(bal2 & lmask);
            if (lane == 0u) {
                wsum[wave] = (uint32_t)__popcll(bal);
                wsum2[wave] = (uint32_t)__popcll(bal2);
            }
            __syncthreads();
            uint32_t woff = 0u, total = 0u, woff2 = 0u, total2 = 0u;
            for (uint32_t w = 0; w < 16u; w++) {
                if (w < wave) {
                    woff += wsum[w];
                    woff2 += wsum2[w];
                }
                total += wsum[w];
                total2 += wsum2[w];
            }
            if (t < ty1 * tw) {
                if (active) {
                    F.tile_list[base + woff + below] = t;
                    F.tile_flags[t] = (fl_t & 95u) | 32u;
                } else {
                    F.tile_flags[t] = fl_t & 31u;
                }
                if (cure) F.cure_list[base2 + woff2 + below2] = t;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                base += total;
                base2 += total2;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *F.active_count = base;
        *F.cure_count = base2;
    }
}

// The same compaction for one canvas in two short launches of one thread per
// tile, for canvases whose tile count would keep one block busy for tens of
// microseconds.  k_compact_bits applies the wake rule to the flags from
// k_mark, writes the tile's flags, one byte (1 active, 2 cure) and the
// block's two counts.  The flag bits other tiles' wake tests read (1..16 and
// MOVED) are the same before and after the write, so tiles need no ordering
// among themselves.  k_compact_list: each block sums the counts of the
// blocks before it, then places its active and cure tiles by a block scan,
// so both lists come out ascending.
__global__ __launch_bounds__(256) void k_compact_bits(FilmDev F) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    bool active = false, cure = false;
    if (t < F.nt16) {
        const uint32_t tw = F.tw, th = F.th;
        const int tx = (int)(t % tw), ty = (int)(t / tw);
        const uint32_t fl_t = F.tile_flags[t];
        uint32_t anyv = fl_t & MOVED;
        for (int dy = -1; dy <= 1; dy++) {
            for (int dxi = -1; dxi <= 1; dxi++) {
                if (dxi == 0 && dy == 0) continue;
                const int xx = tx + dxi, yy = ty + dy;
                if (xx >= 0 && yy >= 0 && xx < (int)tw && yy < (int)th) {
                    const uint32_t fl = F.tile_flags[(uint32_t)yy * tw + (uint32_t)xx];
                    bool need = true;
                    if (dxi == 1 && (fl & 2u) == 0u) need = false;
                    if (dxi == -1 && (fl & 4u) == 0u) need = false;
                    if (dy == 1 && (fl & 8u) == 0u) need = false;
                    if (dy == -1 && (fl & 16u) == 0u) need = false;
                    if (need) anyv |= fl & MOVED;
                }
            }
        }
        active = anyv != 0u;
        cure = (fl_t & NEEDS_CURE) != 0u;
        F.tile_flags[t] = active ? ((fl_t & 95u) | 32u) : (fl_t & 31u);
        F.tile_bits[t] = (uint8_t)((active ? 1u : 0u) | (cure ? 2u : 0u));
    }
    const uint32_t na = (uint32_t)__syncthreads_count(active ? 1 : 0);
    const uint32_t nc = (uint32_t)__syncthreads_count(cure ? 1 : 0);
    if (threadIdx.x == 0) {
        F.compact_cnt[2u * blockIdx.x] = na;
        F.compact_cnt[2u * blockIdx.x + 1u] = nc;
    }
}

__global__ __launch_bounds__(256) void k_compact_list(FilmDev F) {
    __shared__ uint32_t s_base[2], s_wa[4], s_wc[4];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (wave == 0u) {
        uint32_t a = 0u, c = 0u;
        for (uint32_t q = lane; q < blockIdx.x; q += 64u) {
            a += F.compact_cnt[2u * q];
            c += F.compact_cnt[2u * q + 1u];
        }
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            c += __shfl_xor(c, off, 64);
        }
        if (lane == 0u) {
            s_base[0] = a;
            s_base[1] = c;
        }
    }
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t b = t < F.nt16 ? F.tile_bits[t] : 0u;
    const unsigned long long lmask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    const unsigned long long ba = __ballot((b & 1u) != 0u), bc = __ballot((b & 2u) != 0u);
    if (lane == 0u) {
        s_wa[wave] = (uint32_t)__popcll(ba);
        s_wc[wave] = (uint32_t)__popcll(bc);
    }
    __syncthreads();
    uint32_t oa = s_base[0], oc = s_base[1];
    for (uint32_t w = 0; w < wave; w++) {
        oa += s_wa[w];
        oc += s_wc[w];
    }
    if (b & 1u) F.tile_list[oa + (uint32_t)__popcll(ba & lmask)] = t;
    if (b & 2u) F.cure_list[oc + (uint32_t)__popcll(bc & lmask)] = t;
    if (blockIdx.x + 1u == gridDim.x && threadIdx.x == 0u) {
        *F.active_count = s_base[0] + s_wa[0] + s_wa[1] + s_wa[2] + s_wa[3];
        *F.cure_count = s_base[1] + s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
    }
}

// ---------------------------------------------------------------------------
// shade_cell (shade.wgsl:29-83) with load_refl and spectrum_to_rgb
// (km.wgsl:35-62): the display colour and detail channels of every cell.
__device__ __forceinline__ float smoothstep_w(float e0, float e1, float x) {
    const float t = cz_clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

template <class WS>
__global__ __launch_bounds__(256) void k_shade(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        float r[NBAND];
#pragma unroll
        for (uint32_t b = 0; b < 8u; b++) {
            const uint32_t v = F.sub_refl[c * 8u + b];
            r[2u * b] = cz_unpack_lo(v);
            r[2u * b + 1u] = cz_unpack_hi(v);
        }
        const float hv = F.h[c];
        float wet = 0.0f, cure_top = 0.0f;
        const float4 th4 = F.strata_h[c];
        if (hv > 0.0f) {
            const float4 cu4 = F.strata_cure[c];
#pragma unroll 1
            for (uint32_t k = 0; k < 4u; k++) {
                const float t = k == 0u ? th4.x : (k == 1u ? th4.y : (k == 2u ? th4.z : th4.w));
                if (t <= 0.0f) continue;
                float pig[16];
                unpack8(F.strata_pig[(c * 4u + k) * 2u], F.strata_pig[(c * 4u + k) * 2u + 1u], pig);
#pragma unroll
                for (uint32_t b = 0; b < NBAND; b++) {
                    float kk, ss, R, T;
                    layer_ks(F.pigtab, pig, b, kk, ss);
                    km_layer(kk, ss, t, R, T);
                    r[b] = composite(R, T, r[b]);
                }
                const float cu = k == 0u ? cu4.x : (k == 1u ? cu4.y : (k == 2u ? cu4.z : cu4.w));
                wet = 1.0f - cu;
                cure_top = cu;
            }
        }
        const float k1 = 0.04f;
        const float k2 = cz_mix(0.4f, 0.6f, wet);
        float xyz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t b = 0; b < NBAND; b++) {
            const float rb = (1.0f - k1) * (1.0f - k2) * r[b] / cz_max(1.0f - k2 * r[b], 1e-3f);
#pragma unroll
            for (uint32_t q = 0; q < 3u; q++) xyz[q] = xyz[q] + rb * F.cmf[b * 3u + q];
        }
        float rgb[3] = {3.2406f * xyz[0] - 1.5372f * xyz[1] - 0.4986f * xyz[2],
                        -0.9689f * xyz[0] + 1.8758f * xyz[1] + 0.0415f * xyz[2],
                        0.0557f * xyz[0] - 0.2040f * xyz[1] + 1.0570f * xyz[2]};
#pragma unroll
        for (int q = 0; q < 3; q++) rgb[q] = cz_max(rgb[q], 0.0f);
        const float mc = F.m[c];
        const float painted = smoothstep_w(60e-6f, 150e-6f, mc);
        cure_top = cz_max(cure_top, painted * (1.0f - smoothstep_w(0.0f, 50e-6f, hv)));
        float rough = cz_mix(0.06f, 0.35f, sqrtf(cure_top));
        if (hv <= 0.0f) rough = cz_mix(0.85f, 0.35f, painted);
        const uint32_t lv = F.lay[c];
        const float turn = (float)((lv >> 8u) & 0xffffu) / 65536.0f;
        const float thickness = hv + cz_max(mc - 60e-6f, 0.0f);
        if (F.P.flags[0] == 1u && hv > 0.0f) {
            // the strata debug view
            const float ssum = th4.x + th4.y + th4.z + th4.w;
            bool bad = false;
            uint32_t nst = 0u;
#pragma unroll
            for (uint32_t k = 0; k < 4u; k++) {
                const float t = k == 0u ? th4.x : (k == 1u ? th4.y : (k == 2u ? th4.z : th4.w));
                if (t > 0.0f) nst = k + 1u;
                if (t <= 0.0f) continue;
                float pig[16];
                unpack8(F.strata_pig[(c * 4u + k) * 2u], F.strata_pig[(c * 4u + k) * 2u + 1u], pig);
                float cs = 0.0f;
#pragma unroll
                for (int i = 0; i < 13; i++) cs = cs + pig[i];
                if (!(cs > 0.0f)) bad = true;
            }
            if (ssum <= 0.0f) { rgb[0] = 0.0f; rgb[1] = 0.0f; rgb[2] = 1.0f; }
            else if (bad) { rgb[0] = 1.0f; rgb[1] = 0.0f; rgb[2] = 0.0f; }
            else if (cz_abs(ssum - hv) > 0.01f * hv) { rgb[0] = 0.0f; rgb[1] = 1.0f; rgb[2] = 0.0f; }
            else {
#pragma unroll
                for (int q = 0; q < 3; q++) rgb[q] = 0.3f + 0.15f * (float)nst;
            }
        }
        F.color[c] = make_uint2(cz_pack2(rgb[0], rgb[1]), cz_pack2(rgb[2], rough));
        F.detail[c] = make_uint2(cz_pack2((float)(lv & 0xffu) + turn, cure_top), cz_pack2(F.rheo[c].y, thickness));
    }
}

// ---------------------------------------------------------------------------
// cure.wgsl:3-48,62-123

__device__ inline void composition_rheo(const FilmDev& F, const Stack& s, float& mu_o, float& ty_o) {
    float total = 0.0f, lnmu = 0.0f, tau = 0.0f, thin = 0.0f, wsum = 0.0f, cure = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++) {
        float t = s.th[k];
        if (t <= 0.0f) continue;
        total = total + t;
        cure = cure + t * s.cu[k];
        float csum = 0.0f, lm = 0.0f, ty = 0.0f;
#pragma unroll
        for (uint32_t i = 0; i < (uint32_t)NPIG; i++) {
            float cc = s.pg[k][i];
            if (cc > 0.0f) {
                csum = csum + cc;
                lm = lm + cc * cz_logf(cz_max(F.pigtab[i * PIG_STRIDE + 32u], 1e-3f));
                ty = ty + cc * s.pg[k][13];
            }
        }
        if (s.pg[k][14] > 0.0f) lm = csum * (s.pg[k][14] - 10.0f);
        if (csum > 0.0f) {
            lnmu = lnmu + t * lm / csum;
            tau = tau + t * ty / csum;
            wsum = wsum + t;
        }
        thin = thin + t * s.pg[k][12];
    }
    if (total <= 0.0f) { mu_o = 1.0f; ty_o = 0.0f; return; }
    float mu = cz_logf(F.P.misc[2]);
    float ty = 0.0f;
    if (wsum > 0.0f) { mu = lnmu / wsum; ty = tau / wsum; }
    float tf = cz_clamp(thin / total, 0.0f, 0.95f);
    float cb = cure / total;
    mu = mu + (3.0f * cz_logf(1.0f - tf) + F.P.cure[2] * cb * cb);
    ty = ty * cz_powf(1.0f - tf, 4.0f) + F.P.cure[1] * cb * cb;
    mu_o = cz_expf(mu);
    ty_o = ty;
}

__device__ inline float2 packed_rheo(const FilmDev& F, uint32_t c, const float* t, const float* u) {
    float total = 0.0f, lnmu = 0.0f, tau = 0.0f, thin = 0.0f, wsum = 0.0f, cure = 0.0f;
    for (uint32_t k = 0; k < 4u; k++) {
        float tk = t[k];
        if (tk <= 0.0f) continue;
        total = total + tk;
        cure = cure + tk * u[k];
        uint4 a = F.strata_pig[(c * 4u + k) * 2u];
        uint4 b = F.strata_pig[(c * 4u + k) * 2u + 1u];
        float pc[12] = {cz_unpack_lo(a.x), cz_unpack_hi(a.x), cz_unpack_lo(a.y), cz_unpack_hi(a.y),
                        cz_unpack_lo(a.z), cz_unpack_hi(a.z), cz_unpack_lo(a.w), cz_unpack_hi(a.w),
                        cz_unpack_lo(b.x), cz_unpack_hi(b.x), cz_unpack_lo(b.y), cz_unpack_hi(b.y)};
        float exx = cz_unpack_lo(b.z), exy = cz_unpack_hi(b.z);
        float mu14 = cz_unpack_lo(b.w);
        float csum = 0.0f, ty = 0.0f;
        for (uint32_t i = 0; i < (uint32_t)NPIG; i++) {
            float cc = pc[i];
            if (cc > 0.0f) {
                csum = csum + cc;
                ty = ty + cc * exy;
            }
        }
        if (csum > 0.0f) {
            if (mu14 > 0.0f) {
                lnmu = lnmu + tk * (csum * (mu14 - 10.0f)) / csum;
            } else {
                // layer_lnmu (cure.wgsl:125-135)
                float lm = 0.0f;
                for (uint32_t i = 0; i < (uint32_t)NPIG; i++) {
                    float cc = pc[i];
                    if (cc > 0.0f) lm = lm + cc * cz_logf(cz_max(F.pigtab[i * PIG_STRIDE + 32u], 1e-3f));
                }
                lnmu = lnmu + tk * lm / csum;
            }
            tau = tau + tk * ty / csum;
            wsum = wsum + tk;
        }
        thin = thin + tk * exx;
    }
    if (total <= 0.0f) return make_float2(1.0f, 0.0f);
    float mu = cz_logf(F.P.misc[2]);
    float ty = 0.0f;
    if (wsum > 0.0f) { mu = lnmu / wsum; ty = tau / wsum; }
    float tf = cz_clamp(thin / total, 0.0f, 0.95f);
    float cb = cure / total;
    mu = mu + (3.0f * cz_logf(1.0f - tf) + F.P.cure[2] * cb * cb);
    ty = ty * cz_powf(1.0f - tf, 4.0f) + F.P.cure[1] * cb * cb;
    return make_float2(cz_expf(mu), ty);
}

// cure_step (cure.wgsl:119-159) of one wet cell with height hc: cure
// advance, rheology, and the commit of fully cured bottom strata into the
// relief and the substrate reflectance.  Returns the height after commits.
__device__ inline float cure_wet_cell(const FilmDev& F, uint32_t c, float hc) {
    float rate = F.P.cure[0] * F.P.frame_dt;
    float4 t4 = F.strata_h[c];
    float4 cu4 = F.strata_cure[c];
    float t[4] = {t4.x, t4.y, t4.z, t4.w};
    float cu[4] = {cu4.x, cu4.y, cu4.z, cu4.w};
    float u[4];
#pragma unroll
    for (int k = 0; k < 4; k++) u[k] = (t[k] > 0.0f) ? cz_min(cu[k] + rate, 1.0f) : cu[k];
    if (!(t[0] > 0.0f && u[0] >= 1.0f)) {
        F.strata_cure[c] = make_float4(u[0], u[1], u[2], u[3]);
        set_rheo(F, c, packed_rheo(F, c, t, u));
        if (cell_n_on(F)) {
            float n15[4];
#pragma unroll
            for (int k = 0; k < 4; k++) n15[k] = (t[k] > 0.0f) ? cz_unpack_hi(F.strata_pig[(c * 4u + k) * 2u + 1u].w) : 0.0f;
            F.n_cell[c] = strata_n(F, t, n15);
        }
    } else {
        Stack s;
        stack_load(F, c, s);
#pragma unroll
        for (int k = 0; k < 4; k++) s.cu[k] = u[k];
        float committed = 0.0f;
        uint32_t guard = 0;
        float mc = F.m[c];
        while (s.th[0] > 0.0f && s.cu[0] >= 1.0f && guard < 4u) {
            guard++;
            // KM composite of the bottom stratum over the substrate, one
            // packed word (two bands) at a time
#pragma unroll 1
            for (uint32_t w = 0; w < 8u; w++) {
                const uint32_t v = F.sub_refl[c * 8u + w];
                float r2[2] = {cz_unpack_lo(v), cz_unpack_hi(v)};
#pragma unroll
                for (uint32_t h2 = 0; h2 < 2u; h2++) {
                    float kk, ss, R, T;
                    layer_ks(F.pigtab, s.pg[0], 2u * w + h2, kk, ss);
                    km_layer(kk, ss, s.th[0], R, T);
                    r2[h2] = composite(R, T, r2[h2]);
                }
                F.sub_refl[c * 8u + w] = cz_pack2(r2[0], r2[1]);
            }
            committed = committed + s.th[0];
            mc = mc + s.th[0] * (1.0f - OIL_SHRINK) * (1.0f - cz_clamp(s.pg[0][12], 0.0f, 0.95f));
            s.th[0] = 0.0f;
            s.cu[0] = 0.0f;
            stack_compact(F, s);
        }
        F.m[c] = mc;
        if (committed > 0.0f) {
            hc = cz_max(hc - committed, 0.0f);
            F.h[c] = hc;
        }
        float mu, ty;
        composition_rheo(F, s, mu, ty);
        set_rheo(F, c, make_float2(mu, ty));
        if (cell_n_on(F)) {
            const float n15[4] = {s.pg[0][15], s.pg[1][15], s.pg[2][15], s.pg[3][15]};
            F.n_cell[c] = strata_n(F, s.th, n15);
        }
        stack_store_to(F.strata_h, F.strata_cure, F.strata_pig, c, s);
    }
    return hc;
}

// cure_step over the full grid fused with clear_acc (film.wgsl:1427-1434)
// for the cells of active tiles when do_clear (the dry_now pass).
template <class WS>
__global__ __launch_bounds__(256) void k_cure_clear(WS ws, int do_clear) {
    CZ_FOR_TILES(ws) {
        uint32_t i = (_t % F.tw) * 16u + _lx;
        uint32_t j = (_t / F.tw) * 16u + _ly;
        if (i >= F.W || j >= F.H) continue;
        uint32_t c = idx(F, i, j);
        float hc = F.h[c];
        if (hc <= 0.0f) {
            set_rheo(F, c, make_float2(1.0f, 0.0f));
            if (cell_n_on(F)) F.n_cell[c] = F.P.phys[2];
        } else {
            hc = cure_wet_cell(F, c, hc);
        }
        if (do_clear && cell_active(F, i, j)) {
            F.acc[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            F.h_frame[c] = hc;
        }
    }
}

// mark_tiles (film.wgsl:1365-1391), one wave per 16x16 tile: lane (r, q)
// owns cells 4q..4q+3 of tile row r, read as one 16-byte load when rows are
// 4-aligned.  The same pass writes the rheology of dry cells (the dry branch
// of cure_step, cure.wgsl:122-125, which reads nothing but the height; with
// per-cell indices on, a dry cell's index is reset to the film's) and
// marks with NEEDS_CURE the tiles holding cells for the wet branch; those
// run in k_cure over the list the compaction builds.  Each cell's mark bits
// use its height before its cure, as with the passes run in the reference's
// order (both touch only the cell itself).
template <class WS>
__global__ __launch_bounds__(256) void k_mark(WS ws) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = lane >> 2, q = lane & 3u;
    const uint32_t n = ws.count();
    for (uint32_t a = blockIdx.x * 4u + (threadIdx.x >> 6); a < n; a += gridDim.x * 4u) {
        const uint32_t it = ws.item(a);
        const FilmDev& F = ws.film(it);
        const uint32_t t = ws.tile(F, it);
        const uint32_t x0 = (t % F.tw) * 16u + 4u * q, y = (t / F.tw) * 16u + r;
        uint32_t bits = 0u;
        if (y < F.H && x0 < F.W) {
            const uint32_t c0 = idx(F, x0, y);
            const bool vec = (F.W & 3u) == 0u;  // then x0 + 3 < W and c0 is 4-aligned
            float hv[4];
            if (vec) {
                const float4 v = *(const float4*)(F.h + c0);
                hv[0] = v.x; hv[1] = v.y; hv[2] = v.z; hv[3] = v.w;
            } else {
#pragma unroll
                for (uint32_t k = 0; k < 4u; k++) hv[k] = (x0 + k < F.W) ? F.h[c0 + k] : 0.0f;
            }
            bool all_dry = true;
#pragma unroll
            for (uint32_t k = 0; k < 4u; k++) {
                if (x0 + k >= F.W) continue;
                const float hc = hv[k];
                if (hc > 0.0f) {
                    const uint32_t lx = 4u * q + k;
                    bits |= 1u;
                    if (lx < WAKE_BAND) bits |= 2u;
                    if (lx >= 16u - WAKE_BAND) bits |= 4u;
                    if (r < WAKE_BAND) bits |= 8u;
                    if (r >= 16u - WAKE_BAND) bits |= 16u;
                    if (cz_abs(hc - F.h_frame[c0 + k]) > FRAME_EPS) bits |= MOVED;
                }
                if (!(hc <= 0.0f)) {
                    all_dry = false;
                    bits |= NEEDS_CURE;
                }
            }
            if (vec && all_dry) {
                float4* rp = (float4*)(F.rheo + c0);
                rp[0] = make_float4(1.0f, 0.0f, 1.0f, 0.0f);
                rp[1] = make_float4(1.0f, 0.0f, 1.0f, 0.0f);
                *(float4*)(F.lmu_cell + c0) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // ln 1
                if (cell_n_on(F)) {
                    const float nf = F.P.phys[2];
                    *(float4*)(F.n_cell + c0) = make_float4(nf, nf, nf, nf);
                }
            } else {
#pragma unroll
                for (uint32_t k = 0; k < 4u; k++)
                    if (x0 + k < F.W && hv[k] <= 0.0f) {
                        set_rheo(F, c0 + k, make_float2(1.0f, 0.0f));
                        if (cell_n_on(F)) F.n_cell[c0 + k] = F.P.phys[2];
                    }
            }
        }
        uint32_t flags = 0u;
        const uint32_t bl[7] = {1u, 2u, 4u, 8u, 16u, MOVED, NEEDS_CURE};
#pragma unroll
        for (int k = 0; k < 7; k++)
            if (__ballot((bits & bl[k]) != 0u)) flags |= bl[k];
        if (lane == 0u) {
            if (F.tile_moved[t] != 0u) flags |= MOVED;
            F.tile_moved[t] = 0u;
            F.tile_flags[t] = flags;
        }
    }
}

// The wet branch of cure_step (cure.wgsl:126-159) on the tiles k_mark
// listed, one cell per thread.
template <class WS>
__global__ __launch_bounds__(256) void k_cure(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        const float hc = F.h[c];
        if (!(hc <= 0.0f)) cure_wet_cell(F, c, hc);
    }
}

// ---------------------------------------------------------------------------
// One substep's first phase on a tile: [nose of the previous substep]
// (film.wgsl:680-723) + laplacian (film.wgsl:16-42) + plug (film.wgsl:58-87).
//
// The block first stages h, m, the previous offsets, rheology and front
// normals of the tile and a 3-cell ring (a 22x22 window) in LDS with row-wise
// loads, plus the flags of the 3x3 tiles around it.  Then, from LDS:
//  - nose for the tile and a 1-cell ring (active cells; the ring is
//    recomputed identically by the neighbouring tile's block), since the
//    laplacian of a cell reads the new offsets and front normals of its 8
//    neighbours.  Only a nose target beyond the window reads global memory;
//  - plug's hb_flux(g_tilt, h_n, mu_n, tau_n) once per window cell, summed
//    over each 7x7 window in the reference's order;
//  - the laplacian and pressures of the tile's cells.
// The tile's results go to F.delta_out / F.fnorm / F.pres / F.tmp.
constexpr int PLUG_SPAN = 16 + 2 * PLUG_RADIUS;   // = the staged window
// nose compares candidate offsets by their squared length, an integer: the
// offsets are small integers (|cx|, |cy| <= 129), so cx^2 + cy^2 is exact in
// FP32 and its correctly rounded square root is strictly increasing over the
// integers up to 2^24, so the shader's comparisons of length(c) pick the same
// candidate, ties included; length <= 3 is cx^2 + cy^2 <= 9.
constexpr int kNoseNone = 1 << 30;
constexpr int RING_SPAN = 18;
constexpr int WIN = PLUG_SPAN, WIN_R = PLUG_RADIUS;

struct LapWin {
    float h[WIN * WIN];
    float m[WIN * WIN];
    uint32_t d[WIN * WIN];
    float2 rheo[WIN * WIN];
    float2 fn[WIN * WIN];
    float nc[WIN * WIN];    // the cells' power-law index (staged with CZ_FLAG_CELL_N only)
    float lmu[WIN * WIN];   // ln max(mu, 1e-4)
    uint32_t tf[9];
};

struct LapShared {
    LapWin sw;
    // plug per window cell: x = its height (0 dry), negated when it has a
    // yield stress; y = its hb_flux.  One 8-byte read per disc cell.
    float2 s_hq[WIN * WIN];
    uint32_t s_d[RING_SPAN * RING_SPAN];
    float2 s_fn[RING_SPAN * RING_SPAN];
    float s_us[RING_SPAN * RING_SPAN];
};

// One tile of the lap/plug pass (every thread of the block calls it).  BT:
// threads of the block (256, or 1024 for the wide shape, whose window
// phases then take one pass; the per-cell tail runs on threads 0-255).
template <int BT = 256>
__device__ __forceinline__ void lap_plug_tile(const FilmDev& F, uint32_t _t, LapShared& S, int do_nose, int do_plug,
                                              int do_clear) {
    LapWin& sw = S.sw;
    float2* s_hq = S.s_hq;
    uint32_t* s_d = S.s_d;
    float2* s_fn = S.s_fn;
    float* s_us = S.s_us;
    const uint32_t _lx = threadIdx.x & 15u, _ly = (threadIdx.x >> 4) & 15u;
    CZ_TILE_CELL(F)
    const bool cell_th = BT == 256 || threadIdx.x < 256u;
    const float gb = F.k_gb_plug;
    const bool plug_on = do_plug && !(gb < 1e-6f);
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int tx0 = ttx * 16, ty0 = tty * 16;
    const int wx0 = tx0 - WIN_R, wy0 = ty0 - WIN_R;
    const int W_ = (int)F.W, H_ = (int)F.H;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < (uint32_t)(WIN * WIN); q += (uint32_t)BT) {
        const int gx = wx0 + (int)(q % WIN), gy = wy0 + (int)(q / WIN);
        if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
            const uint32_t nn = idx(F, (uint32_t)gx, (uint32_t)gy);
            sw.h[q] = F.h[nn];
            sw.m[q] = F.m[nn];
            sw.d[q] = F.delta[nn];
            sw.rheo[q] = F.rheo[nn];
            sw.fn[q] = F.fnorm[nn];
            if (cell_n_on(F)) sw.nc[q] = F.n_cell[nn];
            sw.lmu[q] = F.lmu_cell[nn];
        } else {
            sw.h[q] = 0.0f;
            sw.m[q] = 0.0f;
            sw.d[q] = 0u;
            sw.rheo[q] = make_float2(0.0f, 0.0f);
            sw.fn[q] = make_float2(0.0f, 0.0f);
        }
    }
    if (threadIdx.x < 9u) {
        const int ax = ttx - 1 + (int)(threadIdx.x % 3u), ay = tty - 1 + (int)(threadIdx.x / 3u);
        sw.tf[threadIdx.x] = (ax >= 0 && ay >= 0 && ax < (int)F.tw && ay < (int)F.th)
                                 ? F.tile_flags[(uint32_t)ax + (uint32_t)ay * F.tw] : 0u;
    }
    __syncthreads();
    // window accessors (callers keep (gx, gy) inside the canvas)
    auto wq = [&](int gx, int gy) -> int { return (gy - wy0) * WIN + (gx - wx0); };
    auto in_win = [&](int gx, int gy) -> bool {
        return gx >= wx0 && gy >= wy0 && gx < wx0 + WIN && gy < wy0 + WIN;
    };
    auto act_w = [&](int gx, int gy) -> bool {
        return (sw.tf[((gy >> 4) - tty + 1) * 3 + ((gx >> 4) - ttx + 1)] & 32u) != 0u;
    };
    // nose (film.wgsl:680-723) of ring cell (i, j): offset to the nearest
    // dry cell from the previous offsets and the heights after apply, and
    // the front normal (film.wgsl:160-180) within 3 cells
    auto nose_w = [&](int i, int j, uint32_t& d_out, float2& fn_out) {
        d_out = 0u;
        fn_out = make_float2(0.0f, 0.0f);
        if (sw.h[wq(i, j)] <= 0.0f) return;
        // The candidates of the 3x3 neighbourhood (k = 3 (dj + 1) + di + 1,
        // the shader's order): a dry neighbour itself, a wet one's target
        // when that target is dry.  A target in the window is read from
        // LDS as its candidate is formed; the targets beyond it are read
        // from HBM for all nine before any is compared, so that those reads
        // are in flight together.  cand[k]: the candidate's offset (10-bit
        // two's complement x and y) and kind (bits 20-21: 0 none, 1 dry
        // neighbour, 2 target in the window, 3 target in HBM).
        uint32_t cand[9];
        float ht[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const int di = k % 3 - 1, dj = k / 3 - 1;
            const int xx = i + di, yy = j + dj;
            uint32_t kind = 0u;
            int cx = di, cy = dj;
            ht[k] = 0.0f;
            if (xx >= 0 && yy >= 0 && xx < W_ && yy < H_) {
                if (sw.h[wq(xx, yy)] > 0.0f) {
                    const uint32_t dv = sw.d[wq(xx, yy)];
                    if ((dv & 0x10000u) != 0u) {
                        int ox, oy;
                        unpack_off(dv, ox, oy);
                        const int tx = i + ox + di, ty = j + oy + dj;
                        if (tx >= 0 && ty >= 0 && tx < W_ && ty < H_) {
                            cx = ox + di;
                            cy = oy + dj;
                            kind = in_win(tx, ty) ? 2u : 3u;
                            if (kind == 2u) ht[k] = sw.h[wq(tx, ty)];
                        }
                    }
                } else if (!(di == 0 && dj == 0)) {
                    kind = 1u;
                }
            }
            cand[k] = ((uint32_t)cx & 0x3ffu) | (((uint32_t)cy & 0x3ffu) << 10) | (kind << 20);
        }
        auto cand_x = [](uint32_t v) { return (int)(v << 22) >> 22; };
        auto cand_y = [](uint32_t v) { return (int)(v << 12) >> 22; };
#pragma unroll
        for (int k = 0; k < 9; k++)
            if ((cand[k] >> 20) == 3u) ht[k] = F.h[idx(F, (uint32_t)(i + cand_x(cand[k])), (uint32_t)(j + cand_y(cand[k])))];
        int bx = 64, by = 64;
        int bd = kNoseNone;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const uint32_t kind = cand[k] >> 20;
            if (kind == 0u || (kind >= 2u && ht[k] > 0.0f)) continue;
            const int cx = cand_x(cand[k]), cy = cand_y(cand[k]);
            const int d = cx * cx + cy * cy;
            if (d < bd) { bd = d; bx = cx; by = cy; }
        }
        if (bd == kNoseNone) return;
        d_out = pack_off(bx, by);
        if (bd <= 9) {
            // the disc's heights (inside the window: the ring cell is at
            // most one cell off the tile) read together, then summed in
            // the shader's order
            float hd[25];
#pragma unroll
            for (int q = 0; q < 25; q++) {
                const int di = q % 5 - 2, dj = q / 5 - 2;
                hd[q] = (di * di + dj * dj <= 5) ? sw.h[wq(i + di, j + dj)] : 0.0f;
            }
            float dx = 0.0f, dy = 0.0f;
#pragma unroll
            for (int dj = -2; dj <= 2; dj++) {
#pragma unroll
                for (int di = -2; di <= 2; di++) {
                    if (di == 0 && dj == 0) continue;
                    const int xx = i + di, yy = j + dj;
                    if (xx < 0 || yy < 0 || xx >= W_ || yy >= H_) continue;
                    if (di * di + dj * dj > 5) continue;
                    if (hd[(dj + 2) * 5 + di + 2] <= F.P.ein[3]) {
                        dx = dx + FN_QUOT.v[di + 2][dj + 2];
                        dy = dy + FN_QUOT.v[dj + 2][di + 2];
                    }
                }
            }
            const float l = cz_len2(dx, dy);
            fn_out = (l < 1e-6f) ? make_float2(0.0f, 0.0f) : make_float2(dx / l, dy / l);
        }
    };
    // A block with a thread for every ring cell and every window cell (the
    // wide shape) gives the ring's nose and usurf and the window's plug
    // fluxes to different threads, so that the two chains run side by side;
    // a 256-thread block walks both in turn.
    constexpr uint32_t RING_N = RING_SPAN * RING_SPAN, PLUG_N = WIN * WIN;
    constexpr bool SPLIT = BT >= (int)(RING_N + PLUG_N);
    const uint32_t ring_end = (!SPLIT || threadIdx.x < RING_N) ? RING_N : 0u;
    const uint32_t plug_q0 = SPLIT ? threadIdx.x - RING_N : threadIdx.x;
    const uint32_t plug_end = (!SPLIT || threadIdx.x >= RING_N) ? PLUG_N : 0u;
    for (uint32_t q = threadIdx.x; q < ring_end; q += (uint32_t)BT) {
        const int gx = tx0 - 1 + (int)(q % RING_SPAN), gy = ty0 - 1 + (int)(q / RING_SPAN);
        uint32_t dv = 0u;
        float2 fn = make_float2(0.0f, 0.0f);
        if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
            if (do_nose && act_w(gx, gy)) {
                nose_w(gx, gy, dv, fn);
            } else {
                dv = sw.d[wq(gx, gy)];
                fn = sw.fn[wq(gx, gy)];
            }
        }
        s_d[q] = dv;
        s_fn[q] = fn;
        // usurf (film.wgsl:147-152) of the ring cell, once for the nine
        // laplacians that read it
        float us = 0.0f;
        if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
            const float hc = sw.h[wq(gx, gy)];
            float hold = 0.0f;
            if (hc > 0.0f) hold = hold_level(F, nose_dist_b(dv));
            us = hc + sw.m[wq(gx, gy)] - hold;
        }
        s_us[q] = us;
    }
    if (plug_on) {
        for (uint32_t q = plug_q0; q < plug_end; q += (uint32_t)BT) {
            float hn = sw.h[q], qv = 0.0f;
            bool yl = false;
            if (hn > 0.0f) {
                const float2 r = sw.rheo[q];
                yl = r.y > 0.0f;
                float oy_, oup_;
                const float nq = cell_n_on(F) ? sw.nc[q] : F.P.phys[2];
                qv = hb_flux(nq, gb, true, F.k_lg_plug, hn, Visc{cz_max(r.x, 1e-4f), sw.lmu[q], true}, r.y, oy_, oup_);
            } else {
                hn = 0.0f;
            }
            s_hq[q] = make_float2(yl ? -hn : hn, qv);
        }
    }
    __syncthreads();
    if (!in_grid || !cell_th) return;
    // clear_acc (film.wgsl:1427-1434) after the cure pass, with the first
    // active-tile kernel of the frame: do_clear 1 = clear, 2 = clear when
    // the canvas has no implicit step this frame (its implicit lap cleared)
    const int i = (int)x, j = (int)y;
    const float hv = sw.h[wq(i, j)];
    if (do_clear == 1 || (do_clear == 2 && F.imp_steps == 0u)) {
        F.acc[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        F.h_frame[c] = hv;
    }
    // the heights the next apply's limiter reads for this cell: apply
    // updates h while neighbouring tiles' blocks evaluate their rings
    if (do_plug) F.h_sub[c] = hv;
    const int own = (int)(_ly + 1u) * RING_SPAN + (int)(_lx + 1u);
    if (do_nose) {
        F.delta_out[c] = s_d[own];
        F.fnorm[c] = s_fn[own];
    }
    // usurf at mirrored coordinates, from the ring values
    auto usurf_l = [&](int ii, int jj) -> float {
        const int ri = (int)refl1_i(F, ii), rj = (int)refl1_j(F, jj);
        return s_us[(rj - ty0 + 1) * RING_SPAN + (ri - tx0 + 1)];
    };
    const float s0 = hv + sw.m[wq(i, j)];
    float hold = 0.0f, pl = 0.0f;
    const float dself = nose_dist_b(s_d[own]);
    if (hv > 0.0f) hold = hold_level(F, dself);
    const float u0 = s0 - hold;
    const float ax = usurf_l(i - 1, j) + usurf_l(i + 1, j) + usurf_l(i, j - 1) + usurf_l(i, j + 1);
    const float dg = usurf_l(i - 1, j - 1) + usurf_l(i + 1, j - 1) + usurf_l(i - 1, j + 1) + usurf_l(i + 1, j + 1);
    const float l = (W_AXIS * ax + W_DIAG_LAP * dg - (4.0f * W_AXIS + 4.0f * W_DIAG_LAP) * u0) * F.P.inv_dx * F.P.inv_dx;
    if (hv > 0.0f) {
        const float d = dself;
        if (d <= 3.0f) {
            const float2 nn = s_fn[own];
            if (nn.x * nn.x + nn.y * nn.y > 0.5f) {
                // line_pressure (film.wgsl:190-201) with the normals from LDS
                float lp = 0.0f;
                if (!(x < 1u || y < 1u || x + 1u >= F.W || y + 1u >= F.H)) {
                    const float2 nl = s_fn[own - 1], nr = s_fn[own + 1];
                    const float2 nd = s_fn[own - RING_SPAN], nu = s_fn[own + RING_SPAN];
                    if (!(nl.x * nl.x + nl.y * nl.y < 0.5f || nr.x * nr.x + nr.y * nr.y < 0.5f ||
                          nd.x * nd.x + nd.y * nd.y < 0.5f || nu.x * nu.x + nu.y * nu.y < 0.5f)) {
                        const float curv = (nr.x - nl.x + nu.y - nd.y) * 0.5f * F.P.inv_dx;
                        lp = F.P.phys[1] * curv * F.P.front2[2];
                    }
                }
                pl = lp * cz_expf(-(d - 1.0f) * 0.7f);
            }
        }
    }
    F.pres[c] = make_float2(-F.P.phys[1] * l + disjoining(F, hv), F.P.phys[0] * F.P.g[2] * (s0 - hold) + pl);
    if (!do_plug) return;
    float tv = 0.0f;
    if (hv > 0.0f && plug_on) {
        float sum = 0.0f, mass = 0.0f;
        bool yielding = false;
        for (int dy = 0; dy < 2 * PLUG_RADIUS + 1; dy++) {
            for (int dxi = 0; dxi < 2 * PLUG_RADIUS + 1; dxi++) {
                const int q = (int)(_ly + dy) * PLUG_SPAN + (int)(_lx + dxi);
                // a dry cell holds (+0, +0): adding it leaves the sums' bits
                // as the shader's skip does (both sums are >= +0)
                const float2 v = s_hq[q];
                yielding = yielding || v.x < 0.0f;
                sum = sum + v.y;
                mass = mass + cz_abs(v.x);
            }
        }
        if (yielding && mass > 0.0f) tv = sum / mass;
    }
    F.tmp[c] = tv;
}

template <class WS>
__global__ __launch_bounds__(256) void k_lap_plug(WS ws, int do_nose, int do_plug, int do_clear) {
    __shared__ LapShared S;
    CZ_FOR_TILES(ws) lap_plug_tile(F, _t, S, do_nose, do_plug, do_clear);
}

// The wide shape: 1024 threads per tile.
template <class WS>
__global__ __launch_bounds__(1024) void k_lap_plug_w(WS ws, int do_nose, int do_plug, int do_clear) {
    __shared__ LapShared S;
    CZ_FOR_TILES(ws) lap_plug_tile<1024>(F, _t, S, do_nose, do_plug, do_clear);
}

// The same pass without plug: the nose of a frame's last substep and the
// laplacian of the implicit steps.  A kernel of its own, so that a kernel
// trace tells the substep passes (k_lap_plug) from these.
template <class WS>
__global__ __launch_bounds__(256) void k_lap_tail(WS ws, int do_nose, int do_clear) {
    __shared__ LapShared S;
    CZ_FOR_TILES(ws) lap_plug_tile(F, _t, S, do_nose, 0, do_clear);
}

// ... in 1024-thread blocks, after substeps that ran in the wide shapes
template <class WS>
__global__ __launch_bounds__(1024) void k_lap_tail_w(WS ws, int do_nose, int do_clear) {
    __shared__ LapShared S;
    CZ_FOR_TILES(ws) lap_plug_tile<1024>(F, _t, S, do_nose, 0, do_clear);
}

// fluxes (film.wgsl:375-455), face-parallel: each block takes one
// (tile, face) pair and reads its fields from HBM through the caches; 4x the
// waves of a cell-per-thread launch to hide the dependent loads of the
// contact-line terms (the shape for very small active sets).
template <class WS>
__global__ __launch_bounds__(256) void k_fluxes_faces(WS ws) {
    const uint32_t nact = ws.count();
    const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4;
    const XcdRange rg = xcd_range<true>(4u * nact);
    for (uint32_t wk = rg.first; wk < rg.end; wk += rg.step) {
        const uint32_t it = ws.item(wk >> 2);
        const FilmDev& F = ws.film(it);
        float* flux = (float*)F.flux;
        const uint32_t t = ws.tile(F, it);
        const int k = (int)(wk & 3u);
        const uint32_t x = (t % F.tw) * 16u + lx, y = (t / F.tw) * 16u + ly;
        if (x >= F.W || y >= F.H) continue;
        const uint32_t c = idx(F, x, y);
        const GlobalRead g{F};
        float2 p00 = pressure_at(F, g, (int)x, (int)y);
        float q = 0.0f;
        FaceGeom fg = face_geom(F, g, (int)x, (int)y, k, true, p00);
        if (fg.has) {
            FaceFlux ff = face(F, g, (int)x, (int)y, fg.ib, fg.jb, fg.gcap, fg.ghyd, fg.gtilt, fg.gtn, fg.diag, fg.dir);
            q = ff.q * ((k < 2) ? W_AXIS : W_DIAG_FLUX);
        }
        flux[4u * c + (uint32_t)k] = q;
    }
}

// The fields of the face terms staged for a tile and a FLUX_R-cell ring
// (front_shape of the dry neighbour reaches 4 cells from the tile); reads
// outside the window (the far steps of cell_excess) go to HBM.
constexpr int FLUX_R = 4;
constexpr int FLUX_WIN = 16 + 2 * FLUX_R;

// The fields the flux of a cell reads, staged for the tile and an R-cell ring
// (the flux of a cell reads cells up to 4 away, plus the far steps of
// cell_excess).
template <int R>
struct FluxWinT {
    static constexpr int N = 16 + 2 * R;
    float h[N * N];
    uint32_t d[N * N];
    float2 rheo[N * N];
    float2 fn[N * N];
    float2 pres[N * N];
    float tmp[N * N];
    float lmu[N * N];   // ln max(mu, 1e-4)
    float fd[N * N];    // front distance (0 when dry)
    float nc[N * N];    // the cells' power-law index (staged with CZ_FLAG_CELL_N only)
    uint32_t tf[9];
};
using FluxWin = FluxWinT<FLUX_R>;

// Accessor of a staged window.  SNAP: the heights are the ones before this
// substep's apply (h_sub for cells of active tiles, h elsewhere, where no
// apply writes), as the fused flux + apply launch needs while other blocks
// apply; otherwise the live heights, which equal them during the flux pass.
template <int R, bool SNAP>
struct WinReadT {
    static constexpr int N = FluxWinT<R>::N;
    static constexpr bool kFarCheap = true;   // far reads: LDS or a plain load
    const FilmDev& F;
    const FluxWinT<R>& w;
    int wx0, wy0, ttx, tty;
    __device__ __forceinline__ int q(int i, int j) const { return (j - wy0) * N + (i - wx0); }
    // cells within R of the tile: always inside the window
    __device__ __forceinline__ float h(int i, int j) const { return w.h[q(i, j)]; }
    __device__ __forceinline__ uint32_t delta(int i, int j) const { return w.d[q(i, j)]; }
    __device__ __forceinline__ float2 rheo(int i, int j) const { return w.rheo[q(i, j)]; }
    __device__ __forceinline__ float2 fnorm(int i, int j) const { return w.fn[q(i, j)]; }
    __device__ __forceinline__ float2 pres(int i, int j) const { return w.pres[q(i, j)]; }
    __device__ __forceinline__ float tmp(int i, int j) const { return w.tmp[q(i, j)]; }
    __device__ __forceinline__ float lmu(int i, int j) const { return w.lmu[q(i, j)]; }
    __device__ __forceinline__ float fdist(int i, int j) const { return w.fd[q(i, j)]; }
    // read when the per-cell index is on only
    __device__ __forceinline__ float ncell(int i, int j) const { return w.nc[q(i, j)]; }
    float hold1_, theta_;
    __device__ __forceinline__ float hold1() const { return hold1_; }
    __device__ __forceinline__ float theta() const { return theta_; }
    // the far steps of cell_excess: the window when inside, else HBM
    __device__ __forceinline__ bool in(int i, int j) const {
        return (unsigned)(i - wx0) < (unsigned)N && (unsigned)(j - wy0) < (unsigned)N;
    }
    __device__ __forceinline__ float h_far(int i, int j) const {
        if (in(i, j)) return w.h[q(i, j)];
        const uint32_t c = (uint32_t)j * F.W + (uint32_t)i;
        if (SNAP && cell_active(F, (uint32_t)i, (uint32_t)j)) return F.h_sub[c];
        return F.h[c];
    }
    __device__ __forceinline__ uint32_t delta_far(int i, int j) const {
        if (in(i, j)) return w.d[q(i, j)];
        return F.delta[(uint32_t)j * F.W + (uint32_t)i];
    }
    // cells within 16 of the tile: one of the 3x3 tiles around it
    __device__ __forceinline__ bool active(int i, int j) const {
        return (w.tf[((j >> 4) - tty + 1) * 3 + ((i >> 4) - ttx + 1)] & 32u) != 0u;
    }
};
using WinRead = WinReadT<FLUX_R, false>;

// Stage the window of tile (ttx, tty) with origin (wx0, wy0).  The 3x3 tile
// flags are loaded first (they pick the heights' source when SNAP).
template <int R, bool SNAP, int BT = 256>
__device__ __forceinline__ void flux_stage(const FilmDev& F, FluxWinT<R>& sw, int ttx, int tty, int wx0, int wy0) {
    constexpr int N = FluxWinT<R>::N;
    const int W_ = (int)F.W, H_ = (int)F.H;
    __syncthreads();
    if (threadIdx.x < 9u) {
        const int ax = ttx - 1 + (int)(threadIdx.x % 3u), ay = tty - 1 + (int)(threadIdx.x / 3u);
        sw.tf[threadIdx.x] = (ax >= 0 && ay >= 0 && ax < (int)F.tw && ay < (int)F.th)
                                 ? F.tile_flags[(uint32_t)ax + (uint32_t)ay * F.tw] : 0u;
    }
    if (SNAP) __syncthreads();
    for (uint32_t q = threadIdx.x; q < (uint32_t)(N * N); q += (uint32_t)BT) {
        const int gx = wx0 + (int)(q % N), gy = wy0 + (int)(q / N);
        if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
            const uint32_t nn = idx(F, (uint32_t)gx, (uint32_t)gy);
            float hv = F.h[nn];
            if (SNAP) {
                const float hs = F.h_sub[nn];
                const bool act = (sw.tf[((gy >> 4) - tty + 1) * 3 + ((gx >> 4) - ttx + 1)] & 32u) != 0u;
                hv = act ? hs : hv;
            }
            sw.h[q] = hv;
            sw.d[q] = F.delta[nn];
            sw.rheo[q] = F.rheo[nn];
            sw.fn[q] = F.fnorm[nn];
            sw.pres[q] = F.pres[nn];
            sw.tmp[q] = F.tmp[nn];
            sw.lmu[q] = F.lmu_cell[nn];
            sw.fd[q] = hv <= 0.0f ? 0.0f : nose_dist_b(sw.d[q]);
            if (cell_n_on(F)) sw.nc[q] = F.n_cell[nn];
        }
    }
    __syncthreads();
}

// fluxes (film.wgsl:375-455) of cell (x, y): its four owned faces
template <class A>
__device__ __forceinline__ float4 cell_fluxes(const FilmDev& F, const A& g, int x, int y) {
    const float2 p00 = pressure_at(F, g, x, y);
    float4 qv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        FaceGeom fg = face_geom(F, g, x, y, k, true, p00);
        if (!fg.has) continue;
        FaceFlux ff = face(F, g, x, y, fg.ib, fg.jb, fg.gcap, fg.ghyd, fg.gtilt, fg.gtn, fg.diag, fg.dir);
        const float q = ff.q * ((k < 2) ? W_AXIS : W_DIAG_FLUX);
        if (k == 0) qv.x = q;
        else if (k == 1) qv.y = q;
        else if (k == 2) qv.z = q;
        else qv.w = q;
    }
    return qv;
}

// One tile of the staged flux pass (every thread of the block calls it): the
// tile's cells, one cell's four faces per thread, the fields in LDS.
__device__ __forceinline__ void flux_tile(const FilmDev& F, uint32_t _t, FluxWin& sw) {
    const uint32_t _lx = threadIdx.x & 15u, _ly = threadIdx.x >> 4;
    CZ_TILE_CELL(F)
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int wx0 = ttx * 16 - FLUX_R, wy0 = tty * 16 - FLUX_R;
    flux_stage<FLUX_R, false>(F, sw, ttx, tty, wx0, wy0);
    if (!in_grid) return;
    const WinRead g{F, sw, wx0, wy0, ttx, tty, F.k_hold1, F.k_theta};
    F.flux[c] = cell_fluxes(F, g, (int)x, (int)y);
}

template <class WS>
__global__ __launch_bounds__(256) void k_fluxes(WS ws) {
    __shared__ FluxWin sw;
    CZ_FOR_TILES(ws) flux_tile(F, _t, sw);
}

// The wide shape of the staged flux pass: 1024 threads per tile, one face of
// one cell per thread (cell = thread % 256, face = thread / 256), each face
// evaluated exactly as cell_fluxes evaluates it.
__device__ __forceinline__ void flux_tile_w(const FilmDev& F, uint32_t _t, FluxWin& sw) {
    const uint32_t k = threadIdx.x >> 8;
    const uint32_t _lx = threadIdx.x & 15u, _ly = (threadIdx.x >> 4) & 15u;
    CZ_TILE_CELL(F)
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int wx0 = ttx * 16 - FLUX_R, wy0 = tty * 16 - FLUX_R;
    flux_stage<FLUX_R, false, 1024>(F, sw, ttx, tty, wx0, wy0);
    if (!in_grid) return;
    const WinRead g{F, sw, wx0, wy0, ttx, tty, F.k_hold1, F.k_theta};
    const float2 p00 = pressure_at(F, g, (int)x, (int)y);
    float q = 0.0f;
    FaceGeom fg = face_geom(F, g, (int)x, (int)y, (int)k, true, p00);
    if (fg.has) {
        FaceFlux ff = face(F, g, (int)x, (int)y, fg.ib, fg.jb, fg.gcap, fg.ghyd, fg.gtilt, fg.gtn, fg.diag, fg.dir);
        q = ff.q * ((k < 2u) ? W_AXIS : W_DIAG_FLUX);
    }
    ((float*)F.flux)[4u * c + k] = q;
}

template <class WS>
__global__ __launch_bounds__(1024) void k_fluxes_w(WS ws) {
    __shared__ FluxWin sw;
    CZ_FOR_TILES(ws) flux_tile_w(F, _t, sw);
}

struct Inflow { float best; uint32_t src; float div; float accv; };

// Copy of the top stratum (and rheology) of src into the newly wetted cell c
// (film.wgsl:847-860, strata.wgsl:312-322).
__device__ inline void inherit_from(const FilmDev& F, uint32_t c, uint32_t src, float hn, bool copy_rheo) {
    if (copy_rheo) {
        F.rheo[c] = F.rheo[src];
        F.lmu_cell[c] = F.lmu_cell[src];
        if (cell_n_on(F)) F.n_cell[c] = F.n_cell[src];
    }
    float4 ts = F.strata_h[src];
    uint32_t ks = 0;
    if (ts.x > 0.0f) ks = 0;
    if (ts.y > 0.0f) ks = 1;
    if (ts.z > 0.0f) ks = 2;
    if (ts.w > 0.0f) ks = 3;
    float4 cs4 = F.strata_cure[src];
    float cs = ks == 0 ? cs4.x : (ks == 1 ? cs4.y : (ks == 2 ? cs4.z : cs4.w));
    F.strata_h[c] = make_float4(hn, 0.0f, 0.0f, 0.0f);
    uint4 a = F.strata_pig[(src * 4u + ks) * 2u];
    uint4 b = F.strata_pig[(src * 4u + ks) * 2u + 1u];
    F.strata_pig[c * 8u] = a;
    F.strata_pig[c * 8u + 1u] = b;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t k = 2u; k < 8u; k++) F.strata_pig[c * 8u + k] = z;
    F.strata_cure[c] = make_float4(cs, 0.0f, 0.0f, 0.0f);
}

// limiter + apply (film.wgsl:521-537, 803-863) on a tile.  apply reads the
// limiter factor of the cell and of its 8 neighbours, so the block evaluates
// the limiter for the tile and a 1-cell ring (active cells) into LDS; the
// factors never leave the chip.  Fluxes and heights of the tile and a 2-cell
// ring (what the ring's limiter reads) are staged in LDS first; the heights
// come from the copy the lap pass wrote (h_sub), since the blocks of the
// neighbouring tiles update h during this pass.
constexpr int AP_R = 2;
constexpr int AP_WIN = 16 + 2 * AP_R;

struct ApplyShared {
    float4 s_f[AP_WIN * AP_WIN];
    float s_hh[AP_WIN * AP_WIN];
    float s_a[RING_SPAN * RING_SPAN];
    uint32_t s_tf[9];
};

// The limiter + apply of a tile from the staged fluxes and heights of the tile
// and a 2-cell ring (S.s_f, S.s_hh) and the 3x3 tile flags (S.s_tf).
template <int BT = 256>
__device__ __forceinline__ void apply_core(const FilmDev& F, uint32_t _t, ApplyShared& S) {
    float4* s_f = S.s_f;
    float* s_hh = S.s_hh;
    float* s_a = S.s_a;
    uint32_t* s_tf = S.s_tf;
    const uint32_t _lx = threadIdx.x & 15u, _ly = (threadIdx.x >> 4) & 15u;
    CZ_TILE_CELL(F)
    const bool cell_th = BT == 256 || threadIdx.x < 256u;
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int tx0 = ttx * 16, ty0 = tty * 16;
    const int wx0 = tx0 - AP_R, wy0 = ty0 - AP_R;
    const int W_ = (int)F.W, H_ = (int)F.H;
    auto wq = [&](int gx, int gy) -> int { return (gy - wy0) * AP_WIN + (gx - wx0); };
    // in the canvas and in an active tile (cells within 2 of the tile)
    auto nb_ok = [&](int gi, int gj) -> bool {
        if (gi < 0 || gj < 0 || gi >= W_ || gj >= H_) return false;
        return (s_tf[((gj >> 4) - tty + 1) * 3 + ((gi >> 4) - ttx + 1)] & 32u) != 0u;
    };
    // limiter (film.wgsl:521-537) of cell (gi, gj)
    auto limiter = [&](int gi, int gj) -> float {
        const float4 f = s_f[wq(gi, gj)];
        float out = cz_max(f.x, 0.0f) + cz_max(f.y, 0.0f) + cz_max(f.z, 0.0f) + cz_max(f.w, 0.0f);
        if (nb_ok(gi - 1, gj)) out = out + cz_max(-s_f[wq(gi - 1, gj)].x, 0.0f);
        if (nb_ok(gi, gj - 1)) out = out + cz_max(-s_f[wq(gi, gj - 1)].y, 0.0f);
        if (nb_ok(gi - 1, gj - 1)) out = out + cz_max(-s_f[wq(gi - 1, gj - 1)].z, 0.0f);
        if (nb_ok(gi - 1, gj + 1)) out = out + cz_max(-s_f[wq(gi - 1, gj + 1)].w, 0.0f);
        float a = 1.0f;
        if (out > 0.0f) a = cz_min(1.0f, 0.5f * s_hh[wq(gi, gj)] * F.P.dx / (F.P.dt * out));
        return a;
    };
    for (uint32_t q = threadIdx.x; q < (uint32_t)(RING_SPAN * RING_SPAN); q += (uint32_t)BT) {
        const int gx = tx0 - 1 + (int)(q % RING_SPAN), gy = ty0 - 1 + (int)(q / RING_SPAN);
        s_a[q] = nb_ok(gx, gy) ? limiter(gx, gy) : 0.0f;
    }
    __syncthreads();
    bool moved = false;
    if (in_grid && cell_th) {
        const int i = (int)x, j = (int)y;
        auto alpha_at = [&](int gi, int gj) -> float { return s_a[(gj - ty0 + 1) * RING_SPAN + (gi - tx0 + 1)]; };
        const float ac = alpha_at(i, j);
        Inflow st = {0.0f, c, 0.0f, 0.0f};
        const float4 f = s_f[wq(i, j)];
        const bool hx = x + 1u < F.W, hy = y + 1u < F.H, hm = y > 0u;
        const int xp = (x + 1u < F.W - 1u) ? i + 1 : W_ - 1;
        const int yp = (y + 1u < F.H - 1u) ? j + 1 : H_ - 1;
        const int ym = hm ? j - 1 : 0;
        float accv[4];
        const float fq[4] = {f.x, f.y, f.z, f.w};
        const bool has[4] = {hx, hy, hx && hy, hx && hm};
        const int bi[4] = {hx ? xp : i, hx ? i : i, (hx && hy) ? xp : i, (hx && hm) ? xp : i};
        const int bj[4] = {j, hy ? yp : j, (hx && hy) ? yp : j, (hx && hm) ? ym : j};
        // own_face (film.wgsl:546-554)
        for (int k = 0; k < 4; k++) {
            const float q = fq[k];
            float ab = ac;
            if (q < 0.0f && has[k]) {
                ab = alpha_at(bi[k], bj[k]);
                if (-q * ab > st.best) { st.best = -q * ab; st.src = idx(F, (uint32_t)bi[k], (uint32_t)bj[k]); }
            }
            st.div = st.div + q * ab;
            accv[k] = q * ab;
        }
        // nb_face (film.wgsl:556-561)
        const int ni[4] = {i - 1, i, i - 1, i - 1};
        const int nj[4] = {j, j - 1, j - 1, j + 1};
        for (int k = 0; k < 4; k++) {
            if (!nb_ok(ni[k], nj[k])) continue;
            const float4 fn = s_f[wq(ni[k], nj[k])];
            const float q = (k == 0) ? fn.x : (k == 1) ? fn.y : (k == 2) ? fn.z : fn.w;
            float an = alpha_at(ni[k], nj[k]);
            if (q < 0.0f) an = ac;
            if (q * an > st.best) { st.best = q * an; st.src = idx(F, (uint32_t)ni[k], (uint32_t)nj[k]); }
            st.div = st.div - q * an;
        }
        const uint32_t src = st.src;
        const float h0 = s_hh[wq(i, j)];
        float hn = cz_max(h0 - F.P.dt * st.div * F.P.inv_dx, 0.0f);
        if (hn < h0 && hn < DRAINED_FILM) hn = 0.0f;
        F.h[c] = hn;
        moved = cz_abs(hn - h0) > MOVE_EPS;
        if (h0 <= 0.0f && src != c) inherit_from(F, c, src, hn, true);
        const float4 a = F.acc[c];
        F.acc[c] = make_float4(a.x + accv[0] * F.P.dt, a.y + accv[1] * F.P.dt, a.z + accv[2] * F.P.dt,
                               a.w + accv[3] * F.P.dt);
    }
    if (__syncthreads_or(moved ? 1 : 0) && threadIdx.x == 0) F.tile_moved[_t] = 1u;
}

// One tile of the limiter + apply pass (every thread of the block calls it).
template <int BT = 256>
__device__ __forceinline__ void apply_tile(const FilmDev& F, uint32_t _t, ApplyShared& S) {
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int wx0 = ttx * 16 - AP_R, wy0 = tty * 16 - AP_R;
    const int W_ = (int)F.W, H_ = (int)F.H;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < (uint32_t)(AP_WIN * AP_WIN); q += (uint32_t)BT) {
        const int gx = wx0 + (int)(q % AP_WIN), gy = wy0 + (int)(q / AP_WIN);
        if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
            const uint32_t nn = idx(F, (uint32_t)gx, (uint32_t)gy);
            S.s_f[q] = F.flux[nn];
            S.s_hh[q] = F.h_sub[nn];
        }
    }
    if (threadIdx.x < 9u) {
        const int ax = ttx - 1 + (int)(threadIdx.x % 3u), ay = tty - 1 + (int)(threadIdx.x / 3u);
        S.s_tf[threadIdx.x] = (ax >= 0 && ay >= 0 && ax < (int)F.tw && ay < (int)F.th)
                                  ? F.tile_flags[(uint32_t)ax + (uint32_t)ay * F.tw] : 0u;
    }
    __syncthreads();
    apply_core<BT>(F, _t, S);
}

// ---------------------------------------------------------------------------
// slump: cone-cap shedding after each substep's nose, on when
// settings.slump (Params.flags[1]; film.wgsl:563-801, 866-876, encoded after
// nose by src/film/mod.rs:1422-1429).  Two launches: k_shed stores each
// active cell's shed in acc_s (as the shader does; strata_split rewrites
// acc_s at the frame's end), then k_slump gathers what the 8 neighbours shed
// and commits the new height at once: a cell's slump reads no other cell's
// height, so commit_tmp's separate pass is not needed (tmp still receives the
// height, as the shader leaves it).

// film.wgsl:563-568: the steepest slope the paint holds at rest
__device__ __forceinline__ float slope_cap(const FilmDev& F, uint32_t c) {
    const float ty = F.rheo[c].y;
    const float hb = cz_max(F.h[c], 1e-4f);
    const float need = cz_max(ty / (F.P.phys[0] * hb) - cz_len2(F.P.g[0], F.P.g[1]), 0.0f);
    return cz_max(F.P.ein[0], need / cz_max(F.P.g[2], 1e-3f));
}

// film.wgsl:587-596: the height above the cone through the nearest dry cell
__device__ __forceinline__ float shed_excess(const FilmDev& F, uint32_t c) {
    const float hc = F.h[c];
    const float hmin = 2.0f * F.P.ein[1];
    if (hc <= hmin) return 0.0f;
    const float cap = slope_cap(F, c) * F.P.dx * cz_max(nose_dist(F, c), 1.0f);
    return cz_max(hc - cap, 0.0f);
}

// film.wgsl:598-608
__device__ __forceinline__ float hat(const FilmDev& F, int i, int j) {
    if (i < 0 || j < 0 || i >= (int)F.W || j >= (int)F.H) return 0.0f;
    if (!cell_active(F, (uint32_t)i, (uint32_t)j)) return 0.0f;
    return F.h[idx(F, (uint32_t)i, (uint32_t)j)];
}
__device__ __forceinline__ float okf(const FilmDev& F, int i, int j) {
    if (i < 0 || j < 0 || i >= (int)F.W || j >= (int)F.H) return 0.0f;
    return cell_active(F, (uint32_t)i, (uint32_t)j) ? 1.0f : 0.0f;
}

struct Shed {
    float amount;
    int i0, j0;
    float4 w;
};

// film.wgsl:621-668: the excess, at most half the drop to the bilinear point
// one cell along the outward direction (the offsets to the dry cells within
// radius sqrt 5, else down the height gradient), over its four cells
__device__ Shed shed_of(const FilmDev& F, int i, int j) {
    Shed o = {0.0f, i, j, make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
    if (!cell_active(F, (uint32_t)i, (uint32_t)j)) return o;
    const uint32_t c = idx(F, (uint32_t)i, (uint32_t)j);
    const float e = shed_excess(F, c);
    if (e <= 0.0f) return o;
    const float hc = F.h[c];
    float dx = 0.0f, dy = 0.0f;
    for (int dj = -2; dj <= 2; dj++)
        for (int di = -2; di <= 2; di++)
            if (di * di + dj * dj <= 5 && hat(F, i + di, j + dj) <= 0.0f) {
                dx = dx + (float)di;
                dy = dy + (float)dj;
            }
    if (cz_len2(dx, dy) < 1e-6f) {
        dx = -(0.5f * (hat(F, i + 1, j) - hat(F, i - 1, j)));
        dy = -(0.5f * (hat(F, i, j + 1) - hat(F, i, j - 1)));
    }
    const float gm = cz_len2(dx, dy);
    if (gm < 1e-9f) return o;
    dx = dx / gm;
    dy = dy / gm;
    const float px = (float)i + dx, py = (float)j + dy;
    const float fx0 = floorf(px), fy0 = floorf(py);
    const float tx = px - fx0, ty = py - fy0;
    const int i0 = (int)fx0, j0 = (int)fy0;
    float4 w = make_float4((1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty);
    w = make_float4(w.x * okf(F, i0, j0), w.y * okf(F, i0 + 1, j0), w.z * okf(F, i0, j0 + 1), w.w * okf(F, i0 + 1, j0 + 1));
    const float ws = w.x + w.y + w.z + w.w;
    if (ws <= 1e-6f) return o;
    w = make_float4(w.x / ws, w.y / ws, w.z / ws, w.w / ws);
    const float ht = w.x * hat(F, i0, j0) + w.y * hat(F, i0 + 1, j0) + w.z * hat(F, i0, j0 + 1) + w.w * hat(F, i0 + 1, j0 + 1);
    const float amount = cz_min(e, 0.5f * cz_max(hc - ht, 0.0f));
    if (amount <= 0.0f) return o;
    o.amount = amount;
    o.i0 = i0;
    o.j0 = j0;
    o.w = w;
    return o;
}

// film.wgsl:670-678
__device__ __forceinline__ float share_to(const Shed& sd, int i, int j) {
    if (sd.amount <= 0.0f) return 0.0f;
    const int di = i - sd.i0, dj = j - sd.j0;
    if (di == 0 && dj == 0) return sd.amount * sd.w.x;
    if (di == 1 && dj == 0) return sd.amount * sd.w.y;
    if (di == 0 && dj == 1) return sd.amount * sd.w.z;
    if (di == 1 && dj == 1) return sd.amount * sd.w.w;
    return 0.0f;
}

// film.wgsl:725-741: acc_s[4c] = (amount, i0, j0, 0), acc_s[4c + 1] = w
__device__ __forceinline__ Shed load_shed(const FilmDev& F, uint32_t c) {
    const float4 a = F.acc_s[4u * c];
    Shed sd;
    sd.amount = a.x;
    sd.i0 = (int)cz_f2bits(a.y);
    sd.j0 = (int)cz_f2bits(a.z);
    sd.w = F.acc_s[4u * c + 1u];
    return sd;
}

// film.wgsl:743-748
template <class WS>
__global__ __launch_bounds__(256) void k_shed(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        const Shed sd = shed_of(F, (int)x, (int)y);
        F.acc_s[4u * c] = make_float4(sd.amount, cz_bits2f((uint32_t)sd.i0), cz_bits2f((uint32_t)sd.j0), 0.0f);
        F.acc_s[4u * c + 1u] = sd.w;
    }
}

// film.wgsl:750-801 and commit_tmp (866-876)
template <class WS>
__global__ __launch_bounds__(256) void k_slump(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        bool moved = false;
        if (in_grid) {
            const int i = (int)x, j = (int)y;
            const float hc = F.h[c];
            const Shed mine = load_shed(F, c);
            const float self_share = share_to(mine, i, j);
            float inn = 0.0f, best_in = 0.0f;
            uint32_t src = c;
            float fx = 0.0f, fy = 0.0f, fd = 0.0f, fe = 0.0f;
            for (int dj = -1; dj <= 1; dj++) {
                for (int di = -1; di <= 1; di++) {
                    if (di == 0 && dj == 0) continue;
                    const int xx = i + di, yy = j + dj;
                    if (xx < 0 || yy < 0 || xx >= (int)F.W || yy >= (int)F.H) continue;
                    if (!cell_active(F, (uint32_t)xx, (uint32_t)yy)) continue;
                    const uint32_t nn = idx(F, (uint32_t)xx, (uint32_t)yy);
                    const Shed sn = load_shed(F, nn);
                    const float give = share_to(sn, i, j);
                    const float take = share_to(mine, xx, yy);
                    inn = inn + give;
                    if (give > best_in) { best_in = give; src = nn; }
                    const float net = take - give;
                    if (di == 1 && dj == 0) fx = net;
                    if (di == 0 && dj == 1) fy = net;
                    if (di == 1 && dj == 1) fd = net;
                    if (di == 1 && dj == -1) fe = net;
                }
            }
            const float outs = mine.amount - self_share;
            const float hn2 = cz_max(hc - outs + inn, 0.0f);
            F.tmp[c] = hn2;
            const float4 a = F.acc[c];
            F.acc[c] = make_float4(a.x + fx * F.P.dx, a.y + fy * F.P.dx, a.z + fd * F.P.dx, a.w + fe * F.P.dx);
            // the donor is wet (it sheds), so no other cell rewrites its
            // rheology or strata during this pass
            if (hc <= 0.0f && hn2 > 0.0f && src != c) inherit_from(F, c, src, hn2, true);
            moved = cz_abs(hn2 - hc) > MOVE_EPS;
            F.h[c] = hn2;
        }
        if (__syncthreads_or(moved ? 1 : 0) && threadIdx.x == 0) F.tile_moved[_t] = 1u;
    }
}

// fluxes + limiter + apply of a tile in one launch, for small active sets
// where the launch between the two passes costs more than the work: the
// block computes the fluxes of the tile and a 2-cell ring itself (the cells
// the tile's limiter ring reads) from a window of the tile and a 6-cell ring,
// instead of reading them from the flux pass.  Other blocks apply while this
// one computes, so the window's heights are the ones before this substep's
// apply (h_sub on active tiles; elsewhere h, which no apply writes).  The
// rheology apply may write (a newly wetted cell inheriting) belongs to a cell
// that was dry before the apply, whose rheology enters the fluxes only with
// weight 0 (film.wgsl:324-327), so its live value gives the same fluxes.
constexpr int FA_R = AP_R + FLUX_R;

struct FluxApplyShared {
    FluxWinT<FA_R> w;
    ApplyShared ap;
};

template <int BT = 256>
__device__ __forceinline__ void flux_apply_tile(const FilmDev& F, uint32_t _t, FluxApplyShared& S) {
    const uint32_t _lx = threadIdx.x & 15u, _ly = (threadIdx.x >> 4) & 15u;
    CZ_TILE_CELL(F)
    const bool cell_th = BT == 256 || threadIdx.x < 256u;
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int wx0 = ttx * 16 - FA_R, wy0 = tty * 16 - FA_R;
    const int W_ = (int)F.W, H_ = (int)F.H;
    flux_stage<FA_R, true, BT>(F, S.w, ttx, tty, wx0, wy0);
    if (threadIdx.x < 9u) S.ap.s_tf[threadIdx.x] = S.w.tf[threadIdx.x];
    const WinReadT<FA_R, true> g{F, S.w, wx0, wy0, ttx, tty, F.k_hold1, F.k_theta};
    // the fluxes of the active cells of the tile + 2 ring into the apply's window
    if (BT == 256) {
        for (uint32_t q = threadIdx.x; q < (uint32_t)(AP_WIN * AP_WIN); q += 256u) {
            const int gx = ttx * 16 - AP_R + (int)(q % AP_WIN), gy = tty * 16 - AP_R + (int)(q / AP_WIN);
            if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
                S.ap.s_hh[q] = S.w.h[g.q(gx, gy)];
                float4 qv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (g.active(gx, gy)) qv = cell_fluxes(F, g, gx, gy);
                S.ap.s_f[q] = qv;
            }
        }
    } else {
        // one face of one window cell per item, as cell_fluxes evaluates it
        for (uint32_t it = threadIdx.x; it < (uint32_t)(AP_WIN * AP_WIN * 4); it += (uint32_t)BT) {
            const uint32_t q = it >> 2, k = it & 3u;
            const int gx = ttx * 16 - AP_R + (int)(q % AP_WIN), gy = tty * 16 - AP_R + (int)(q / AP_WIN);
            if (gx >= 0 && gy >= 0 && gx < W_ && gy < H_) {
                if (k == 0u) S.ap.s_hh[q] = S.w.h[g.q(gx, gy)];
                float qk = 0.0f;
                if (g.active(gx, gy)) {
                    const float2 p00 = pressure_at(F, g, gx, gy);
                    FaceGeom fg = face_geom(F, g, gx, gy, (int)k, true, p00);
                    if (fg.has) {
                        FaceFlux ff = face(F, g, gx, gy, fg.ib, fg.jb, fg.gcap, fg.ghyd, fg.gtilt, fg.gtn, fg.diag, fg.dir);
                        qk = ff.q * ((k < 2u) ? W_AXIS : W_DIAG_FLUX);
                    }
                }
                ((float*)&S.ap.s_f[q])[k] = qk;
            }
        }
    }
    __syncthreads();
    if (in_grid && cell_th) F.flux[c] = S.ap.s_f[(int)(_ly + AP_R) * AP_WIN + (int)(_lx + AP_R)];
    apply_core<BT>(F, _t, S.ap);
}

template <class WS>
__global__ __launch_bounds__(256) void k_flux_apply(WS ws) {
    __shared__ FluxApplyShared S;
    CZ_FOR_TILES(ws) flux_apply_tile(F, _t, S);
}

template <class WS>
__global__ __launch_bounds__(1024) void k_flux_apply_w(WS ws) {
    __shared__ FluxApplyShared S;
    CZ_FOR_TILES(ws) flux_apply_tile<1024>(F, _t, S);
}

template <class WS>
__global__ __launch_bounds__(256) void k_limit_apply(WS ws) {
    __shared__ ApplyShared S;
    CZ_FOR_TILES(ws) apply_tile(F, _t, S);
}

template <class WS>
__global__ __launch_bounds__(1024) void k_limit_apply_w(WS ws) {
    __shared__ ApplyShared S;
    CZ_FOR_TILES(ws) apply_tile<1024>(F, _t, S);
}

// ---------------------------------------------------------------------------
// A whole substep of a tile in one launch: [nose of the previous substep] +
// laplacian + plug + fluxes + limiter + apply (film.wgsl:16-87, 375-455,
// 521-537, 680-723, 803-863), for small active sets where the launch between
// the lap/plug pass and the flux + apply pass costs more than recomputing a
// wider ring.  The apply of the tile needs fluxes of the tile + 2 ring, whose
// window (tile + 6) needs pressures and plug means there, which need the
// nose of the tile + 7 ring and per-cell plug fluxes of the tile + 9 ring:
// the block stages that 34x34 window and computes every ring itself.
//
// Blocks of one launch apply while others read, so the heights before this
// substep are read from hr (cells of active tiles; inactive tiles keep F.h,
// which no apply writes) and the new heights written to hw; the caller
// alternates the two buffers over the frame's substeps and ends in F.h.  The
// offsets are read from F.delta (before the nose) and written to
// F.delta_out; the flux terms read, like the separate flux pass after the
// offset flip, the new offsets of active cells (from the block's nose, or
// recomputed from hr and F.delta for the far steps of cell_excess beyond the
// window) and F.delta_out at cells of inactive tiles.
constexpr int SB_R = 9, SB_N = 16 + 2 * SB_R;   // staged window
constexpr int SN_R = 7, SN_N = 16 + 2 * SN_R;   // nose, usurf
constexpr int SL_R = AP_R + FLUX_R;             // lap, plug means, flux window

struct SubstepWin {
    float h[SB_N * SB_N];
    float m[SB_N * SB_N];
    uint32_t d[SB_N * SB_N];
    float2 rheo[SB_N * SB_N];
    float2 fn[SB_N * SB_N];
    float2 hq[SB_N * SB_N];       // plug: height (negated with a yield stress), hb_flux
    uint32_t nd[SN_N * SN_N];     // offsets the laplacian reads (new on active cells)
    float2 nfn[SN_N * SN_N];      // front normals likewise
    float us[SN_N * SN_N];        // usurf
    uint32_t tf[9];
};

struct SubstepFusedShared {
    SubstepWin s;
    FluxWinT<SL_R> w;
    ApplyShared ap;
};

// heights before this substep's apply at any canvas cell
__device__ __forceinline__ float h_before(const FilmDev& F, const float* hr, int i, int j) {
    const uint32_t c = idx(F, (uint32_t)i, (uint32_t)j);
    return cell_active(F, (uint32_t)i, (uint32_t)j) ? hr[c] : F.h[c];
}

// nose (film.wgsl:680-723) of one active cell from HBM: the offset only
__device__ uint32_t nose_offset_global(const FilmDev& F, const float* hr, int i, int j) {
    const int W_ = (int)F.W, H_ = (int)F.H;
    if (h_before(F, hr, i, j) <= 0.0f) return 0u;
    int bx = 64, by = 64;
    int bd = kNoseNone;
    for (int dj = -1; dj <= 1; dj++) {
        for (int di = -1; di <= 1; di++) {
            const int xx = i + di, yy = j + dj;
            if (xx < 0 || yy < 0 || xx >= W_ || yy >= H_) continue;
            int cx = di, cy = dj;
            if (h_before(F, hr, xx, yy) > 0.0f) {
                const uint32_t dv = F.delta[idx(F, (uint32_t)xx, (uint32_t)yy)];
                if ((dv & 0x10000u) == 0u) continue;
                int ox, oy;
                unpack_off(dv, ox, oy);
                cx = ox + di;
                cy = oy + dj;
                const int tx = i + cx, ty = j + cy;
                if (tx < 0 || ty < 0 || tx >= W_ || ty >= H_) continue;
                if (h_before(F, hr, tx, ty) > 0.0f) continue;
            } else if (di == 0 && dj == 0) {
                continue;
            }
            const int d = cx * cx + cy * cy;
            if (d < bd) { bd = d; bx = cx; by = cy; }
        }
    }
    if (bd == kNoseNone) return 0u;
    return pack_off(bx, by);
}

// The flux window reader of the fused substep: as WinReadT, with the far
// steps of cell_excess served as described above.
struct FusedRead {
    static constexpr int N = FluxWinT<SL_R>::N;
    // a far offset may be a nose evaluated in HBM: cell_excess reads the far
    // steps one at a time and stops at the first dry one
    static constexpr bool kFarCheap = false;
    const FilmDev& F;
    const FluxWinT<SL_R>& w;
    const float* hr;
    int wx0, wy0, ttx, tty;
    float hold1_, theta_;
    int do_nose;
    __device__ __forceinline__ int q(int i, int j) const { return (j - wy0) * N + (i - wx0); }
    __device__ __forceinline__ float h(int i, int j) const { return w.h[q(i, j)]; }
    __device__ __forceinline__ uint32_t delta(int i, int j) const { return w.d[q(i, j)]; }
    __device__ __forceinline__ float2 rheo(int i, int j) const { return w.rheo[q(i, j)]; }
    __device__ __forceinline__ float2 fnorm(int i, int j) const { return w.fn[q(i, j)]; }
    __device__ __forceinline__ float2 pres(int i, int j) const { return w.pres[q(i, j)]; }
    __device__ __forceinline__ float tmp(int i, int j) const { return w.tmp[q(i, j)]; }
    __device__ __forceinline__ float lmu(int i, int j) const { return w.lmu[q(i, j)]; }
    __device__ __forceinline__ float fdist(int i, int j) const { return w.fd[q(i, j)]; }
    __device__ __forceinline__ float ncell(int i, int j) const { return F.n_cell[(uint32_t)j * F.W + (uint32_t)i]; }
    __device__ __forceinline__ float hold1() const { return hold1_; }
    __device__ __forceinline__ float theta() const { return theta_; }
    __device__ __forceinline__ bool in(int i, int j) const {
        return (unsigned)(i - wx0) < (unsigned)N && (unsigned)(j - wy0) < (unsigned)N;
    }
    __device__ __forceinline__ float h_far(int i, int j) const {
        if (in(i, j)) return w.h[q(i, j)];
        return h_before(F, hr, i, j);
    }
    __device__ __forceinline__ uint32_t delta_far(int i, int j) const {
        if (in(i, j)) return w.d[q(i, j)];
        const uint32_t c = (uint32_t)j * F.W + (uint32_t)i;
        if (!do_nose) return F.delta[c];
        return cell_active(F, (uint32_t)i, (uint32_t)j) ? nose_offset_global(F, hr, i, j) : F.delta_out[c];
    }
    __device__ __forceinline__ bool active(int i, int j) const {
        return (w.tf[((j >> 4) - tty + 1) * 3 + ((i >> 4) - ttx + 1)] & 32u) != 0u;
    }
};

__device__ __forceinline__ void substep_fused_tile(const FilmDev& F, const float* hr, float* hw, uint32_t _t,
                                                   SubstepFusedShared& S, int do_nose, int do_clear) {
    constexpr int BT = 1024;
    SubstepWin& sw = S.s;
    FluxWinT<SL_R>& fw = S.w;
    const uint32_t _lx = threadIdx.x & 15u, _ly = (threadIdx.x >> 4) & 15u;
    CZ_TILE_CELL(F)
    const bool cell_th = threadIdx.x < 256u;
    const float gb = F.k_gb_plug;
    const bool plug_on = !(gb < 1e-6f);
    const int ttx = (int)(_t % F.tw), tty = (int)(_t / F.tw);
    const int tx0 = ttx * 16, ty0 = tty * 16;
    const int bx0 = tx0 - SB_R, by0 = ty0 - SB_R;     // staged window origin
    const int nx0 = tx0 - SN_R, ny0 = ty0 - SN_R;     // nose region origin
    const int lx0 = tx0 - SL_R, ly0 = ty0 - SL_R;     // lap / flux window origin
    const int W_ = (int)F.W, H_ = (int)F.H;
    __syncthreads();
    if (threadIdx.x < 9u) {
        const int ax = ttx - 1 + (int)(threadIdx.x % 3u), ay = tty - 1 + (int)(threadIdx.x / 3u);
        const uint32_t fl = (ax >= 0 && ay >= 0 && ax < (int)F.tw && ay < (int)F.th)
                                ? F.tile_flags[(uint32_t)ax + (uint32_t)ay * F.tw] : 0u;
        sw.tf[threadIdx.x] = fl;
        fw.tf[threadIdx.x] = fl;
        S.ap.s_tf[threadIdx.x] = fl;
    }
    __syncthreads();
    auto act_w = [&](int gx, int gy) -> bool {
        return (sw.tf[((gy >> 4) - tty + 1) * 3 + ((gx >> 4) - ttx + 1)] & 32u) != 0u;
    };
    auto in_canvas = [&](int gx, int gy) -> bool { return gx >= 0 && gy >= 0 && gx < W_ && gy < H_; };
    // stage the window: heights before the apply, relief, offsets before the
    // nose, rheology, stored front normals
    for (uint32_t q = threadIdx.x; q < (uint32_t)(SB_N * SB_N); q += (uint32_t)BT) {
        const int gx = bx0 + (int)(q % SB_N), gy = by0 + (int)(q / SB_N);
        if (in_canvas(gx, gy)) {
            const uint32_t nn = idx(F, (uint32_t)gx, (uint32_t)gy);
            sw.h[q] = act_w(gx, gy) ? hr[nn] : F.h[nn];
            sw.m[q] = F.m[nn];
            sw.d[q] = F.delta[nn];
            sw.rheo[q] = F.rheo[nn];
            sw.fn[q] = F.fnorm[nn];
        } else {
            sw.h[q] = 0.0f;
            sw.m[q] = 0.0f;
            sw.d[q] = 0u;
            sw.rheo[q] = make_float2(0.0f, 0.0f);
            sw.fn[q] = make_float2(0.0f, 0.0f);
        }
    }
    __syncthreads();
    auto bq = [&](int gx, int gy) -> int { return (gy - by0) * SB_N + (gx - bx0); };
    auto in_b = [&](int gx, int gy) -> bool {
        return gx >= bx0 && gy >= by0 && gx < bx0 + SB_N && gy < by0 + SB_N;
    };
    // nose of cell (i, j) from the window (as lap_plug_tile's)
    auto nose_w = [&](int i, int j, uint32_t& d_out, float2& fn_out) {
        d_out = 0u;
        fn_out = make_float2(0.0f, 0.0f);
        if (sw.h[bq(i, j)] <= 0.0f) return;
        int bxx = 64, byy = 64;
        int bd = kNoseNone;
        for (int dj = -1; dj <= 1; dj++) {
            for (int di = -1; di <= 1; di++) {
                const int xx = i + di, yy = j + dj;
                if (xx < 0 || yy < 0 || xx >= W_ || yy >= H_) continue;
                int cx = di, cy = dj;
                if (sw.h[bq(xx, yy)] > 0.0f) {
                    const uint32_t dv = sw.d[bq(xx, yy)];
                    if ((dv & 0x10000u) == 0u) continue;
                    int ox, oy;
                    unpack_off(dv, ox, oy);
                    cx = ox + di;
                    cy = oy + dj;
                    const int tx = i + cx, ty = j + cy;
                    if (tx < 0 || ty < 0 || tx >= W_ || ty >= H_) continue;
                    const float ht = in_b(tx, ty) ? sw.h[bq(tx, ty)] : h_before(F, hr, tx, ty);
                    if (ht > 0.0f) continue;
                } else if (di == 0 && dj == 0) {
                    continue;
                }
                const int d = cx * cx + cy * cy;
                if (d < bd) { bd = d; bxx = cx; byy = cy; }
            }
        }
        if (bd == kNoseNone) return;
        d_out = pack_off(bxx, byy);
        if (bd <= 9) {
            float dx = 0.0f, dy = 0.0f;
            for (int dj = -2; dj <= 2; dj++) {
                for (int di = -2; di <= 2; di++) {
                    if (di == 0 && dj == 0) continue;
                    const int xx = i + di, yy = j + dj;
                    if (xx < 0 || yy < 0 || xx >= W_ || yy >= H_) continue;
                    if (di * di + dj * dj > 5) continue;
                    if (sw.h[bq(xx, yy)] <= F.P.ein[3]) {
                        dx = dx + FN_QUOT.v[di + 2][dj + 2];
                        dy = dy + FN_QUOT.v[dj + 2][di + 2];
                    }
                }
            }
            const float l = cz_len2(dx, dy);
            fn_out = (l < 1e-6f) ? make_float2(0.0f, 0.0f) : make_float2(dx / l, dy / l);
        }
    };
    // nose region: the offsets and normals the laplacian reads, usurf
    for (uint32_t q = threadIdx.x; q < (uint32_t)(SN_N * SN_N); q += (uint32_t)BT) {
        const int gx = nx0 + (int)(q % SN_N), gy = ny0 + (int)(q / SN_N);
        uint32_t dv = 0u;
        float2 fn = make_float2(0.0f, 0.0f);
        float us = 0.0f;
        if (in_canvas(gx, gy)) {
            if (do_nose && act_w(gx, gy)) {
                nose_w(gx, gy, dv, fn);
            } else {
                dv = sw.d[bq(gx, gy)];
                fn = sw.fn[bq(gx, gy)];
            }
            const float hc = sw.h[bq(gx, gy)];
            float hold = 0.0f;
            if (hc > 0.0f) hold = hold_level(F, nose_dist_b(dv));
            us = hc + sw.m[bq(gx, gy)] - hold;
        }
        sw.nd[q] = dv;
        sw.nfn[q] = fn;
        sw.us[q] = us;
    }
    // plug: hb_flux once per window cell
    if (plug_on) {
        for (uint32_t q = threadIdx.x; q < (uint32_t)(SB_N * SB_N); q += (uint32_t)BT) {
            float hn = sw.h[q], qv = 0.0f;
            bool yl = false;
            if (hn > 0.0f) {
                const float2 r = sw.rheo[q];
                yl = r.y > 0.0f;
                float oy_, oup_;
                float nq = F.P.phys[2];
                if (cell_n_on(F))
                    nq = F.n_cell[idx(F, (uint32_t)(bx0 + (int)(q % SB_N)), (uint32_t)(by0 + (int)(q / SB_N)))];
                const float lmq = F.lmu_cell[idx(F, (uint32_t)(bx0 + (int)(q % SB_N)), (uint32_t)(by0 + (int)(q / SB_N)))];
                qv = hb_flux(nq, gb, true, F.k_lg_plug, hn, Visc{cz_max(r.x, 1e-4f), lmq, true}, r.y, oy_, oup_);
            } else {
                hn = 0.0f;
            }
            sw.hq[q] = make_float2(yl ? -hn : hn, qv);
        }
    }
    __syncthreads();
    auto nqi = [&](int gx, int gy) -> int { return (gy - ny0) * SN_N + (gx - nx0); };
    // laplacian, pressures and plug means over the flux window (active
    // cells; cells of inactive tiles keep their stored values), and the flux
    // window's other fields
    constexpr int LN = FluxWinT<SL_R>::N;
    for (uint32_t q = threadIdx.x; q < (uint32_t)(LN * LN); q += (uint32_t)BT) {
        const int i = lx0 + (int)(q % LN), j = ly0 + (int)(q / LN);
        if (!in_canvas(i, j)) continue;
        const uint32_t c = idx(F, (uint32_t)i, (uint32_t)j);
        const bool act = act_w(i, j);
        const float hv = sw.h[bq(i, j)];
        const uint32_t dl = sw.nd[nqi(i, j)];
        fw.h[q] = hv;
        fw.rheo[q] = sw.rheo[bq(i, j)];
        fw.lmu[q] = F.lmu_cell[c];
        const uint32_t dflux = act ? dl : (do_nose ? F.delta_out[c] : sw.d[bq(i, j)]);
        fw.d[q] = dflux;
        fw.fn[q] = sw.nfn[nqi(i, j)];
        fw.fd[q] = hv <= 0.0f ? 0.0f : nose_dist_b(dflux);
        if (!act) {
            fw.pres[q] = F.pres[c];
            fw.tmp[q] = F.tmp[c];
            continue;
        }
        auto usurf_l = [&](int ii, int jj) -> float {
            return sw.us[nqi((int)refl1_i(F, ii), (int)refl1_j(F, jj))];
        };
        const float s0 = hv + sw.m[bq(i, j)];
        float hold = 0.0f, pl = 0.0f;
        const float dself = nose_dist_b(dl);
        if (hv > 0.0f) hold = hold_level(F, dself);
        const float u0 = s0 - hold;
        const float ax = usurf_l(i - 1, j) + usurf_l(i + 1, j) + usurf_l(i, j - 1) + usurf_l(i, j + 1);
        const float dg = usurf_l(i - 1, j - 1) + usurf_l(i + 1, j - 1) + usurf_l(i - 1, j + 1) + usurf_l(i + 1, j + 1);
        const float l = (W_AXIS * ax + W_DIAG_LAP * dg - (4.0f * W_AXIS + 4.0f * W_DIAG_LAP) * u0) * F.P.inv_dx * F.P.inv_dx;
        if (hv > 0.0f) {
            const float d = dself;
            if (d <= 3.0f) {
                const int own = nqi(i, j);
                const float2 nn = sw.nfn[own];
                if (nn.x * nn.x + nn.y * nn.y > 0.5f) {
                    float lp = 0.0f;
                    if (!(i < 1 || j < 1 || i + 1 >= W_ || j + 1 >= H_)) {
                        const float2 nl = sw.nfn[own - 1], nr = sw.nfn[own + 1];
                        const float2 nd = sw.nfn[own - SN_N], nu = sw.nfn[own + SN_N];
                        if (!(nl.x * nl.x + nl.y * nl.y < 0.5f || nr.x * nr.x + nr.y * nr.y < 0.5f ||
                              nd.x * nd.x + nd.y * nd.y < 0.5f || nu.x * nu.x + nu.y * nu.y < 0.5f)) {
                            const float curv = (nr.x - nl.x + nu.y - nd.y) * 0.5f * F.P.inv_dx;
                            lp = F.P.phys[1] * curv * F.P.front2[2];
                        }
                    }
                    pl = lp * cz_expf(-(d - 1.0f) * 0.7f);
                }
            }
        }
        fw.pres[q] = make_float2(-F.P.phys[1] * l + disjoining(F, hv), F.P.phys[0] * F.P.g[2] * (s0 - hold) + pl);
        float tv = 0.0f;
        if (hv > 0.0f && plug_on) {
            float sum = 0.0f, mass = 0.0f;
            bool yielding = false;
            const int wq0 = (j - PLUG_RADIUS - by0) * SB_N + (i - PLUG_RADIUS - bx0);
            for (int dy = 0; dy < 2 * PLUG_RADIUS + 1; dy++) {
                for (int dxi = 0; dxi < 2 * PLUG_RADIUS + 1; dxi++) {
                    const float2 v = sw.hq[wq0 + dy * SB_N + dxi];   // dry cells: (+0, +0)
                    yielding = yielding || v.x < 0.0f;
                    sum = sum + v.y;
                    mass = mass + cz_abs(v.x);
                }
            }
            if (yielding && mass > 0.0f) tv = sum / mass;
        }
        fw.tmp[q] = tv;
    }
    __syncthreads();
    // the tile's own outputs of the lap/plug part
    if (in_grid && cell_th) {
        const float hv = sw.h[bq((int)x, (int)y)];
        if (do_clear == 1 || (do_clear == 2 && F.imp_steps == 0u)) {
            F.acc[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            F.h_frame[c] = hv;
        }
        if (do_nose) {
            F.delta_out[c] = sw.nd[nqi((int)x, (int)y)];
            F.fnorm[c] = sw.nfn[nqi((int)x, (int)y)];
        }
        F.tmp[c] = fw.tmp[((int)y - ly0) * LN + ((int)x - lx0)];
    }
    // fluxes of the tile + 2 ring, one face per thread
    const FusedRead g{F, fw, hr, lx0, ly0, ttx, tty, F.k_hold1, F.k_theta, do_nose};
    for (uint32_t it = threadIdx.x; it < (uint32_t)(AP_WIN * AP_WIN * 4); it += (uint32_t)BT) {
        const uint32_t q = it >> 2, k = it & 3u;
        const int gx = tx0 - AP_R + (int)(q % AP_WIN), gy = ty0 - AP_R + (int)(q / AP_WIN);
        if (in_canvas(gx, gy)) {
            if (k == 0u) S.ap.s_hh[q] = fw.h[g.q(gx, gy)];
            float qk = 0.0f;
            if (g.active(gx, gy)) {
                const float2 p00 = pressure_at(F, g, gx, gy);
                FaceGeom fg = face_geom(F, g, gx, gy, (int)k, true, p00);
                if (fg.has) {
                    FaceFlux ff = face(F, g, gx, gy, fg.ib, fg.jb, fg.gcap, fg.ghyd, fg.gtilt, fg.gtn, fg.diag, fg.dir);
                    qk = ff.q * ((k < 2u) ? W_AXIS : W_DIAG_FLUX);
                }
            }
            ((float*)&S.ap.s_f[q])[k] = qk;
        }
    }
    __syncthreads();
    FilmDev Fw = F;
    Fw.h = hw;
    apply_core<BT>(Fw, _t, S.ap);
}

template <class WS>
__global__ __launch_bounds__(1024) void k_substep_fused(WS ws, const float* hr, float* hw, int do_nose, int do_clear) {
    __shared__ SubstepFusedShared S;
    CZ_FOR_TILES(ws) substep_fused_tile(F, hr, hw, _t, S, do_nose, do_clear);
}

// ---------------------------------------------------------------------------
// All substeps of a frame in one launch, for small active sets, where the
// three launches per substep would cost more in dispatch gaps than in work:
// the lap/plug, flux and limiter+apply passes of every substep run over the
// active list in turn with a grid-wide barrier between passes (the three
// global synchronisations of a substep), then the nose of the last substep.
// Each pass calls the tile function of its own kernel, so the results are
// those of the separate launches.  Fa and Fb are the film with the two
// offset buffers in either order: the nose of substeps 1.. writes the other
// buffer, which the rest of the substep reads.
//
// The grid is small (at most a few hundred blocks of 256 threads, resident
// together on the 256 CUs).  A barrier's wait is bounded: past the spin
// limit the block sets the timeout word and stops waiting, every later wait
// returns at once, k_frame_stats reports the word to the host, and the frame
// is reported as failed.
struct SubstepShared {
    union {
        LapShared lap;
        FluxWin flux;
        ApplyShared ap;
    };
};

constexpr uint32_t kBarrierSpins = 1u << 20;

// Block arrival is counted with relaxed atomics after a release fence; the
// last block to arrive resets the count and publishes the next generation;
// the others poll the generation with relaxed loads (an acquire per poll
// would invalidate the L2 on every iteration) and take one acquire fence
// when it changes.
__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t nblocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived + 1u == nblocks) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0u;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
                if (++spins > kBarrierSpins) {
                    __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_substeps(FilmDev Fa, FilmDev Fb, uint32_t substeps, int clear_first) {
    __shared__ SubstepShared S;
    const uint32_t n = *Fa.active_count;
    const uint32_t G = gridDim.x;
    uint32_t* bar = Fa.grid_sync;
    bool flip = false;
    for (uint32_t s = 0; s < substeps; s++) {
        {
            const FilmDev& F = flip ? Fb : Fa;
            for (uint32_t a = blockIdx.x; a < n; a += G)
                lap_plug_tile(F, F.tile_list[a], S.lap, s > 0 ? 1 : 0, 1, (s == 0 && clear_first) ? 1 : 0);
        }
        grid_barrier(bar, G);
        if (s > 0) flip = !flip;
        const FilmDev& F = flip ? Fb : Fa;
        for (uint32_t a = blockIdx.x; a < n; a += G) flux_tile(F, F.tile_list[a], S.flux);
        grid_barrier(bar, G);
        for (uint32_t a = blockIdx.x; a < n; a += G) apply_tile(F, F.tile_list[a], S.ap);
        grid_barrier(bar, G);
    }
    const FilmDev& F = flip ? Fb : Fa;
    for (uint32_t a = blockIdx.x; a < n; a += G) lap_plug_tile(F, F.tile_list[a], S.lap, 1, 0, 0);
}

// ---------------------------------------------------------------------------
// strata_split (film.wgsl:457-514)

__device__ inline float4 band_split(const FilmDev& F, uint32_t cu, float hf, float y, float up, float total, float n) {
    float4 sh = F.strata_h[cu];
    float hu = sh.x + sh.y + sh.z + sh.w;
    if (hu <= 0.0f) return make_float4(total, 0.0f, 0.0f, 0.0f);
    float fr[4] = {sh.x / hu, sh.y / hu, sh.z / hu, sh.w / hu};
    if (y > 0.0f && up > 0.0f) {
        float k = hf / hu;
        float z1 = sh.x * k;
        float z2 = z1 + sh.y * k;
        float z3 = z2 + sh.z * k;
        float z4 = z3 + sh.w * k;
        float b[4] = {band_flux(n, up, y, 0.0f, z1), band_flux(n, up, y, z1, z2), band_flux(n, up, y, z2, z3),
                      band_flux(n, up, y, z3, z4)};
        float bs = b[0] + b[1] + b[2] + b[3];
        if (bs > 0.0f)
            for (int q = 0; q < 4; q++) fr[q] = b[q] / bs;
    }
    return make_float4(fr[0] * total, fr[1] * total, fr[2] * total, fr[3] * total);
}

// strata_split (film.wgsl:457-514), face-parallel: each block takes one
// (tile, face) pair, as k_fluxes_faces does; a cell's four faces are split
// independently.  Against a cell's four faces per thread: 4x the waves and a
// quarter of each thread's chain of face and band evaluations (2048^2 blob
// canvas: 26-28 -> 8 us per frame at 100-250 active tiles, 343 -> 235 us with
// every tile active).  The launch also copies the offsets of active cells
// from the buffer the last nose wrote (F.delta, which split reads) to the
// other one (F.delta_out, which nothing in the launch reads), so that both
// hold the same values everywhere between frames.
template <class WS>
__global__ __launch_bounds__(256) void k_strata_split(WS ws) {
    const uint32_t nact = ws.count();
    const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4;
    const XcdRange rg = xcd_range<true>(4u * nact);
    for (uint32_t wk = rg.first; wk < rg.end; wk += rg.step) {
        const uint32_t it = ws.item(wk >> 2);
        const FilmDev& F = ws.film(it);
        const uint32_t t = ws.tile(F, it);
        const int k = (int)(wk & 3u);
        const uint32_t x = (t % F.tw) * 16u + lx, y = (t / F.tw) * 16u + ly;
        if (x >= F.W || y >= F.H) continue;
        const uint32_t c = idx(F, x, y);
        if (k == 0) F.delta_out[c] = F.delta[c];
        const float4 a4 = F.acc[c];
        const float av = (k == 0) ? a4.x : (k == 1) ? a4.y : (k == 2) ? a4.z : a4.w;
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const bool exists = (k == 0) ? (x + 1u < F.W)
                          : (k == 1) ? (y + 1u < F.H)
                          : (k == 2) ? (x + 1u < F.W && y + 1u < F.H)
                                     : (x + 1u < F.W && y > 0u);
        if (av != 0.0f && exists) {
            const GlobalRead g{F};
            const float2 p00 = pressure_at(F, g, (int)x, (int)y);
            FaceGeom fg = face_geom(F, g, (int)x, (int)y, k, true, p00);
            FaceFlux ff = {0.0f, 0.0f, 0.0f};
            if (fg.has) ff = face(F, g, (int)x, (int)y, fg.ib, fg.jb, fg.gcap, fg.ghyd, fg.gtilt, fg.gtn, fg.diag, fg.dir);
            const uint32_t cb = (k == 0) ? c + 1u : (k == 1) ? c + F.W : (k == 2) ? c + F.W + 1u : c - F.W + 1u;
            const uint32_t cu = (av < 0.0f) ? cb : c;
            o = band_split(F, cu, 0.5f * (F.h[c] + F.h[cb]), ff.y, ff.up, av, ff.n);
        }
        F.acc_s[4u * c + (uint32_t)k] = o;
    }
}

// ---------------------------------------------------------------------------
// reduce_tiles (film.wgsl:1330-1361) fused with pool8 (film.wgsl:1304-1324):
// thread (lx, ly) of the block for 64-tile (tx, ty) owns the same 4x4 cells
// in both, i.e. pool cell (16 tx + lx, 16 ty + ly); each of its four rows is
// one 16-byte load of h and of m when rows are 4-aligned.
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = cz_max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = cz_min(v, __shfl_xor(v, off, 64));
    return v;
}

template <class WS>
__global__ __launch_bounds__(256) void k_reduce_pool(WS ws) {
    __shared__ float s_h[4];
    __shared__ float s_mu[4];
    __shared__ float s_s[4];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    CZ_FOR_TILES(ws) {
        const uint32_t lx = _lx, ly = _ly;
        const uint32_t bx = _t % F.P.tiles[0], by = _t / F.P.tiles[0];
        const uint32_t i0 = bx * TILE64 + lx * 4u;
        const bool vec = (F.W & 3u) == 0u;
        const float hmin4 = 4.0f * F.P.ein[3];
        float hm = 0.0f, mm = 1e9f, mx = 0.0f, ms = 0.0f;
        for (uint32_t dy = 0; dy < 4u; dy++) {
            const uint32_t j = by * TILE64 + ly * 4u + dy;
            if (j >= F.H || i0 >= F.W) continue;
            const uint32_t c0 = idx(F, i0, j);
            float hv[4], mv[4];
            if (vec) {
                const float4 a4 = *(const float4*)(F.h + c0);
                const float4 m4 = *(const float4*)(F.m + c0);
                hv[0] = a4.x; hv[1] = a4.y; hv[2] = a4.z; hv[3] = a4.w;
                mv[0] = m4.x; mv[1] = m4.y; mv[2] = m4.z; mv[3] = m4.w;
            } else {
                for (uint32_t k = 0; k < 4u; k++) {
                    hv[k] = (i0 + k < F.W) ? F.h[c0 + k] : 0.0f;
                    mv[k] = (i0 + k < F.W) ? F.m[c0 + k] : 0.0f;
                }
            }
            for (uint32_t k = 0; k < 4u; k++) {
                if (i0 + k >= F.W) continue;
                if (hv[k] > hmin4) {
                    hm = cz_max(hm, hv[k]);
                    mm = cz_min(mm, F.rheo[c0 + k].x);
                }
                mx = cz_max(mx, hv[k] + mv[k]);
                ms = cz_max(ms, mv[k]);
            }
        }
        const uint32_t pi = bx * 16u + lx, pj = by * 16u + ly;
        // s_max of poll_stats (src/film/mod.rs:961): the largest finite pool surface
        float sm = 0.0f;
        if (pi < F.pw && pj < F.ph) {
            F.pool_out[pj * F.pw + pi] = mx;
            F.pool_out[F.pw * F.ph + pj * F.pw + pi] = ms;
            if (isfinite(mx)) sm = mx;
        }
        hm = wave_max(hm);
        mm = wave_min(mm);
        sm = wave_max(sm);
        __syncthreads();
        if (lane == 0u) {
            s_h[wave] = hm;
            s_mu[wave] = mm;
            s_s[wave] = sm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            F.stats_out[_t] = make_float2(cz_max(cz_max(s_h[0], s_h[1]), cz_max(s_h[2], s_h[3])),
                                          cz_min(cz_min(s_mu[0], s_mu[1]), cz_min(s_mu[2], s_mu[3])));
            F.smax_out[_t] = cz_max(cz_max(s_s[0], s_s[1]), cz_max(s_s[2], s_s[3]));
        }
    }
}

// Fold of the per-64-tile stats (poll_stats, src/film/mod.rs:944-951) and the
// frame scalars the host reads back: [0] h_max, [1] mu_min, [2] active count
// (bits), [3] cells of the active tiles (bits), [4..7] cg_s[1], [8] s_max.
template <class CS>
__global__ __launch_bounds__(1024) void k_frame_stats(CS cs) {
    // one 1024-thread block per canvas: each thread takes a strided share of
    // the 64-tile stats and of the active list, then wave and block
    // reductions (max and min are exact in any order, the cell count is an
    // integer sum)
    const FilmDev& F = cs.film(blockIdx.x);
    __shared__ float s_h[16], s_mu[16], s_s[16];
    __shared__ uint32_t s_cells[16];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t nt = F.t64x * F.t64y;
    float hm = 0.0f, mm = 1e9f, sm = 0.0f;
    for (uint32_t t = threadIdx.x; t < nt; t += 1024u) {
        const float2 v = F.stats_out[t];
        hm = cz_max(hm, v.x);
        mm = cz_min(mm, v.y);
        sm = cz_max(sm, F.smax_out[t]);
    }
    uint32_t cells = 0u;
    const uint32_t nact = *F.active_count;
    for (uint32_t a = threadIdx.x; a < nact; a += 1024u) {
        const uint32_t t = F.tile_list[a];
        const uint32_t tx = t % F.tw, ty = t / F.tw;
        const uint32_t cw = F.W - tx * 16u < 16u ? F.W - tx * 16u : 16u;
        const uint32_t ch = F.H - ty * 16u < 16u ? F.H - ty * 16u : 16u;
        cells += cw * ch;
    }
    hm = wave_max(hm);
    mm = wave_min(mm);
    sm = wave_max(sm);
    for (int off = 32; off > 0; off >>= 1) cells += __shfl_xor(cells, off, 64);
    if (lane == 0u) {
        s_h[wave] = hm;
        s_mu[wave] = mm;
        s_s[wave] = sm;
        s_cells[wave] = cells;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float h = s_h[0], mu = s_mu[0], smax = s_s[0];
        uint32_t cs_ = s_cells[0];
        for (int w = 1; w < 16; w++) {
            h = cz_max(h, s_h[w]);
            mu = cz_min(mu, s_mu[w]);
            smax = cz_max(smax, s_s[w]);
            cs_ += s_cells[w];
        }
        F.frame_out[8] = smax;
        F.frame_out[0] = h;
        F.frame_out[1] = mu;
        F.frame_out[2] = cz_bits2f(nact);
        F.frame_out[3] = cz_bits2f(cs_);
        for (int k = 0; k < 4; k++) F.frame_out[4 + k] = F.cg_s[4 + k];
        F.frame_out[15] = cz_bits2f(F.grid_sync ? F.grid_sync[2] : 0u);
    }
}

// ---------------------------------------------------------------------------
// strata advection (strata.wgsl:157-323)

template <class WS>
__global__ __launch_bounds__(256) void k_strata_limit(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        float4 s0 = F.acc_s[4u * c], s1 = F.acc_s[4u * c + 1u], s2 = F.acc_s[4u * c + 2u], s3 = F.acc_s[4u * c + 3u];
        float out[4] = {cz_max(s0.x, 0.0f) + cz_max(s1.x, 0.0f) + cz_max(s2.x, 0.0f) + cz_max(s3.x, 0.0f),
                        cz_max(s0.y, 0.0f) + cz_max(s1.y, 0.0f) + cz_max(s2.y, 0.0f) + cz_max(s3.y, 0.0f),
                        cz_max(s0.z, 0.0f) + cz_max(s1.z, 0.0f) + cz_max(s2.z, 0.0f) + cz_max(s3.z, 0.0f),
                        cz_max(s0.w, 0.0f) + cz_max(s1.w, 0.0f) + cz_max(s2.w, 0.0f) + cz_max(s3.w, 0.0f)};
        float4 q[4];
        bool has[4] = {x > 0u && cell_active(F, x - 1u, y), y > 0u && cell_active(F, x, y - 1u),
                       x > 0u && y > 0u && cell_active(F, x - 1u, y - 1u),
                       x > 0u && y + 1u < F.H && cell_active(F, x - 1u, y + 1u)};
        if (has[0]) q[0] = F.acc_s[4u * (c - 1u)];
        if (has[1]) q[1] = F.acc_s[4u * (c - F.W) + 1u];
        if (has[2]) q[2] = F.acc_s[4u * (c - F.W - 1u) + 2u];
        if (has[3]) q[3] = F.acc_s[4u * (c + F.W - 1u) + 3u];
        for (int k = 0; k < 4; k++) {
            if (!has[k]) continue;
            out[0] = out[0] + cz_max(-q[k].x, 0.0f);
            out[1] = out[1] + cz_max(-q[k].y, 0.0f);
            out[2] = out[2] + cz_max(-q[k].z, 0.0f);
            out[3] = out[3] + cz_max(-q[k].w, 0.0f);
        }
        float4 th4 = F.strata_h[c];
        float th[4] = {th4.x, th4.y, th4.z, th4.w};
        float a[4];
        for (int k = 0; k < 4; k++) {
            a[k] = 1.0f;
            if (out[k] > 0.0f) a[k] = cz_min(1.0f, th[k] * F.P.dx / out[k]);
        }
        F.alpha_s[c] = make_float4(a[0], a[1], a[2], a[3]);
    }
}

__device__ __forceinline__ float sel4(float4 v, uint32_t k) {
    return k == 0u ? v.x : (k == 1u ? v.y : (k == 2u ? v.z : v.w));
}

// The tile's cells and a 1-cell ring staged in LDS for strata_update: the
// limiter factors, strata heights, cure and pigment words of every cell a
// tile cell takes an inflow from.  The eight inflows of a cell then read
// their donor's strata from LDS; read from HBM, each inflow's and each
// stratum's loads waited behind the previous merge.
constexpr int SU_N = 18, SU_CELLS = SU_N * SU_N;
struct StrataUpdShared {
    float4 an[SU_CELLS];
    float4 th[SU_CELLS];
    float4 cu[SU_CELLS];
    uint4 pg[SU_CELLS * 8];
};

// inflow_from (strata.wgsl:167-188): the strata of the neighbour in slot qn
// of the staged ring moving into the cell with face fluxes f (one per
// stratum).  The stratum loop stays a loop (one copy of stack_add_band) and
// reads its per-stratum values by select, so nothing is indexed at run time.
__device__ inline void inflow_from_lds(const FilmDev& F, Stack& s, const StrataUpdShared& S, int qn, float4 f) {
    const float4 an4 = S.an[qn];
    const float4 thn4 = S.th[qn];
    const float4 cu4 = S.cu[qn];
    float z0 = 0.0f;
#pragma unroll 1
    for (uint32_t k = 0; k < 4u; k++) {
        const float thk = sel4(thn4, k);
        float z1 = z0 + thk;
        float v = sel4(f, k) * sel4(an4, k) * F.P.inv_dx;
        if (v > 0.0f && thk > 0.0f) {
            float pig[16];
            unpack8(S.pg[qn * 8 + (int)k * 2], S.pg[qn * 8 + (int)k * 2 + 1], pig);
            stack_add_band(F, s, z0, z1, v, pig, sel4(cu4, k));
        }
        z0 = z1;
    }
}

__device__ __forceinline__ bool zero4(const float4 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f && v.w == 0.0f; }
__device__ __forceinline__ void to4(const float4 v, float* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }

// strata_update (strata.wgsl:190-274); also records the new stack sum of each
// active cell for the fill passes.
template <class WS>
__global__ __launch_bounds__(256) void k_strata_update(WS ws) {
    __shared__ StrataUpdShared S;
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        const int gx0 = (int)(_t % F.tw) * 16 - 1, gy0 = (int)(_t / F.tw) * 16 - 1;
        const int W_ = (int)F.W, H_ = (int)F.H;
        __syncthreads();  // the previous tile's cells are done with S
        {
            // every load first, then every store: a ring cell beyond the
            // canvas (or a slot past the ring) reads a clamped cell, which
            // nothing uses, so the loads carry no branch between them
            constexpr int R1 = (SU_CELLS + 255) / 256, R2 = (SU_CELLS * 8 + 255) / 256;
            float4 va[R1], vt[R1], vc[R1];
            uint4 vp[R2];
#pragma unroll
            for (int r = 0; r < R1; r++) {
                const int q = min((int)threadIdx.x + r * 256, SU_CELLS - 1);
                const int gx = min(max(gx0 + q % SU_N, 0), W_ - 1), gy = min(max(gy0 + q / SU_N, 0), H_ - 1);
                const uint32_t cn = idx(F, (uint32_t)gx, (uint32_t)gy);
                va[r] = F.alpha_s[cn];
                vt[r] = F.strata_h[cn];
                vc[r] = F.strata_cure[cn];
            }
#pragma unroll
            for (int r = 0; r < R2; r++) {
                const int q = min((int)threadIdx.x + r * 256, SU_CELLS * 8 - 1);
                const int cell = q >> 3;
                const int gx = min(max(gx0 + cell % SU_N, 0), W_ - 1), gy = min(max(gy0 + cell / SU_N, 0), H_ - 1);
                vp[r] = F.strata_pig[idx(F, (uint32_t)gx, (uint32_t)gy) * 8u + (uint32_t)(q & 7)];
            }
#pragma unroll
            for (int r = 0; r < R1; r++) {
                const int q = (int)threadIdx.x + r * 256;
                if (q < SU_CELLS) {
                    S.an[q] = va[r];
                    S.th[q] = vt[r];
                    S.cu[q] = vc[r];
                }
            }
#pragma unroll
            for (int r = 0; r < R2; r++) {
                const int q = (int)threadIdx.x + r * 256;
                if (q < SU_CELLS * 8) S.pg[q] = vp[r];
            }
        }
        __syncthreads();
        if (!in_grid) continue;
        // slots of the cell and of its neighbours in S
        const int own = ((int)_ly + 1) * SU_N + (int)_lx + 1;
        float4 fr = F.acc_s[4u * c], fu = F.acc_s[4u * c + 1u], fdg = F.acc_s[4u * c + 2u], feg = F.acc_s[4u * c + 3u];
        const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 fl = z4, fd = z4, fll = z4, flu = z4;
        uint32_t cl = c, cd = c, cll = c, clu = c;
        if (x > 0u && cell_active(F, x - 1u, y)) { cl = c - 1u; fl = F.acc_s[4u * cl]; }
        if (y > 0u && cell_active(F, x, y - 1u)) { cd = c - F.W; fd = F.acc_s[4u * cd + 1u]; }
        if (x > 0u && y > 0u && cell_active(F, x - 1u, y - 1u)) { cll = c - F.W - 1u; fll = F.acc_s[4u * cll + 2u]; }
        if (x > 0u && y + 1u < F.H && cell_active(F, x - 1u, y + 1u)) { clu = c + F.W - 1u; flu = F.acc_s[4u * clu + 3u]; }
        if (zero4(fr) && zero4(fu) && zero4(fdg) && zero4(feg) && zero4(fl) && zero4(fd) && zero4(fll) && zero4(flu)) {
            float hc = F.h[c];
            float4 t4 = S.th[own];
            float th[4] = {cz_max(t4.x, 0.0f), cz_max(t4.y, 0.0f), cz_max(t4.z, 0.0f), cz_max(t4.w, 0.0f)};
            float nt[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            float4 cu = S.cu[own];
            if (hc <= 0.0f) {
                cu = z4;
            } else {
                float sum = th[0] + th[1] + th[2] + th[3];
                #pragma unroll
                for (int k = 0; k < 4; k++) nt[k] = th[k];
                if (sum > 0.0f)
                    #pragma unroll
                    for (int k = 0; k < 4; k++) nt[k] = th[k] * (hc / sum);
            }
            bool plain = true;
            for (uint32_t k = 0; k < 3u; k++) {
                if (nt[k] <= 0.0f && nt[k + 1u] > 0.0f) plain = false;
                if (nt[k] > 0.0f && nt[k] < F.P.misc[0] && nt[k + 1u] > 0.0f) plain = false;
            }
            if (plain) {
                F.strata_h2[c] = make_float4(nt[0], nt[1], nt[2], nt[3]);
                F.strata_cure2[c] = cu;
                for (uint32_t k = 0; k < 8u; k++) F.strata_pig2[c * 8u + k] = S.pg[own * 8 + (int)k];
                F.tsum_a[c] = nt[0] + nt[1] + nt[2] + nt[3];
                continue;
            }
        }
        Stack s;
        {
            const float4 t = S.th[own], u = S.cu[own];
            s.th[0] = t.x; s.th[1] = t.y; s.th[2] = t.z; s.th[3] = t.w;
            s.cu[0] = u.x; s.cu[1] = u.y; s.cu[2] = u.z; s.cu[3] = u.w;
#pragma unroll
            for (int k = 0; k < 4; k++) unpack8(S.pg[own * 8 + k * 2], S.pg[own * 8 + k * 2 + 1], s.pg[k]);
        }
        float4 a4 = S.an[own];
        float a[4] = {a4.x, a4.y, a4.z, a4.w};
        float vr[4], vu[4], vdg[4], veg[4], vl[4], vd[4], vll[4], vlu[4];
        to4(fr, vr); to4(fu, vu); to4(fdg, vdg); to4(feg, veg); to4(fl, vl); to4(fd, vd); to4(fll, vll); to4(flu, vlu);
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            float out = cz_max(vr[k], 0.0f) + cz_max(vu[k], 0.0f) + cz_max(vdg[k], 0.0f) + cz_max(veg[k], 0.0f) +
                        cz_max(-vl[k], 0.0f) + cz_max(-vd[k], 0.0f) + cz_max(-vll[k], 0.0f) + cz_max(-vlu[k], 0.0f);
            s.th[k] = cz_max(s.th[k] - out * a[k] * F.P.inv_dx, 0.0f);
        }
        // the eight inflows in the reference's order, one call site
#pragma unroll 1
        for (uint32_t q = 0; q < 8u; q++) {
            bool ok = false;
            int qn = own;
            float4 nf = z4;
            switch (q) {
            case 0: ok = x + 1u < F.W; qn = own + 1; nf = make_float4(-fr.x, -fr.y, -fr.z, -fr.w); break;
            case 1: ok = y + 1u < F.H; qn = own + SU_N; nf = make_float4(-fu.x, -fu.y, -fu.z, -fu.w); break;
            case 2: ok = x + 1u < F.W && y + 1u < F.H; qn = own + SU_N + 1;
                    nf = make_float4(-fdg.x, -fdg.y, -fdg.z, -fdg.w); break;
            case 3: ok = x + 1u < F.W && y > 0u; qn = own - SU_N + 1;
                    nf = make_float4(-feg.x, -feg.y, -feg.z, -feg.w); break;
            case 4: ok = cl != c; qn = own - 1; nf = fl; break;
            case 5: ok = cd != c; qn = own - SU_N; nf = fd; break;
            case 6: ok = cll != c; qn = own - SU_N - 1; nf = fll; break;
            default: ok = clu != c; qn = own + SU_N - 1; nf = flu; break;
            }
            if (ok) inflow_from_lds(F, s, S, qn, nf);
        }
        float hc = F.h[c];
        float sum = s.th[0] + s.th[1] + s.th[2] + s.th[3];
        if (hc <= 0.0f) {
            #pragma unroll
            for (int k = 0; k < 4; k++) { s.th[k] = 0.0f; s.cu[k] = 0.0f; }
        } else if (sum > 0.0f) {
            float k = hc / sum;
            #pragma unroll
            for (int q = 0; q < 4; q++) s.th[q] = s.th[q] * k;
        }
        stack_compact(F, s);
        stack_store_to(F.strata_h2, F.strata_cure2, F.strata_pig2, c, s);
        F.tsum_a[c] = s.th[0] + s.th[1] + s.th[2] + s.th[3];
    }
}

// strata_copy_back (strata.wgsl:276-286): the pigment words of a tile row are
// one 2 KiB run, copied by consecutive threads.
// strata_copy_back + the six strata_fill passes (strata.wgsl:276-323) in one
// launch per active tile.
//
// A fill pass gives a wet cell of an active tile whose stack is empty the top
// stratum of the neighbour (x-1, y-1, x+1, y+1 order, strictly largest stack
// sum) with a non-empty stack, as one stratum of the cell's height; the
// passes are Jacobi steps (each reads the sums before it).  A cell filled in
// pass p holds exactly what its source held in its top stratum, so a cell
// filled from a filled cell ends with the stratum of the first cell of the
// chain (its root), which no pass writes.  The block runs the six passes on
// the sums of the tile and a 6-cell ring in LDS (a pass moves information one
// cell, so the tile's results are exact), tracking each cell's root, and then
// writes the tile: filled cells from their root (the post-update buffers if
// the root's tile is active: what copy_back would have put there), every
// other cell by the copy_back of the post-update stack.
constexpr int FL_R = 6;
constexpr int FL_WIN = 16 + 2 * FL_R;
constexpr int FL_CELLS = FL_WIN * FL_WIN;

template <class WS>
__global__ __launch_bounds__(256) void k_strata_fill(WS ws) {
    __shared__ float s_t[2][FL_CELLS];       // stack sums, ping-pong over the passes
    __shared__ uint16_t s_root[2][FL_CELLS]; // window index of the cell's root
    __shared__ float s_h[FL_CELLS];
    __shared__ uint8_t s_kind[FL_CELLS];     // 1 in canvas, 2 active tile
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        const int wx0 = (int)((_t % F.tw) * 16u) - FL_R, wy0 = (int)((_t / F.tw) * 16u) - FL_R;
        __syncthreads();  // the previous tile's window is no longer read
        for (int q = (int)threadIdx.x; q < FL_CELLS; q += 256) {
            const int gx = wx0 + q % FL_WIN, gy = wy0 + q / FL_WIN;
            float t = 0.0f, hv = 0.0f;
            uint8_t kind = 0u;
            if (gx >= 0 && gy >= 0 && gx < (int)F.W && gy < (int)F.H) {
                const uint32_t cg = idx(F, (uint32_t)gx, (uint32_t)gy);
                kind = 1u;
                if (cell_active(F, (uint32_t)gx, (uint32_t)gy)) {
                    kind |= 2u;
                    t = F.tsum_a[cg];
                } else {
                    const float4 v = F.strata_h[cg];
                    t = v.x + v.y + v.z + v.w;
                }
                hv = F.h[cg];
            }
            s_t[0][q] = t;
            s_root[0][q] = (uint16_t)q;
            s_h[q] = hv;
            s_kind[q] = kind;
        }
        __syncthreads();
#pragma unroll 1
        for (int pass = 0; pass < 6; pass++) {
            const int in = pass & 1, out = in ^ 1;
            for (int q = (int)threadIdx.x; q < FL_CELLS; q += 256) {
                float own = s_t[in][q];
                uint16_t root = s_root[in][q];
                const float hv = s_h[q];
                if (s_kind[q] == 3u && !(hv <= 0.0f) && !(own > 0.0f)) {
                    const int qx = q % FL_WIN, gx = wx0 + qx, gy = wy0 + q / FL_WIN;
                    int src = -1;
                    float best = 0.0f;
                    // neighbours outside the window only reach cells within
                    // 6 of its edge, whose results the tile never uses
                    if (gx > 0 && qx > 0) { const float v = s_t[in][q - 1]; if (v > best) { best = v; src = q - 1; } }
                    if (gy > 0 && q >= FL_WIN) { const float v = s_t[in][q - FL_WIN]; if (v > best) { best = v; src = q - FL_WIN; } }
                    if (gx + 1 < (int)F.W && qx + 1 < FL_WIN) { const float v = s_t[in][q + 1]; if (v > best) { best = v; src = q + 1; } }
                    if (gy + 1 < (int)F.H && q + FL_WIN < FL_CELLS) { const float v = s_t[in][q + FL_WIN]; if (v > best) { best = v; src = q + FL_WIN; } }
                    if (src >= 0) {
                        own = hv;
                        root = s_root[in][src];
                    }
                }
                s_t[out][q] = own;
                s_root[out][q] = root;
            }
            __syncthreads();
        }
        // six passes end in buffer 0
        const int qc = (_ly + FL_R) * FL_WIN + (_lx + FL_R);
        const int rq = s_root[0][qc];
        if (in_grid) {
            if (rq != qc) {
                const uint32_t r = idx(F, (uint32_t)(wx0 + rq % FL_WIN), (uint32_t)(wy0 + rq / FL_WIN));
                const bool ra = (s_kind[rq] & 2u) != 0u;
                const float4 ts = ra ? F.strata_h2[r] : F.strata_h[r];
                uint32_t ks = 0;
                if (ts.x > 0.0f) ks = 0;
                if (ts.y > 0.0f) ks = 1;
                if (ts.z > 0.0f) ks = 2;
                if (ts.w > 0.0f) ks = 3;
                const float4 cs4 = ra ? F.strata_cure2[r] : F.strata_cure[r];
                const float cs = ks == 0 ? cs4.x : (ks == 1 ? cs4.y : (ks == 2 ? cs4.z : cs4.w));
                const uint4* pg = ra ? F.strata_pig2 : F.strata_pig;
                const uint4 a = pg[(r * 4u + ks) * 2u];
                const uint4 b = pg[(r * 4u + ks) * 2u + 1u];
                F.strata_h[c] = make_float4(s_h[qc], 0.0f, 0.0f, 0.0f);
                F.strata_cure[c] = make_float4(cs, 0.0f, 0.0f, 0.0f);
                F.strata_pig[c * 8u] = a;
                F.strata_pig[c * 8u + 1u] = b;
                const uint4 z = make_uint4(0u, 0u, 0u, 0u);
                for (uint32_t k = 2u; k < 8u; k++) F.strata_pig[c * 8u + k] = z;
            } else {
                F.strata_h[c] = F.strata_h2[c];
                F.strata_cure[c] = F.strata_cure2[c];
            }
        }
        // copy_back of the pigment words of the cells not filled: a tile row
        // is one 2 KiB run, copied by consecutive threads
        const uint32_t x0 = (_t % F.tw) * 16u, y0 = (_t / F.tw) * 16u;
        for (uint32_t k = 0; k < 8u; k++) {
            const uint32_t q = k * 256u + threadIdx.x;
            const uint32_t row = q >> 7, off = q & 127u;
            const uint32_t xx = x0 + (off >> 3), yy = y0 + row;
            const int wq = ((int)row + FL_R) * FL_WIN + (int)(off >> 3) + FL_R;
            if (xx < F.W && yy < F.H && s_root[0][wq] == (uint16_t)wq) {
                const uint32_t base = idx(F, x0, yy) * 8u + off;
                F.strata_pig[base] = F.strata_pig2[base];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// implicit hydrostatic CG (film.wgsl:895-1302)

// film.wgsl:895-920
__device__ inline float face_mob(const FilmDev& F, uint32_t ca, uint32_t cb, float2 gcap, float2 ghyd, float2 gtilt,
                                 bool diag) {
    float ha = F.h[ca], hb = F.h[cb];
    if (cz_min(ha, hb) <= 0.0f) return 0.0f;
    float hf = 0.5f * (ha + hb);
    float2 ra = F.rheo[ca], rb = F.rheo[cb];
    float wa = ha / (ha + hb);
    float wb = 1.0f - wa;
    const Visc mu{0.0f, wa * F.lmu_cell[ca] + wb * F.lmu_cell[cb], false};
    float ty = wa * ra.y + wb * rb.y;
    const float nf = cell_n_on(F) ? wa * F.n_cell[ca] + wb * F.n_cell[cb] : F.P.phys[2];
    float thin = cz_min(ha, hb) / hf;
    float fcap = thin * thin * thin;
    float2 yc = yield_cap(F, make_float2(gcap.x * fcap, gcap.y * fcap), ty);
    float2 g = make_float2(yc.x + ghyd.x + gtilt.x, yc.y + ghyd.y + gtilt.y);
    float gmag = cz_len2(g.x, g.y);
    float ydep = (ty <= 0.0f) ? hf : ((g.x >= 0.0f) ? ha : hb);
    float qy, qu;
    float q = hb_flux(nf, gmag, false, 0.0f, ydep, mu, ty, qy, qu);
    if (q <= 0.0f) return 0.0f;
    float mob = cz_min(cz_max(q / gmag - F.P.misc[3], 0.0f), F.P.front[2]);
    if (cz_min(ha, hb) < 0.5f * cz_max(ha, hb)) {
        if (diag) return 0.0f;
        uint32_t tgt = cb;
        if (hb > ha) tgt = ca;
        mob = mob / cz_feed_pow(feeders(F, GlobalRead{F}, (int)(tgt % F.W), (int)(tgt / F.W)));
    }
    return mob;
}

__device__ __forceinline__ float kappa(const FilmDev& F) {
    return F.P.front[3] * F.P.phys[0] * F.P.g[2] * F.P.inv_dx * F.P.inv_dx;
}
__device__ __forceinline__ float hold_u(const FilmDev& F, uint32_t c) {
    float hold = 0.0f;
    if (F.h[c] > 0.0f) hold = hold_level(F, nose_dist(F, c));
    return F.m[c] - hold;
}

template <class WS>
__global__ __launch_bounds__(256) void k_cg_setup(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        const GlobalRead g{F};
        float2 p00 = pressure_at(F, g, (int)x, (int)y);
        float mob[4];
        for (int k = 0; k < 4; k++) {
            mob[k] = 0.0f;
            FaceGeom fg = face_geom(F, g, (int)x, (int)y, k, false, p00);
            if (fg.has) mob[k] = face_mob(F, c, idx(F, (uint32_t)fg.ib, (uint32_t)fg.jb), fg.gcap, fg.ghyd, fg.gtilt, fg.diag);
        }
        F.cg_m[c] = make_float4(mob[0], mob[1], mob[2], mob[3]);
    }
}

// film.wgsl:987-1003
__device__ __forceinline__ void cg_nb(const FilmDev& F, uint32_t k, uint32_t i, uint32_t j, uint32_t c, uint32_t& oc,
                                      float& om) {
    oc = 0u;
    om = 0.0f;
    switch (k) {
    case 0: if (i + 1u < F.W) { oc = c + 1u; om = W_AXIS * F.cg_m[c].x; } break;
    case 1: if (j + 1u < F.H) { oc = c + F.W; om = W_AXIS * F.cg_m[c].y; } break;
    case 2: if (i + 1u < F.W && j + 1u < F.H) { oc = c + F.W + 1u; om = W_DIAG_LAP * F.cg_m[c].z; } break;
    case 3: if (i + 1u < F.W && j > 0u) { oc = c - F.W + 1u; om = W_DIAG_LAP * F.cg_m[c].w; } break;
    case 4: if (i > 0u) { oc = c - 1u; om = W_AXIS * F.cg_m[oc].x; } break;
    case 5: if (j > 0u) { oc = c - F.W; om = W_AXIS * F.cg_m[oc].y; } break;
    case 6: if (i > 0u && j > 0u) { oc = c - F.W - 1u; om = W_DIAG_LAP * F.cg_m[oc].z; } break;
    default: if (i > 0u && j + 1u < F.H) { oc = c + F.W - 1u; om = W_DIAG_LAP * F.cg_m[oc].w; } break;
    }
}

__device__ __forceinline__ float lap_src(const FilmDev& F, uint32_t c, uint32_t which) {
    switch (which) {
    case 0: return F.cg_p[c];
    case 1: return F.h[c];
    case 2: return F.cg_t2[c];
    case 3: return F.cg_t[c];
    case 4: return F.cg_x[c] + hold_u(F, c);
    default: return hold_u(F, c);
    }
}

// film.wgsl:1016-1036
__device__ inline float lap_w(const FilmDev& F, uint32_t i, uint32_t j, uint32_t c, uint32_t which) {
    int ii = (int)i, jj = (int)j;
    float vc = lap_src(F, c, which);
    float acc_v = 0.0f;
    for (int dj = -1; dj <= 1; dj++) {
        for (int di = -1; di <= 1; di++) {
            if (di == 0 && dj == 0) continue;
            int xx = ii + di, yy = jj + dj;
            if (xx < 0 || yy < 0 || xx >= (int)F.W || yy >= (int)F.H) continue;
            if (!cell_active(F, (uint32_t)xx, (uint32_t)yy)) continue;
            uint32_t n = idx(F, (uint32_t)xx, (uint32_t)yy);
            if (F.h[n] <= 0.0f) continue;
            float w = W_AXIS;
            if (di != 0 && dj != 0) w = W_DIAG_LAP;
            acc_v = acc_v + w * (lap_src(F, n, which) - vc);
        }
    }
    return acc_v;
}

// film.wgsl:1038-1058
__device__ inline float lap_diag(const FilmDev& F, uint32_t i, uint32_t j) {
    int ii = (int)i, jj = (int)j;
    float s1 = 0.0f, s2 = 0.0f;
    for (int dj = -1; dj <= 1; dj++) {
        for (int di = -1; di <= 1; di++) {
            if (di == 0 && dj == 0) continue;
            int xx = ii + di, yy = jj + dj;
            if (xx < 0 || yy < 0 || xx >= (int)F.W || yy >= (int)F.H) continue;
            if (!cell_active(F, (uint32_t)xx, (uint32_t)yy)) continue;
            if (F.h[idx(F, (uint32_t)xx, (uint32_t)yy)] <= 0.0f) continue;
            float w = W_AXIS;
            if (di != 0 && dj != 0) w = W_DIAG_LAP;
            s1 = s1 + w;
            s2 = s2 + w * w;
        }
    }
    return s1 * s1 + s2;
}

// Block sum of one double per thread (wave64 shuffles, then LDS).
__device__ inline double block_sum_d(double v) {
    __shared__ double part[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0u) part[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (part[0] + part[1]) + (part[2] + part[3]);
    return r;
}

// which: 1 = cg_lap_h (t <- L h, t2 <- L hold_u), 0 = cg_lap_p (t <- L p),
// 4 = cg_lap_u (t <- L (x + hold_u))   (film.wgsl:1060-1083)
template <class WS>
__global__ __launch_bounds__(256) void k_cg_lap(WS ws, uint32_t which) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        if (which == 1u) {
            float a = lap_w(F, x, y, c, 1u);
            float b = lap_w(F, x, y, c, 5u);
            F.cg_t[c] = a;
            F.cg_t2[c] = b;
        } else {
            F.cg_t[c] = lap_w(F, x, y, c, which);
        }
    }
}

// cg_init (film.wgsl:1110-1144)
template <class WS>
__global__ __launch_bounds__(256) void k_cg_init(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        double v = 0.0;
        if (in_grid) {
            float k = kappa(F);
            float hc = F.h[c];
            float uc = hold_u(F, c);
            float diag = 1.0f, b = hc, ax = hc;
            for (uint32_t f = 0; f < 8u; f++) {
                uint32_t nc;
                float nm;
                cg_nb(F, f, x, y, c, nc, nm);
                if (nm > 0.0f) {
                    diag = diag + k * nm;
                    b = b + k * nm * (hold_u(F, nc) - uc);
                    ax = ax + k * nm * (hc - F.h[nc]);
                }
            }
            if (hc > 0.0f) {
                float cc = F.P.front3[0];
                diag = diag + cc * lap_diag(F, x, y);
                b = b - cc * lap_w(F, x, y, c, 2u);
                ax = ax + cc * lap_w(F, x, y, c, 3u);
            }
            float r = b - ax;
            F.cg_d[c] = diag;
            F.cg_x[c] = hc;
            F.cg_r[c] = r;
            F.cg_p[c] = r / diag;
            v = (double)(r * r / diag);
        }
        double s = block_sum_d(v);
        if (threadIdx.x == 0) F.cg_part[_a] = s;
    }
}

// cg_iter_a (film.wgsl:1146-1157): q = A p, partial p.q
template <class WS>
__global__ __launch_bounds__(256) void k_cg_iter_a(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        double v = 0.0;
        if (in_grid) {
            float k = kappa(F);
            float pc = F.cg_p[c];
            float accv = pc;
            for (uint32_t f = 0; f < 8u; f++) {
                uint32_t nc;
                float nm;
                cg_nb(F, f, x, y, c, nc, nm);
                if (nm > 0.0f) accv = accv + k * nm * (pc - F.cg_p[nc]);
            }
            if (F.h[c] > 0.0f) accv = accv + F.P.front3[0] * lap_w(F, x, y, c, 3u);
            F.cg_q[c] = accv;
            v = (double)(pc * accv);
        }
        double s = block_sum_d(v);
        if (threadIdx.x == 0) F.cg_part[_a] = s;
    }
}

// cg_iter_b (film.wgsl:1159-1172)
template <class WS>
__global__ __launch_bounds__(256) void k_cg_iter_b(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        double v = 0.0;
        if (in_grid) {
            float alpha = F.cg_s[2];
            float r = F.cg_r[c] - alpha * F.cg_q[c];
            F.cg_x[c] = F.cg_x[c] + alpha * F.cg_p[c];
            F.cg_r[c] = r;
            v = (double)(r * r / F.cg_d[c]);
        }
        double s = block_sum_d(v);
        if (threadIdx.x == 0) F.cg_part[_a] = s;
    }
}

// cg_iter_c (film.wgsl:1174-1181)
template <class WS>
__global__ __launch_bounds__(256) void k_cg_iter_c(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        float beta = F.cg_s[3];
        F.cg_p[c] = F.cg_r[c] / F.cg_d[c] + beta * F.cg_p[c];
    }
}

// Single-block FP64 fold of the per-tile partials (film.wgsl:1183-1255):
// mode 0 = cg_reduce_init, 1 = cg_reduce_a, 2 = cg_reduce_b, 3 = cg_reduce_mass.
template <class CS>
__global__ __launch_bounds__(256) void k_cg_reduce(CS cs, int mode) {
    const FilmDev& F = cs.film(blockIdx.x);
    const double* part = F.cg_part + cs.base(blockIdx.x);
    const double* part2 = F.cg_part2 + cs.base(blockIdx.x);
    uint32_t n = *F.active_count;
    double a = 0.0, b = 0.0;
    for (uint32_t t = threadIdx.x; t < n; t += 256u) {
        a += part[t];
        if (mode == 3) b += part2[t];
    }
    double sa = block_sum_d(a);
    double sb = (mode == 3) ? block_sum_d(b) : 0.0;
    if (threadIdx.x != 0) return;
    float* s = F.cg_s;
    if (mode == 0) {
        float rz = (float)sa;
        s[0] = rz; s[1] = 0.0f; s[2] = 0.0f; s[3] = 0.0f;
        s[4] = rz; s[5] = 0.0f; s[6] = 0.0f; s[7] = 0.0f;
    } else if (mode == 1) {
        float pq = (float)sa;
        float alpha = 0.0f;
        if (pq > 0.0f) alpha = s[0] / pq;
        s[1] = pq; s[2] = alpha; s[3] = 0.0f;
    } else if (mode == 2) {
        float rz = (float)sa;
        float beta = 0.0f;
        if (s[0] > 0.0f) beta = rz / s[0];
        s[0] = rz; s[3] = beta;
    } else {
        float d = (float)sa, tot = (float)sb;
        s[5] = d; s[6] = tot; s[7] = s[0];
    }
}

// cg_finish (film.wgsl:1261-1302); the new height goes to cg_q so that
// neighbours read the height from before the pass; k_cg_rescale commits it.
template <class WS>
__global__ __launch_bounds__(256) void k_cg_finish(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        bool moved = false;
        double dv = 0.0, tv = 0.0;
        if (in_grid) {
            float h0 = F.h[c];
            float hn = cz_max(F.cg_x[c], 0.0f);
            float rg = F.P.phys[0] * F.P.g[2] * F.P.inv_dx;
            float pc = F.cg_x[c] + hold_u(F, c);
            float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            float4 mob = F.cg_m[c];
            if (mob.x > 0.0f) { uint32_t n = c + 1u; f[0] = W_AXIS * mob.x * rg * (pc - (F.cg_x[n] + hold_u(F, n))); }
            if (mob.y > 0.0f) { uint32_t n = c + F.W; f[1] = W_AXIS * mob.y * rg * (pc - (F.cg_x[n] + hold_u(F, n))); }
            if (mob.z > 0.0f) { uint32_t n = c + F.W + 1u; f[2] = W_DIAG_FLUX * mob.z * rg * (pc - (F.cg_x[n] + hold_u(F, n))) * 0.70710678f; }
            if (mob.w > 0.0f) { uint32_t n = c - F.W + 1u; f[3] = W_DIAG_FLUX * mob.w * rg * (pc - (F.cg_x[n] + hold_u(F, n))) * 0.70710678f; }
            if (h0 > 0.0f) {
                float kc = F.P.front3[1];
                float tc = F.cg_t[c];
                if (x + 1u < F.W) { uint32_t n = c + 1u; if (F.h[n] > 0.0f && cell_active(F, x + 1u, y)) f[0] = f[0] + W_AXIS * kc * (F.cg_t[n] - tc); }
                if (y + 1u < F.H) { uint32_t n = c + F.W; if (F.h[n] > 0.0f && cell_active(F, x, y + 1u)) f[1] = f[1] + W_AXIS * kc * (F.cg_t[n] - tc); }
                if (x + 1u < F.W && y + 1u < F.H) { uint32_t n = c + F.W + 1u; if (F.h[n] > 0.0f && cell_active(F, x + 1u, y + 1u)) f[2] = f[2] + W_DIAG_FLUX * kc * (F.cg_t[n] - tc) * 0.70710678f; }
                if (x + 1u < F.W && y > 0u) { uint32_t n = c - F.W + 1u; if (F.h[n] > 0.0f && cell_active(F, x + 1u, y - 1u)) f[3] = f[3] + W_DIAG_FLUX * kc * (F.cg_t[n] - tc) * 0.70710678f; }
            }
            float4 a = F.acc[c];
            float fw = F.P.front[3];
            F.acc[c] = make_float4(a.x + f[0] * fw, a.y + f[1] * fw, a.z + f[2] * fw, a.w + f[3] * fw);
            F.cg_q[c] = hn;
            moved = cz_abs(hn - h0) > MOVE_EPS;
            dv = (double)(hn - h0);
            tv = (double)hn;
        }
        double sd = block_sum_d(dv);
        double st = block_sum_d(tv);
        if (threadIdx.x == 0) {
            F.cg_part[_a] = sd;
            F.cg_part2[_a] = st;
        }
        if (__syncthreads_or(moved ? 1 : 0) && threadIdx.x == 0) F.tile_moved[_t] = 1u;
    }
}

// cg_rescale (film.wgsl:1224-1233) applied to the height left by k_cg_finish.
template <class WS>
__global__ __launch_bounds__(256) void k_cg_rescale(WS ws) {
    CZ_FOR_TILES(ws) {
        CZ_TILE_CELL(F)
        if (!in_grid) continue;
        float hn = F.cg_q[c];
        float s1y = F.cg_s[5], s1z = F.cg_s[6];
        if (s1z <= 0.0f) {
            F.h[c] = hn;
            continue;
        }
        float fct = 1.0f - s1y / s1z;
        F.h[c] = cz_max(hn * fct, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// The item lists of one batched frame (one block per position k of the
// substep order): member order[k] appends its active tiles at the sum of the
// active counts of the members before it, so the items of the members still
// stepping at substep s are a prefix of the list; the same for the implicit
// order.  Block 0 writes the prefix lengths.
constexpr uint32_t CZ_PLAN_MAX = 1024u;

__global__ __launch_bounds__(256) void k_batch_plan(BatchPlan B) {
    __shared__ uint32_t cnt[CZ_PLAN_MAX];
    __shared__ uint32_t icnt[CZ_PLAN_MAX];
    __shared__ uint32_t ccnt[CZ_PLAN_MAX];
    __shared__ uint32_t red[3][256];
    const uint32_t k = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < B.members; j += 256u) {
        const FilmDev& G = desc_at(B.films, B.order[j] * CZ_VARIANTS);
        cnt[j] = *G.active_count;
        ccnt[j] = *G.cure_count;
        icnt[j] = (B.imp[j] > 0u) ? *desc_at(B.films, B.iorder[j] * CZ_VARIANTS).active_count : 0u;
    }
    __syncthreads();
    uint32_t so = 0u, io = 0u, co = 0u;
    for (uint32_t j = threadIdx.x; j < k; j += 256u) {
        so += cnt[j];
        io += icnt[j];
        co += ccnt[j];
    }
    red[0][threadIdx.x] = so;
    red[1][threadIdx.x] = io;
    red[2][threadIdx.x] = co;
    __syncthreads();
    for (uint32_t w = 128u; w > 0u; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    const uint32_t off_s = red[0][0], off_i = red[1][0], off_c = red[2][0];
    {
        const uint32_t m = B.order[k];
        const FilmDev& F = desc_at(B.films, m * CZ_VARIANTS);
        for (uint32_t a = threadIdx.x; a < cnt[k]; a += 256u)
            B.items_sub[off_s + a] = (m << CZ_ITEM_TILE_BITS) | F.tile_list[a];
        for (uint32_t a = threadIdx.x; a < ccnt[k]; a += 256u)
            B.items_cure[off_c + a] = (m << CZ_ITEM_TILE_BITS) | F.cure_list[a];
        if (k + 1u == B.members && threadIdx.x == 0) *B.n_cure = off_c + ccnt[k];
    }
    if (B.imp[k] > 0u) {
        const uint32_t m = B.iorder[k];
        const FilmDev& F = desc_at(B.films, m * CZ_VARIANTS);
        for (uint32_t a = threadIdx.x; a < icnt[k]; a += 256u)
            B.items_imp[off_i + a] = (m << CZ_ITEM_TILE_BITS) | F.tile_list[a];
        if (threadIdx.x == 0) {
            B.imp_members[k] = m;
            B.imp_bases[k] = off_i;
        }
    }
    if (k != 0u) return;
    for (uint32_t s = threadIdx.x; s <= B.max_sub; s += 256u) {
        uint32_t n = 0u;
        for (uint32_t j = 0; j < B.members; j++)
            if (s == B.max_sub || B.sub[j] > s) n += cnt[j];
        B.lim_sub[s] = n;
    }
    for (uint32_t s = threadIdx.x; s < B.max_imp; s += 256u) {
        uint32_t n = 0u;
        for (uint32_t j = 0; j < B.members; j++)
            if (B.imp[j] > s) n += icnt[j];
        B.lim_imp[s] = n;
    }
}

// Instantiations: full-grid kernels over one canvas's tiles or a batch's
// static all-tile lists, tile kernels over one canvas's active list or a
// batch's union lists.
#define CZ_INST1(name, WSX, ...) template __global__ void name<WSX>(WSX, ##__VA_ARGS__);
CZ_INST1(k_cure_clear, OneFilmAll, int)
CZ_INST1(k_cure_clear, ManyFilms, int)
CZ_INST1(k_reduce_pool, OneFilmAll)
CZ_INST1(k_reduce_pool, ManyFilms)
#define CZ_INST2(name, ...) CZ_INST1(name, OneFilm, ##__VA_ARGS__) CZ_INST1(name, ManyFilms, ##__VA_ARGS__)
CZ_INST2(k_lap_plug, int, int, int)
CZ_INST2(k_lap_tail, int, int)
CZ_INST2(k_lap_plug_w, int, int, int)
CZ_INST2(k_fluxes_w)
CZ_INST2(k_limit_apply_w)
CZ_INST2(k_flux_apply_w)
CZ_INST1(k_substep_fused, OneFilm, const float*, float*, int, int)
CZ_INST1(k_lap_tail_w, OneFilm, int, int)
CZ_INST1(k_shed, OneFilm)
CZ_INST1(k_slump, OneFilm)
CZ_INST1(k_mark, OneFilmAll)
CZ_INST1(k_shade, OneFilmAll)
CZ_INST1(k_mark, ManyFilms)
CZ_INST1(k_cure, OneFilmCure)
CZ_INST1(k_cure, ManyFilms)
CZ_INST2(k_fluxes)
CZ_INST2(k_fluxes_faces)
CZ_INST2(k_limit_apply)
CZ_INST2(k_flux_apply)
CZ_INST2(k_strata_split)
CZ_INST2(k_strata_limit)
CZ_INST2(k_strata_update)
CZ_INST2(k_strata_fill)
CZ_INST2(k_cg_setup)
CZ_INST2(k_cg_lap, uint32_t)
CZ_INST2(k_cg_init)
CZ_INST2(k_cg_iter_a)
CZ_INST2(k_cg_iter_b)
CZ_INST2(k_cg_iter_c)
CZ_INST2(k_cg_finish)
CZ_INST2(k_cg_rescale)
#define CZ_INST3(name, ...) CZ_INST1(name, OneCanvas, ##__VA_ARGS__) CZ_INST1(name, ManyCanvases, ##__VA_ARGS__)
CZ_INST3(k_compact_tiles)
CZ_INST3(k_frame_stats)
CZ_INST3(k_cg_reduce, int)

}  // namespace cz
