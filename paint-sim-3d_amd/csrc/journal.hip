// journal.hip - the undo journal of the film step (gfx950): the pass at the
// start of a frame that saves the tiles the frame may change, once per
// generation, into a ring of tile records, and the restore that brings them
// back.  shaders/journal.wgsl and src/film/mod.rs:357-443, 1246-1275.
//
// Record layout: slot s holds one 16x16-cell tile, 53 words per cell, as
// the eight state parts one after another, each part's 256 cells in tile
// row order (the part's word k of cell li at s*13568 + 256*w0 + li*n + k,
// w0 the part's first word in save_cell order and n its words per cell).
// A tile row of a part is then contiguous on both sides of the copy.  The
// reference keeps the same 53 words cell by cell; the layout is private to
// the ring.
//
// Slot assignment is ordered: the tiles a frame journals take the sequence
// numbers head, head+1, ... in ascending tile order (k_journal_mark counts
// per block of 256 tiles, k_journal_list places them by a prefix over the
// blocks), so the ring's contents are a function of the inputs.  The
// reference takes them by an atomic add per workgroup, in whatever order the
// workgroups run.
#include "film_internal.h"

namespace {

constexpr uint32_t kCellWords = 53u;  // h, m, rheo 2, strata_h 4, strata_pig 32, strata_cure 4, sub_refl 8, lay
constexpr uint64_t kTileBytes = (uint64_t)kCellWords * 4u * 256u;

struct JournalDev {
    uint32_t* meta;   // [cap] uint4: tile, generation, sequence, valid
    uint32_t* data;   // [cap][53][256]
    uint32_t* mark;   // [tiles16] generation the tile was last saved in
    uint32_t* touch;  // [tiles16] touched since the last frame (consumed by k_journal_mark)
    uint32_t* best;   // [tiles16] restore: oldest sequence per tile
    uint8_t* bits;    // [tiles16] saved this frame
    uint32_t* cnt;    // [blocks of 256 tiles]
    uint2* list;      // [tiles16] (tile, slot) of the records to write
    uint32_t* ctl;    // [4] head, records to write, generation, tiles saved this frame
    uint32_t cap;
};

JournalDev jdev(const cz_film* f) {
    return JournalDev{f->j_meta, f->j_data, f->j_mark, f->j_touch, f->j_best, f->j_bits, f->j_cnt, f->j_list, f->j_ctl,
                      f->j_cap};
}

// tile_busy (journal.wgsl:20-33): the tile or one of its eight neighbours
// was active in the last frame (tile flag 32).
__device__ bool tile_busy(const FilmDev& F, uint32_t t) {
    const int tx = (int)(t % F.tw), ty = (int)(t / F.tw);
    bool busy = false;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int x = tx + dx, y = ty + dy;
            if (x < 0 || y < 0 || x >= (int)F.tw || y >= (int)F.th) continue;
            busy = busy || (F.tile_flags[(uint32_t)x + (uint32_t)y * F.tw] & 32u) != 0u;
        }
    return busy;
}

// journal.wgsl:79-90: a tile not yet saved in this generation is saved if
// it was touched or is busy.  Its mark is set here; the touches are cleared
// as they are read (the host sends only new ones).
__global__ __launch_bounds__(256) void k_journal_mark(FilmDev F, JournalDev J) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    bool take = false;
    if (t < F.nt16) {
        const uint32_t gen = J.ctl[2];
        const uint32_t touched = J.touch[t];
        take = J.mark[t] != gen && (touched != 0u || tile_busy(F, t));
        if (touched != 0u) J.touch[t] = 0u;
        if (take) J.mark[t] = gen;
        J.bits[t] = take ? 1u : 0u;
    }
    const uint32_t nb = (uint32_t)__syncthreads_count(take ? 1 : 0);
    if (threadIdx.x == 0u) J.cnt[blockIdx.x] = nb;
}

// Sequence numbers in ascending tile order and the slot records
// (journal.wgsl:86-89).  When a frame saves more tiles than the ring holds,
// only the last cap of them keep their slots (as if the earlier ones had
// been written and overwritten in sequence order).
__global__ __launch_bounds__(256) void k_journal_list(FilmDev F, JournalDev J) {
    __shared__ uint32_t s_base, s_total, s_w[4];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (wave == 0u) {
        uint32_t before = 0u, total = 0u;
        for (uint32_t q = lane; q < gridDim.x; q += 64u) {
            const uint32_t c = J.cnt[q];
            total += c;
            if (q < blockIdx.x) before += c;
        }
        for (int off = 32; off > 0; off >>= 1) {
            before += __shfl_xor(before, off, 64);
            total += __shfl_xor(total, off, 64);
        }
        if (lane == 0u) {
            s_base = before;
            s_total = total;
        }
    }
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const bool take = t < F.nt16 && J.bits[t] != 0u;
    const unsigned long long bm = __ballot(take);
    if (lane == 0u) s_w[wave] = (uint32_t)__popcll(bm);
    __syncthreads();
    const unsigned long long lmask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    uint32_t k = s_base + (uint32_t)__popcll(bm & lmask);
    for (uint32_t w = 0; w < wave; w++) k += s_w[w];
    const uint32_t total = s_total;
    const uint32_t skip = total > J.cap ? total - J.cap : 0u;
    if (take && k >= skip) {
        const uint32_t seq = J.ctl[0] + k;
        const uint32_t slot = seq % J.cap;
        reinterpret_cast<uint4*>(J.meta)[slot] = make_uint4(t, J.ctl[2], seq, 1u);
        J.list[k - skip] = make_uint2(t, slot);
    }
    if (blockIdx.x + 1u == gridDim.x && threadIdx.x == 0u) {
        J.ctl[1] = total - skip;
        J.ctl[3] = total;
    }
}

// A record keeps the tile's cells part by part, in save_cell's order
// (journal.wgsl:35-52): h, m, rheo, strata_h, strata_pig, strata_cure,
// sub_refl, lay, each part as 256 cells x its bytes per cell, tile row by
// tile row.  A tile row of a part is then one contiguous run both in the
// state buffer and in the record, and save and restore are copies of those
// runs in 16-byte (or 8- or 4-byte) lanes.
template <uint32_t B>
struct PartUnit;
template <> struct PartUnit<4> { using T = uint32_t; };
template <> struct PartUnit<8> { using T = uint2; };
template <> struct PartUnit<16> { using T = uint4; };

// Copy one part of one tile between the state (row pitch W cells) and a
// record (row pitch 16 cells); TO_RECORD picks the direction.  rc, rr: the
// tile's columns and rows inside the canvas.
template <uint32_t B, bool TO_RECORD>
__device__ __forceinline__ void copy_part(void* state, uint32_t* rec_words, uint32_t W, uint32_t x0, uint32_t y0,
                                          uint32_t rc, uint32_t rr) {
    constexpr uint32_t U = B < 16u ? B : 16u;   // bytes per lane
    constexpr uint32_t UB = B / U;              // lanes per cell
    constexpr uint32_t L = 16u * UB;            // lanes per tile row
    using T = typename PartUnit<U>::T;
    T* st = reinterpret_cast<T*>(state);
    T* rec = reinterpret_cast<T*>(rec_words);
#pragma unroll
    for (uint32_t it = 0; it < UB; it++) {
        const uint32_t e = threadIdx.x + it * 256u;   // 256 * UB lanes per tile
        const uint32_t row = e / L, col = e % L;
        if (row < rr && col / UB < rc) {
            const size_t s_idx = ((size_t)(y0 + row) * W + x0) * UB + col;
            if (TO_RECORD) rec[e] = st[s_idx];
            else st[s_idx] = rec[e];
        }
    }
}

// The word offset of each part in a record (256 cells x the part's words).
constexpr uint32_t kOffM = 256u * 1u, kOffRheo = 256u * 2u, kOffSh = 256u * 4u, kOffPig = 256u * 8u,
                   kOffCure = 256u * 40u, kOffSub = 256u * 44u, kOffLay = 256u * 52u;

template <bool TO_RECORD>
__device__ __forceinline__ void copy_tile(const FilmDev& F, uint32_t t, uint32_t* rec) {
    const uint32_t x0 = (t % F.tw) * 16u, y0 = (t / F.tw) * 16u;
    const uint32_t rc = min(16u, F.W - x0), rr = min(16u, F.H - y0);
    copy_part<4, TO_RECORD>(F.h, rec, F.W, x0, y0, rc, rr);
    copy_part<4, TO_RECORD>(F.m, rec + kOffM, F.W, x0, y0, rc, rr);
    copy_part<8, TO_RECORD>(F.rheo, rec + kOffRheo, F.W, x0, y0, rc, rr);
    copy_part<16, TO_RECORD>(F.strata_h, rec + kOffSh, F.W, x0, y0, rc, rr);
    copy_part<128, TO_RECORD>(F.strata_pig, rec + kOffPig, F.W, x0, y0, rc, rr);
    copy_part<16, TO_RECORD>(F.strata_cure, rec + kOffCure, F.W, x0, y0, rc, rr);
    copy_part<32, TO_RECORD>(F.sub_refl, rec + kOffSub, F.W, x0, y0, rc, rr);
    copy_part<4, TO_RECORD>(F.lay, rec + kOffLay, F.W, x0, y0, rc, rr);
}

// journal.wgsl:92-97: one block per record, one thread per cell; block 0
// advances the head once every block has read its list entries' count.
__global__ __launch_bounds__(256) void k_journal_save(FilmDev F, JournalDev J) {
    const uint32_t n = J.ctl[1];
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint2 e = J.list[k];
        copy_tile<true>(F, e.x, J.data + (size_t)e.y * kCellWords * 256u);
    }
    // no block of this launch reads the head; the next launch that does
    // (k_journal_list of the next frame, the restore) is stream-ordered
    if (blockIdx.x == 0u && threadIdx.x == 0u) J.ctl[0] = J.ctl[0] + J.ctl[3];
}

// slot_valid (journal.wgsl:73-78)
__device__ __forceinline__ bool slot_valid(const JournalDev& J, uint32_t s, uint4& mt) {
    mt = reinterpret_cast<const uint4*>(J.meta)[s];
    const uint32_t head = J.ctl[0];
    return mt.w == 1u && mt.z < head && head - mt.z <= J.cap;
}

__global__ __launch_bounds__(256) void k_journal_best(JournalDev J, uint32_t nt) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < nt) J.best[t] = 0xffffffffu;
}

// restore_scan (journal.wgsl:100-107)
__global__ __launch_bounds__(256) void k_journal_scan(FilmDev F, JournalDev J, uint32_t target) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= J.cap) return;
    uint4 mt;
    if (!slot_valid(J, s, mt) || mt.y < target || mt.x >= F.nt16) return;
    atomicMin(&J.best[mt.x], mt.z);
}

// restore_apply (journal.wgsl:109-123)
__global__ __launch_bounds__(256) void k_journal_apply(FilmDev F, JournalDev J, uint32_t target) {
    const uint32_t s = blockIdx.x;
    uint4 mt;
    if (!slot_valid(J, s, mt) || mt.y < target || mt.x >= F.nt16 || J.best[mt.x] != mt.z) return;
    if (threadIdx.x == 0u) F.tile_moved[mt.x] = 1u;
    copy_tile<false>(F, mt.x, J.data + (size_t)s * kCellWords * 256u);
}

uint32_t cap_for(uint64_t budget) {
    const uint64_t c = budget / kTileBytes;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(c, 64u), 1u << 26);
}

template <typename T>
int jalloc(cz_film* f, T** p, size_t count) {
    void* v = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess) return set_err(CZ_ENOMEM, "journal: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    f->allocs.push_back(v);
    f->bytes += bytes;
    *p = (T*)v;
    HIP_TRY(hipMemsetAsync(v, 0, bytes, f->stream));
    return CZ_OK;
}

// Journal::new (src/film/mod.rs:372-403), at the first nonzero generation.
// The ring exists once j_cap is set; a failed allocation leaves j_cap 0 and
// a later call allocates what is still missing.
int journal_alloc(cz_film* f) {
    if (f->j_cap) return CZ_OK;
    const uint32_t cap = cap_for(f->j_budget);
    const size_t nt = f->cap_nt16;
    int rc = CZ_OK;
    if (!rc && !f->j_meta) rc = jalloc(f, &f->j_meta, (size_t)cap * 4u);
    if (!rc && !f->j_data) rc = jalloc(f, &f->j_data, (size_t)cap * kCellWords * 256u);
    if (!rc && !f->j_mark) rc = jalloc(f, &f->j_mark, nt);
    if (!rc && !f->j_touch) rc = jalloc(f, &f->j_touch, nt);
    if (!rc && !f->j_best) rc = jalloc(f, &f->j_best, nt);
    if (!rc && !f->j_bits) rc = jalloc(f, &f->j_bits, nt);
    if (!rc && !f->j_cnt) rc = jalloc(f, &f->j_cnt, (nt + 255u) / 256u);
    if (!rc && !f->j_list) rc = jalloc(f, &f->j_list, nt);
    if (!rc && !f->j_ctl) rc = jalloc(f, &f->j_ctl, 4);
    if (!rc) {
        hipError_t e = hipSuccess;
        if (!f->j_ctl_h) e = hipHostMalloc((void**)&f->j_ctl_h, 4 * sizeof(uint32_t), hipHostMallocDefault);
        if (e == hipSuccess && !f->j_touch_h)
            e = hipHostMalloc((void**)&f->j_touch_h, nt * sizeof(uint32_t), hipHostMallocDefault);
        if (e != hipSuccess) rc = set_err(CZ_ENOMEM, "journal: hipHostMalloc: %s", hipGetErrorString(e));
    }
    if (rc) return rc;  // what was allocated stays with the film and is freed with it
    f->j_touch_host.assign(nt, 0u);
    f->j_touched = false;
    rc = czi::wait_idle(f);
    if (!rc) f->j_cap = cap;
    return rc;
}

}  // namespace

namespace czi {

void journal_record(cz_film* f) {
    const FilmDev F = f->dev();
    const JournalDev J = jdev(f);
    const uint32_t blocks = (f->nt16 + 255u) / 256u;
    hipLaunchKernelGGL(k_journal_mark, dim3(blocks), dim3(256), 0, f->stream, F, J);
    hipLaunchKernelGGL(k_journal_list, dim3(blocks), dim3(256), 0, f->stream, F, J);
    hipLaunchKernelGGL(k_journal_save, dim3(std::min(f->nt16, czi::kMaxTileBlocks)), dim3(256), 0, f->stream, F, J);
}

int journal_flush(cz_film* f) {
    if (f->j_gen == 0u || !f->j_touched) return CZ_OK;
    int rc = wait_idle(f);  // the pinned staging may still feed the last copy
    if (rc) return rc;
    memcpy(f->j_touch_h, f->j_touch_host.data(), (size_t)f->nt16 * sizeof(uint32_t));
    HIP_TRY(hipMemcpyAsync(f->j_touch, f->j_touch_h, (size_t)f->nt16 * sizeof(uint32_t), hipMemcpyHostToDevice,
                           f->stream));
    std::fill(f->j_touch_host.begin(), f->j_touch_host.end(), 0u);
    f->j_touched = false;
    return CZ_OK;
}

void journal_touch_rect(cz_film* f, float minx, float miny, float maxx, float maxy) {
    if (f->j_gen == 0u || !(std::isfinite(minx) && std::isfinite(miny) && std::isfinite(maxx) && std::isfinite(maxy)))
        return;
    const float tile_m = 16.0f * f->dx;
    const float mx = (float)(f->tw - 1u), my = (float)(f->th - 1u);
    // ((v / tile_m).floor() as i64).clamp(0, n - 1): the cast saturates, so
    // the clamp can be taken in float
    const int64_t x0 = (int64_t)fminf(fmaxf(floorf(minx / tile_m), 0.0f), mx);
    const int64_t x1 = (int64_t)fminf(fmaxf(floorf(maxx / tile_m), 0.0f), mx);
    const int64_t y0 = (int64_t)fminf(fmaxf(floorf(miny / tile_m), 0.0f), my);
    const int64_t y1 = (int64_t)fminf(fmaxf(floorf(maxy / tile_m), 0.0f), my);
    if (maxx < 0.0f || maxy < 0.0f) return;
    for (int64_t y = y0; y <= y1; y++)
        for (int64_t x = x0; x <= x1; x++) f->j_touch_host[(size_t)(y * f->tw + x)] = 1u;
    f->j_touched = true;
}

int journal_reset(cz_film* f) {
    f->j_gen = 0u;
    if (!f->j_cap) return CZ_OK;
    HIP_TRY(hipMemsetAsync(f->j_meta, 0, (size_t)f->j_cap * 16u, f->stream));
    HIP_TRY(hipMemsetAsync(f->j_mark, 0, (size_t)f->cap_nt16 * 4u, f->stream));
    HIP_TRY(hipMemsetAsync(f->j_touch, 0, (size_t)f->cap_nt16 * 4u, f->stream));
    HIP_TRY(hipMemsetAsync(f->j_ctl, 0, 16u, f->stream));
    std::fill(f->j_touch_host.begin(), f->j_touch_host.end(), 0u);
    f->j_touched = false;
    return CZ_OK;
}

void journal_free_host(cz_film* f) {
    if (f->j_ctl_h) hipHostFree(f->j_ctl_h);
    if (f->j_touch_h) hipHostFree(f->j_touch_h);
    f->j_ctl_h = f->j_touch_h = nullptr;
}

}  // namespace czi

using namespace czi;

extern "C" {

int cz_film_journal_gen(cz_film* f, uint32_t gen) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    if (f->batch) return set_err(CZ_EINVAL, "journal_gen: the film is stepped by a batch (cz_batch_remove it first)");
    if (gen != 0u) {
        int rc = journal_alloc(f);
        if (rc) return rc;
    }
    if (f->j_cap) {
        int rc = wait_idle(f);
        if (rc) return rc;
        f->j_ctl_h[2] = gen;
        HIP_TRY(hipMemcpyAsync(f->j_ctl + 2, f->j_ctl_h + 2, sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
        rc = wait_idle(f);
        if (rc) return rc;
    }
    f->j_gen = gen;
    return CZ_OK;
}

int cz_film_journal_budget(cz_film* f, uint64_t bytes) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    if (f->j_cap) return set_err(CZ_EINVAL, "journal_budget: the ring exists (set the budget before the first nonzero generation)");
    f->j_budget = bytes;
    return CZ_OK;
}

int cz_film_journal_head(cz_film* f, uint32_t* head, uint32_t* cap) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    uint32_t h = 0u;
    if (f->j_cap) {
        int rc = wait_idle(f);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(f->j_ctl_h, f->j_ctl, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, f->stream));
        rc = wait_idle(f);
        if (rc) return rc;
        h = f->j_ctl_h[0];
    }
    if (head) *head = h;
    if (cap) *cap = f->j_cap ? f->j_cap : cap_for(f->j_budget);
    return CZ_OK;
}

int cz_film_touch_rect(cz_film* f, float minx, float miny, float maxx, float maxy) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    if (f->batch) return set_err(CZ_EINVAL, "touch_rect: the film is stepped by a batch (cz_batch_remove it first)");
    journal_touch_rect(f, minx, miny, maxx, maxy);
    return CZ_OK;
}

int cz_film_restore_journal(cz_film* f, uint32_t target) {
    if (!f) return set_err(CZ_EINVAL, "null film");
    if (f->batch) return set_err(CZ_EINVAL, "restore_journal: the film is stepped by a batch (cz_batch_remove it first)");
    if (!f->j_cap) return CZ_OK;  // nothing was ever recorded
    const FilmDev F = f->dev();
    const JournalDev J = jdev(f);
    drop_stale_error();
    hipLaunchKernelGGL(k_journal_best, dim3((f->nt16 + 255u) / 256u), dim3(256), 0, f->stream, J, f->nt16);
    hipLaunchKernelGGL(k_journal_scan, dim3((f->j_cap + 255u) / 256u), dim3(256), 0, f->stream, F, J, target);
    hipLaunchKernelGGL(k_journal_apply, dim3(f->j_cap), dim3(256), 0, f->stream, F, J, target);
    HIP_TRY(hipGetLastError());
    return CZ_OK;
}

}  // extern "C"
