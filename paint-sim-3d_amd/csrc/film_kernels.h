// film_kernels.h - declarations of the gfx950 kernels of the film step and
// the work sources they are instantiated for.
#pragma once
#include "film_dev.h"

namespace cz {

// A tile kernel walks the items of its launch with a grid-stride loop; one
// 256-thread block takes one 16x16 tile (or one 64x64 tile for the pool) per
// item.  The position `a` of an item in the launch is also its slot for
// per-item partial sums.  Work sources say where the items come from:

// one canvas, its ordered list of active tiles
struct OneFilm {
    static constexpr bool kXcdChunks = true;   // blocks: a contiguous chunk of the list per XCD
    FilmDev F;
    __device__ __forceinline__ uint32_t count() const { return *F.active_count; }
    __device__ __forceinline__ uint32_t item(uint32_t a) const { return a; }
    __device__ __forceinline__ const FilmDev& film(uint32_t) const { return F; }
    __device__ __forceinline__ uint32_t tile(const FilmDev& G, uint32_t it) const { return G.tile_list[it]; }
};

// one canvas, every tile 0..n-1
struct OneFilmAll {
    static constexpr bool kXcdChunks = false;  // blocks: the plain stride over every tile
    FilmDev F;
    uint32_t n;
    __device__ __forceinline__ uint32_t count() const { return n; }
    __device__ __forceinline__ uint32_t item(uint32_t a) const { return a; }
    __device__ __forceinline__ const FilmDev& film(uint32_t) const { return F; }
    __device__ __forceinline__ uint32_t tile(const FilmDev&, uint32_t it) const { return it; }
};

// Batched launches read the per-member descriptors from a device table that
// no kernel writes; device code reads it as constant memory (address space
// 4), so the descriptor fields load with scalar instructions and stay in
// SGPRs across the kernels' stores.
typedef const __attribute__((address_space(4))) FilmDev DescConst;
__device__ __forceinline__ const FilmDev& desc_at(const FilmDev* table, uint32_t k) {
    return *(const FilmDev*)&((DescConst*)table)[k];
}

// one canvas, the tiles its compaction listed for the cure pass
struct OneFilmCure {
    static constexpr bool kXcdChunks = true;   // blocks: a contiguous chunk of the list per XCD
    FilmDev F;
    __device__ __forceinline__ uint32_t count() const { return *F.cure_count; }
    __device__ __forceinline__ uint32_t item(uint32_t a) const { return a; }
    __device__ __forceinline__ const FilmDev& film(uint32_t) const { return F; }
    __device__ __forceinline__ uint32_t tile(const FilmDev& G, uint32_t it) const { return G.cure_list[it]; }
};

// Item word of the batched lists: member slot in the high bits, tile index
// in the low CZ_ITEM_TILE_BITS.
constexpr uint32_t CZ_ITEM_TILE_BITS = 20u;
constexpr uint32_t CZ_ITEM_TILE_MASK = (1u << CZ_ITEM_TILE_BITS) - 1u;
// Descriptor variants per member: 0/1 = offset-buffer parity of the frame
// start flipped 0/1 times (the substep loop), 2 = parity of the last
// substep's nose, 3 = parity after the frame.
constexpr uint32_t CZ_VARIANTS = 4u;

// many canvases: a device list of (member, tile) words and its device count
struct ManyFilms {
    static constexpr bool kXcdChunks = true;   // blocks: a contiguous chunk of the list per XCD
    const FilmDev* films;   // [member][CZ_VARIANTS]
    const uint32_t* items;
    const uint32_t* count_p;
    uint32_t variant;
    __device__ __forceinline__ uint32_t count() const { return *count_p; }
    __device__ __forceinline__ uint32_t item(uint32_t a) const { return items[a]; }
    __device__ __forceinline__ const FilmDev& film(uint32_t it) const {
        return desc_at(films, (it >> CZ_ITEM_TILE_BITS) * CZ_VARIANTS + variant);
    }
    __device__ __forceinline__ uint32_t tile(const FilmDev&, uint32_t it) const { return it & CZ_ITEM_TILE_MASK; }
};

// Single-block-per-canvas kernels (compaction, reductions, frame scalars):
// block b works on canvas `film(b)`; its per-item partial sums start at
// `base(b)` of the canvas's partial arrays.
struct OneCanvas {
    FilmDev F;
    __device__ __forceinline__ const FilmDev& film(uint32_t) const { return F; }
    __device__ __forceinline__ uint32_t base(uint32_t) const { return 0u; }
};
struct ManyCanvases {
    const FilmDev* films;      // [member][CZ_VARIANTS]
    const uint32_t* members;   // block -> member slot (NULL: identity)
    const uint32_t* bases;     // block -> first partial (NULL: 0)
    uint32_t variant;
    __device__ __forceinline__ const FilmDev& film(uint32_t b) const {
        return desc_at(films, (members ? members[b] : b) * CZ_VARIANTS + variant);
    }
    __device__ __forceinline__ uint32_t base(uint32_t b) const { return bases ? bases[b] : 0u; }
};

// Items of the batched launches of one frame, built on the device from each
// member's active count (one block per member).
//   order[k]  : member slots by substeps, descending; sub[k] their substeps
//   iorder[k] : member slots by implicit steps, descending; imp[k] their
//               implicit steps (0 for the members without)
struct BatchPlan {
    const FilmDev* films;
    const uint32_t* order;
    const uint32_t* sub;
    const uint32_t* iorder;
    const uint32_t* imp;
    uint32_t members;
    uint32_t max_sub;
    uint32_t max_imp;
    uint32_t* items_sub;   // union in substep order
    uint32_t* items_imp;   // union of the members with implicit steps, by implicit steps
    uint32_t* items_cure;  // union of the cure lists
    uint32_t* n_cure;      // its length
    uint32_t* lim_sub;     // [max_sub + 1]: items of substep s; [max_sub] = all
    uint32_t* lim_imp;     // [max_imp]
    uint32_t* imp_members; // [members] block -> member slot of the implicit reductions
    uint32_t* imp_bases;   // [members] first partial of that member in items_imp
};

__global__ void k_deposit(FilmDev F);
__global__ void k_substeps(FilmDev Fa, FilmDev Fb, uint32_t substeps, int clear_first);
__global__ void k_compact_bits(FilmDev F);
__global__ void k_compact_list(FilmDev F);
constexpr uint32_t kDepRectBlocks = 64u;  // grid of k_dep_rects (a grid-stride loop over the deposits)
__global__ void k_dep_rects(FilmDev F, float width_m, float height_m, float tile_m);
__global__ void k_dep_scan(FilmDev F);
__global__ void k_dep_fill(FilmDev F);
__global__ void k_batch_plan(BatchPlan B);

#define CZ_TILE_KERNELS(X)                                          \
    X(k_mark, (WS ws), (ws))                                        \
    X(k_shade, (WS ws), (ws))                                       \
    X(k_cure, (WS ws), (ws))                                        \
    X(k_cure_clear, (WS ws, int do_clear), (ws, do_clear))          \
    X(k_lap_plug, (WS ws, int do_nose, int do_plug, int do_clear), (ws, do_nose, do_plug, do_clear)) \
    X(k_lap_tail, (WS ws, int do_nose, int do_clear), (ws, do_nose, do_clear)) \
    X(k_lap_tail_w, (WS ws, int do_nose, int do_clear), (ws, do_nose, do_clear)) \
    X(k_lap_plug_w, (WS ws, int do_nose, int do_plug, int do_clear), (ws, do_nose, do_plug, do_clear)) \
    X(k_fluxes_w, (WS ws), (ws))                                    \
    X(k_limit_apply_w, (WS ws), (ws))                               \
    X(k_flux_apply_w, (WS ws), (ws))                                \
    X(k_fluxes, (WS ws), (ws))                                      \
    X(k_fluxes_faces, (WS ws), (ws))                                \
    X(k_limit_apply, (WS ws), (ws))                                 \
    X(k_flux_apply, (WS ws), (ws))                                  \
    X(k_substep_fused, (WS ws, const float* hr, float* hw, int do_nose, int do_clear), (ws, hr, hw, do_nose, do_clear)) \
    X(k_shed, (WS ws), (ws))                                        \
    X(k_slump, (WS ws), (ws))                                       \
    X(k_strata_split, (WS ws), (ws))                                \
    X(k_reduce_pool, (WS ws), (ws))                                 \
    X(k_strata_limit, (WS ws), (ws))                                \
    X(k_strata_update, (WS ws), (ws))                               \
    X(k_strata_fill, (WS ws), (ws))                                 \
    X(k_cg_setup, (WS ws), (ws))                                    \
    X(k_cg_lap, (WS ws, uint32_t which), (ws, which))               \
    X(k_cg_init, (WS ws), (ws))                                     \
    X(k_cg_iter_a, (WS ws), (ws))                                   \
    X(k_cg_iter_b, (WS ws), (ws))                                   \
    X(k_cg_iter_c, (WS ws), (ws))                                   \
    X(k_cg_finish, (WS ws), (ws))                                   \
    X(k_cg_rescale, (WS ws), (ws))

#define CZ_CANVAS_KERNELS(X)                                        \
    X(k_compact_tiles, (CS cs), (cs))                               \
    X(k_frame_stats, (CS cs), (cs))                                 \
    X(k_cg_reduce, (CS cs, int mode), (cs, mode))

#define CZ_DECLARE_WS(name, params, args) template <class WS> __global__ void name params;
#define CZ_DECLARE_CS(name, params, args) template <class CS> __global__ void name params;
CZ_TILE_KERNELS(CZ_DECLARE_WS)
CZ_CANVAS_KERNELS(CZ_DECLARE_CS)
#undef CZ_DECLARE_WS
#undef CZ_DECLARE_CS

}  // namespace cz
