// film_kernels.h - declarations of the gfx950 kernels of the film step.
#pragma once
#include "film_dev.h"

namespace cz {
__global__ void k_deposit(FilmDev F);
__global__ void k_mark_tiles(FilmDev F);
__global__ void k_compact_tiles(FilmDev F);
__global__ void k_cure_clear(FilmDev F, int do_clear);
__global__ void k_lap_plug(FilmDev F, int do_nose, int do_plug);
__global__ void k_fluxes(FilmDev F);
__global__ void k_limit_apply(FilmDev F);
__global__ void k_delta_sync(FilmDev F, uint32_t* dst);
__global__ void k_strata_split(FilmDev F);
__global__ void k_reduce_pool(FilmDev F);
__global__ void k_frame_stats(FilmDev F);
__global__ void k_strata_limit(FilmDev F);
__global__ void k_strata_update(FilmDev F);
__global__ void k_strata_copy_back(FilmDev F);
__global__ void k_strata_fill(FilmDev F, const float* tin, float* tout);
__global__ void k_cg_setup(FilmDev F);
__global__ void k_cg_lap(FilmDev F, uint32_t which);
__global__ void k_cg_init(FilmDev F);
__global__ void k_cg_iter_a(FilmDev F);
__global__ void k_cg_iter_b(FilmDev F);
__global__ void k_cg_iter_c(FilmDev F);
__global__ void k_cg_reduce(FilmDev F, int mode);
__global__ void k_cg_finish(FilmDev F);
__global__ void k_cg_rescale(FilmDev F);
}  // namespace cz
