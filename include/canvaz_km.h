/*
 * canvaz_km.h - Kubelka-Munk spectral mix on gfx950 MFMA.
 *
 * The pigment mix of the reference's layer_ks (shaders/km.wgsl:17-33) is the
 * contraction of per-cell pigment concentrations (cells x 12) with the
 * per-pigment absorption and scattering tables (12 x bands); each (cell,
 * band) is then mapped to the infinite-layer reflectance R_inf(K/S), the
 * thick-film limit of km_layer (shaders/km.wgsl:1-11).  In the reference this
 * runs per cell inside shade_cell (shaders/shade.wgsl:35-50) and the cure
 * commit (shaders/cure.wgsl:139-153) on 16 bands; here it is a standalone
 * batched operator for any band count up to 48.
 *
 * Rules as in canvaz_film.h: 0 = ok, negative = CZ_E* error.
 */
#ifndef CANVAZ_KM_H
#define CANVAZ_KM_H

#include <stdint.h>

#include "canvaz_film.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host side: the BF16 hi/lo operand tiles of the FP32 tables k_tab, s_tab
 * [12][bands] (bands <= 48).  ops receives 2 * ceil(bands/16) * 64 * 8 u16;
 * *ntiles = ceil(bands/16). */
int cz_km_prepare(const float* k_tab, const float* s_tab, uint32_t bands, uint16_t* ops, uint32_t* ntiles);

/* Device side: out[c][b] = R_inf(K_b(c) / S_b(c)), K_b = sum_i conc[c][i] K[i][b]
 * (likewise S), conc BF16 bits [cells][12] (8-byte aligned), ops from
 * cz_km_prepare (device copy), out FP32 [cells][bands] (16-byte aligned);
 * enqueued on `stream` (hipStream_t, NULL = default). */
int cz_km_mix(const uint16_t* conc_bf16, const uint16_t* ops, uint64_t cells, uint32_t bands, float* out_refl,
              void* stream);

const char* cz_km_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CANVAZ_KM_H */
