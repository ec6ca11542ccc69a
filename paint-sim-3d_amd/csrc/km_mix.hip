// km_mix.hip - Kubelka-Munk spectral mix on gfx950 MFMA.
//
// The pigment mix of km.wgsl:17-33 is the contraction
//   K[c][b] = sum_i conc[c][i] K_tab[i][b],  S[c][b] = sum_i conc[c][i] S_tab[i][b]
// of (cells x 12) by (12 x bands), followed per (cell, band) by the
// infinite-layer reflectance R_inf = 1 + a - sqrt(a^2 + 2a), a = K/S (the
// t -> infinity limit of km_layer, km.wgsl:1-11), evaluated in the
// cancellation-free form 1 / (1 + a + sqrt(a (a + 2))).
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, one wave64 per 16-cell group):
//   A (16 cells x 32): k = 0..11 concentrations, 12..15 zero, 16..27 the same
//     concentrations again, 28..31 zero (BF16 inputs, exact);
//   B (32 x 16 columns): k = 0..11 the BF16 high part of the FP32 table,
//     16..27 its BF16 residual (table - high), so that one instruction sums
//     conc x (high + low) with FP32 accumulation;
//   columns: NT tiles of 16 bands for K, then NT tiles for S, so that K and S
//     of one band land in the same lane of the accumulator.
// The op is HBM-bound (24 B read + 4*bands B written per cell); results are
// staged through LDS so that each wave writes its 16 x bands floats as one
// contiguous run with 16-byte non-temporal stores.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/canvaz_km.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

namespace {

constexpr int KM_NPIG = 12;
constexpr int KM_MAX_TILES = 3;  // bands <= 48

__device__ __forceinline__ float r_inf(float k, float s) {
    if (!(s > 0.0f)) return 0.0f;
    float a = k / s;
    return 1.0f / (1.0f + a + sqrtf(a * (a + 2.0f)));
}

// A operand of one lane for the 16-cell group at `base`: 8 concentrations of
// the lane's k-range (groups 0/2: pigments 0-7, groups 1/3: pigments 8-11).
__device__ __forceinline__ u16x8 load_a(const uint16_t* __restrict__ conc, uint64_t cell, uint64_t cells,
                                        uint32_t grp) {
    u16x8 av = {0, 0, 0, 0, 0, 0, 0, 0};
    if (cell < cells) {
        const uint16_t* cp = conc + cell * KM_NPIG;
        if ((grp & 1u) == 0u) {
            const u32x2 v0 = __builtin_nontemporal_load((const u32x2*)(cp));
            const u32x2 v1 = __builtin_nontemporal_load((const u32x2*)(cp + 4));
            av[0] = (unsigned short)(v0.x & 0xffffu); av[1] = (unsigned short)(v0.x >> 16);
            av[2] = (unsigned short)(v0.y & 0xffffu); av[3] = (unsigned short)(v0.y >> 16);
            av[4] = (unsigned short)(v1.x & 0xffffu); av[5] = (unsigned short)(v1.x >> 16);
            av[6] = (unsigned short)(v1.y & 0xffffu); av[7] = (unsigned short)(v1.y >> 16);
        } else {
            const u32x2 v2 = __builtin_nontemporal_load((const u32x2*)(cp + 8));
            av[0] = (unsigned short)(v2.x & 0xffffu); av[1] = (unsigned short)(v2.x >> 16);
            av[2] = (unsigned short)(v2.y & 0xffffu); av[3] = (unsigned short)(v2.y >> 16);
        }
    }
    return av;
}

// Each wave takes KM_BATCH consecutive 16-cell groups per iteration and issues
// all of their loads before computing, to keep enough bytes in flight.
constexpr int KM_BATCH = 4;

template <int NT>
__global__ __launch_bounds__(256) void k_km_mix(const uint16_t* __restrict__ conc, const u16x8* __restrict__ bops,
                                                uint64_t cells, uint32_t bands, float* __restrict__ out) {
    __shared__ float stage[4][16 * 16 * KM_MAX_TILES];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 15u, grp = lane >> 4;
    // B operands: constant for the whole launch
    bf16x8 b[2 * NT];
#pragma unroll
    for (int t = 0; t < 2 * NT; t++) {
        u16x8 v = bops[t * 64 + lane];
        b[t] = __builtin_bit_cast(bf16x8, v);
    }
    const uint64_t groups = (cells + 15u) / 16u;
    const uint64_t batches = (groups + KM_BATCH - 1) / KM_BATCH;
    const uint64_t wave_id = (uint64_t)blockIdx.x * 4u + wave;
    const uint64_t nwaves = (uint64_t)gridDim.x * 4u;
    float* st = stage[wave];
    for (uint64_t bi = wave_id; bi < batches; bi += nwaves) {
        u16x8 av[KM_BATCH];
#pragma unroll
        for (int g = 0; g < KM_BATCH; g++) {
            const uint64_t base = (bi * KM_BATCH + g) * 16u;
            av[g] = load_a(conc, base + row, cells, grp);
        }
#pragma unroll
        for (int g = 0; g < KM_BATCH; g++) {
            const uint64_t base = (bi * KM_BATCH + g) * 16u;
            if (base >= cells) break;
            const bf16x8 a = __builtin_bit_cast(bf16x8, av[g]);
            f32x4 acc[2 * NT];
#pragma unroll
            for (int t = 0; t < 2 * NT; t++) {
                f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], z, 0, 0, 0);
            }
            // accumulator: lane holds rows 4*grp + j (cells), column `row` (band in tile)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const uint32_t band = 16u * t + row;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t cl = 4u * grp + j;
                    if (band < bands) st[cl * bands + band] = r_inf(acc[t][j], acc[NT + t][j]);
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            // contiguous write-out of the group's 16 x bands floats
            const uint64_t ncell = (cells - base) < 16u ? (cells - base) : 16u;
            const uint32_t nflt = (uint32_t)ncell * bands;
            float* dst = out + base * bands;
            if ((bands & 3u) == 0u) {
                for (uint32_t q = lane; q < nflt / 4u; q += 64u) {
                    f32x4 v = *(const f32x4*)(st + 4u * q);
                    __builtin_nontemporal_store(v, (f32x4*)(dst + 4u * q));
                }
            } else {
                for (uint32_t q = lane; q < nflt; q += 64u) __builtin_nontemporal_store(st[q], dst + q);
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        }
    }
}

// FP32 -> BF16 bits, round to nearest even (finite inputs).
uint16_t bf16_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

thread_local char g_km_err[256] = "";

}  // namespace

extern "C" {

const char* cz_km_last_error(void) { return g_km_err; }

// B operands of the mix for ktab/stab [12][bands] (host side): 2*NT tiles of
// 64 lanes x 8 BF16 values.
int cz_km_prepare(const float* k_tab, const float* s_tab, uint32_t bands, uint16_t* ops, uint32_t* ntiles) {
    if (!k_tab || !s_tab || !ops || bands == 0 || bands > 16u * KM_MAX_TILES) {
        snprintf(g_km_err, sizeof(g_km_err), "cz_km_prepare: bands must be 1..%d", 16 * KM_MAX_TILES);
        return CZ_EINVAL;
    }
    const uint32_t NT = (bands + 15u) / 16u;
    for (uint32_t t = 0; t < 2 * NT; t++) {
        const float* tab = t < NT ? k_tab : s_tab;
        const uint32_t tt = t < NT ? t : t - NT;
        for (uint32_t lane = 0; lane < 64; lane++) {
            const uint32_t col = lane & 15u, grp = lane >> 4;
            const uint32_t band = 16u * tt + col;
            for (uint32_t j = 0; j < 8; j++) {
                const uint32_t k = 8u * grp + j;
                uint16_t v = 0;
                if (band < bands) {
                    if (k < 12u) {
                        v = bf16_rne(tab[k * bands + band]);
                    } else if (k >= 16u && k < 28u) {
                        float x = tab[(k - 16u) * bands + band];
                        float hi = bf16_to_f32(bf16_rne(x));
                        v = bf16_rne(x - hi);
                    }
                }
                ops[(t * 64u + lane) * 8u + j] = v;
            }
        }
    }
    if (ntiles) *ntiles = NT;
    return CZ_OK;
}

int cz_km_mix(const uint16_t* conc_bf16, const uint16_t* ops, uint64_t cells, uint32_t bands, float* out_refl,
              void* stream) {
    if (!conc_bf16 || !ops || !out_refl || bands == 0 || bands > 16u * KM_MAX_TILES) {
        snprintf(g_km_err, sizeof(g_km_err), "cz_km_mix: bad arguments");
        return CZ_EINVAL;
    }
    if (((uintptr_t)conc_bf16 & 7u) != 0 || ((uintptr_t)out_refl & 15u) != 0) {
        snprintf(g_km_err, sizeof(g_km_err), "cz_km_mix: conc must be 8-byte and out 16-byte aligned");
        return CZ_EINVAL;
    }
    if (cells == 0) return CZ_OK;
    const uint32_t NT = (bands + 15u) / 16u;
    const uint64_t groups = (cells + 15u) / 16u;
    const uint64_t want = (groups + 4u * KM_BATCH - 1u) / (4u * KM_BATCH);
    const uint32_t grid = (uint32_t)(want < 4096u ? want : 4096u);
    hipStream_t s = (hipStream_t)stream;
    const u16x8* b = (const u16x8*)ops;
    if (NT == 1) hipLaunchKernelGGL(k_km_mix<1>, dim3(grid), dim3(256), 0, s, conc_bf16, b, cells, bands, out_refl);
    else if (NT == 2) hipLaunchKernelGGL(k_km_mix<2>, dim3(grid), dim3(256), 0, s, conc_bf16, b, cells, bands, out_refl);
    else hipLaunchKernelGGL(k_km_mix<3>, dim3(grid), dim3(256), 0, s, conc_bf16, b, cells, bands, out_refl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_km_err, sizeof(g_km_err), "cz_km_mix launch: %s", hipGetErrorString(e));
        return CZ_EHIP;
    }
    return CZ_OK;
}

}  // extern "C"
