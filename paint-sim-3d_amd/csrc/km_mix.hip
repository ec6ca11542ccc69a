// km_mix.hip - Kubelka-Munk spectral mix on gfx950 MFMA.
//
// The pigment mix of km.wgsl:17-33 is the contraction
//   K[c][b] = sum_i conc[c][i] K_tab[i][b],  S[c][b] = sum_i conc[c][i] S_tab[i][b]
// of (cells x 12) by (12 x bands), followed per (cell, band) by the
// infinite-layer reflectance R_inf = 1 + a - sqrt(a^2 + 2a), a = K/S (the
// t -> infinity limit of km_layer, km.wgsl:1-11), evaluated in the
// cancellation-free form 1 / (1 + a + sqrt(a (a + 2))) with the hardware
// reciprocal and square root.
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, one wave64 per 16-cell group):
//   concentrations (32 x 16 cells): k = 0..11 concentrations, 12..15 zero,
//     16..27 the same concentrations again, 28..31 zero (BF16 inputs, exact);
//   table (16 bands x 32): k = 0..11 the BF16 high part of the FP32 table,
//     16..27 its BF16 residual (table - high), so that one instruction sums
//     conc x (high + low) with FP32 accumulation;
//   tiles: NT tiles of 16 bands for K, then NT tiles for S, so that K and S
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

// R_inf = 1 / (1 + a + sqrt(a (a + 2))) with a = K / S, multiplied through
// by S: S / (S + K + sqrt(K (K + 2 S))), exact algebra that needs one
// reciprocal and one square root per (cell, band) instead of two
// reciprocals and a square root.  Both are the hardware instructions
// (v_rcp_f32, v_sqrt_f32, each within 1 ulp): the relative error stays within
// a few 1e-7, far inside the 1e-4 bar (the BF16 table split contributes
// ~1e-6).  The transcendental unit is the kernel's VALU bound, so the count of
// these two instructions is what sets its speed.  A cell without pigment
// (S = K = 0) gives 0, as km_layer does for k + s < 1e-6 (km.wgsl:3); the
// guard is a select, not a branch.
__device__ __forceinline__ float r_inf(float k, float s) {
    const float q = __builtin_amdgcn_sqrtf(k * fmaf(2.0f, s, k));
    const float r = s * __builtin_amdgcn_rcpf(s + k + q);
    return (s > 0.0f) ? r : 0.0f;
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

// The product is formed transposed, D = table^T (16 bands x 32) . conc^T
// (32 x 16 cells): the prepared table tile is the A operand and the lane's
// concentrations the B operand (the two operand layouts of the instruction
// are the same, lane (grp, i) holding k = 8 grp .. 8 grp + 7 of row/column
// i).  The accumulator of lane (grp, i) then holds bands 4 grp .. 4 grp + 3
// of the tile for cell i, K and S of a band in the same lane and slot: a lane
// puts its four reflectances into the wave's LDS stage with one 16-byte
// write.  The stage (the group's 16 cells x bands floats, row-major as in
// HBM) is then written out as one contiguous run of 16-byte stores, whole
// cache lines: storing each lane's 16 bytes straight to HBM leaves partial
// lines and doubles the bytes written.
template <int NT>
__global__ __launch_bounds__(256) void k_km_mix(const uint16_t* __restrict__ conc, const u16x8* __restrict__ bops,
                                                uint64_t cells, uint32_t bands, float* __restrict__ out) {
    __shared__ __align__(16) float stage[4][16 * 16 * KM_MAX_TILES];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 15u, grp = lane >> 4;
    // table operands: constant for the whole launch
    bf16x8 tab[2 * NT];
#pragma unroll
    for (int t = 0; t < 2 * NT; t++) {
        u16x8 v = bops[t * 64 + lane];
        tab[t] = __builtin_bit_cast(bf16x8, v);
    }
    const uint64_t groups = (cells + 15u) / 16u;
    const uint64_t batches = (groups + KM_BATCH - 1) / KM_BATCH;
    const uint64_t wave_id = (uint64_t)blockIdx.x * 4u + wave;
    const uint64_t nwaves = (uint64_t)gridDim.x * 4u;
    const bool quads = (bands & 3u) == 0u;
    float* st = stage[wave];
    for (uint64_t bi = wave_id; bi < batches; bi += nwaves) {
        u16x8 cv[KM_BATCH];
#pragma unroll
        for (int g = 0; g < KM_BATCH; g++) {
            const uint64_t base = (bi * KM_BATCH + g) * 16u;
            cv[g] = load_a(conc, base + row, cells, grp);
        }
#pragma unroll
        for (int g = 0; g < KM_BATCH; g++) {
            const uint64_t base = (bi * KM_BATCH + g) * 16u;
            if (base >= cells) break;
            const bf16x8 c = __builtin_bit_cast(bf16x8, cv[g]);
            f32x4 acc[2 * NT];
#pragma unroll
            for (int t = 0; t < 2 * NT; t++) {
                f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tab[t], c, z, 0, 0, 0);
            }
            float* sr = st + row * bands;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const uint32_t b0 = 16u * t + 4u * grp;
                if (b0 >= bands) continue;
                if (quads) {
                    f32x4 v;
                    v[0] = r_inf(acc[t][0], acc[NT + t][0]);
                    v[1] = r_inf(acc[t][1], acc[NT + t][1]);
                    v[2] = r_inf(acc[t][2], acc[NT + t][2]);
                    v[3] = r_inf(acc[t][3], acc[NT + t][3]);
                    *(f32x4*)(sr + b0) = v;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (b0 + j < bands) sr[b0 + j] = r_inf(acc[t][j], acc[NT + t][j]);
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            // contiguous write-out of the group's 16 x bands floats
            const uint64_t ncell = (cells - base) < 16u ? (cells - base) : 16u;
            const uint32_t nflt = (uint32_t)ncell * bands;
            float* dst = out + base * bands;
            if (quads) {
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
    // hipGetLastError reports the last failed runtime call of the thread,
    // whoever made it: drop what an earlier call left before launching, so
    // that the check below reports this launch only
    (void)hipGetLastError();
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
