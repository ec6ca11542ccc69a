// scene_gen.h - the seeded blob canvas, one cell at a time, shared by the
// device generator (k_seed_blobs) and its host twin (cz_scene_cell_host).
//
// A canvas of the plain workload: height a sum of Gaussian blobs truncated
// at 3 sigma and zeroed below 1e-9 m; every wet cell holds one stratum with
// Dirichlet(alpha) concentrations of the 12 pigments, a log-uniform yield
// stress (slot 13) and a log-uniform viscosity (slot 14 = ln mu + 10), stored
// as f16 pairs; rheology from the stored values; relief = the reference's
// canvas weave (src/film/mod.rs:212-233); substrate = the raw canvas
// reflectance (src/film/mod.rs:188-210).
//
// Randomness is counter based: every draw of cell c is SplitMix64 (Steele,
// Lea and Flood, "Fast splittable pseudorandom number generators", 2014) of
// a key derived from (seed, c) plus a draw counter, so a cell's values do not
// depend on how cells are spread over threads.  Gamma variates use Marsaglia
// and Tsang ("A simple method for generating gamma variables", 2000) with
// the shape-boost U^(1/a) for a < 1; normals use Marsaglia's polar method.
// All float arithmetic is single IEEE operations and the czmath functions,
// so host and device give the same bits.
#pragma once
#include <stdint.h>

#include "czmath.h"

namespace cz {

constexpr int SCENE_MAX_BLOBS = 64;

struct SceneBlob {
    float cx, cy;    // centre, in cells
    float nhalf;     // -0.5 / sigma^2 (cells^-2)
    float r2;        // (3 sigma)^2
    float pk;        // peak height, m
};

struct ScenePlan {
    uint64_t seed;
    uint32_t nblobs;
    uint32_t w, h;
    float ltau0, ltau1;  // ln of the yield-stress range, Pa
    float lmu0, lmu1;    // ln of the viscosity range, Pa s
    float alpha;         // Dirichlet parameter
    float n0, n1;        // per-cell power-law index range; n1 = 0: none
    float weave_amp, dx, wk;
    uint32_t sub[8];     // packed raw canvas reflectance
    SceneBlob blobs[SCENE_MAX_BLOBS];
};

CZ_HD inline uint64_t cz_splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct CellRng {
    uint64_t key;
    uint32_t ctr;
    // uniform in (0, 1): 24 random bits, centred in their bin
    CZ_HD float uniform() {
        uint64_t u = cz_splitmix64(key + 0x632BE59BD9B4E019ull * (uint64_t)(ctr++));
        return ((float)(uint32_t)(u >> 40) + 0.5f) * 5.9604644775390625e-8f;
    }
    CZ_HD float normal() {
        for (int t = 0; t < 64; t++) {
            float a = 2.0f * uniform() - 1.0f, b = 2.0f * uniform() - 1.0f;
            float s = a * a + b * b;
            if (s > 0.0f && s < 1.0f) return a * sqrtf(-2.0f * cz_logf(s) / s);
        }
        return 0.0f;
    }
    // Gamma(a), a > 0
    CZ_HD float gamma(float a) {
        float boost = 1.0f;
        if (a < 1.0f) {
            boost = cz_expf(cz_logf(uniform()) / a);
            a = a + 1.0f;
        }
        float d = a - 1.0f / 3.0f;
        float c = 1.0f / sqrtf(9.0f * d);
        for (int t = 0; t < 64; t++) {
            float x = normal();
            float v = 1.0f + c * x;
            if (v <= 0.0f) continue;
            v = v * v * v;
            float u = uniform();
            float x2 = x * x;
            if (u < 1.0f - 0.0331f * x2 * x2) return d * v * boost;
            if (cz_logf(u) < 0.5f * x2 + d * (1.0f - v + cz_logf(v))) return d * v * boost;
        }
        return d * boost;
    }
};

struct SceneCell {
    float h, m;
    float rheo[2];
    uint32_t pig0[8];   // stratum 0 (the others are empty)
};

CZ_HD inline void scene_cell(const ScenePlan& P, uint32_t i, uint32_t j, SceneCell& o) {
    const float xc = (float)i + 0.5f, yc = (float)j + 0.5f;
    float hv = 0.0f;
    for (uint32_t k = 0; k < P.nblobs; k++) {
        const SceneBlob& b = P.blobs[k];
        float dxc = xc - b.cx, dyc = yc - b.cy;
        float d2 = dxc * dxc + dyc * dyc;
        if (d2 <= b.r2) hv = hv + b.pk * cz_expf(d2 * b.nhalf);
    }
    if (hv < 1e-9f) hv = 0.0f;
    o.h = hv;
    // relief: the canvas weave (src/film/mod.rs:212-233)
    {
        float x = xc * P.dx, y = yc * P.dx;
        uint32_t hsh = (i * 73856093u) ^ (j * 19349663u);
        hsh ^= hsh >> 13;
        hsh = hsh * 0x5bd1e995u;
        hsh ^= hsh >> 15;
        float r = (float)(hsh & 0xffffu) / 65535.0f * 2.0f - 1.0f;
        float s = (float)sin((double)(P.wk * x)) * (float)sin((double)(P.wk * y));
        o.m = P.weave_amp > 0.0f ? P.weave_amp * (1.0f + 0.6f * s + 0.4f * r) : 0.0f;
    }
    for (int q = 0; q < 8; q++) o.pig0[q] = 0u;
    o.rheo[0] = 1.0f;
    o.rheo[1] = 0.0f;
    if (!(hv > 0.0f)) return;
    CellRng rng;
    rng.key = cz_splitmix64(P.seed ^ cz_splitmix64(((uint64_t)j << 32) | (uint64_t)i));
    rng.ctr = 0u;
    float g[12];
    float gs = 0.0f;
    for (int p = 0; p < 12; p++) {
        g[p] = rng.gamma(P.alpha);
        gs = gs + g[p];
    }
    float frac[16];
    for (int p = 0; p < 12; p++) frac[p] = gs > 0.0f ? g[p] / gs : (p == 0 ? 1.0f : 0.0f);
    float tau = cz_expf(P.ltau0 + (P.ltau1 - P.ltau0) * rng.uniform());
    float lmu = P.lmu0 + (P.lmu1 - P.lmu0) * rng.uniform();
    frac[12] = 0.0f;
    frac[13] = tau;
    frac[14] = lmu + 10.0f;
    frac[15] = P.n1 > 0.0f ? P.n0 + (P.n1 - P.n0) * rng.uniform() : 0.0f;
    for (int q = 0; q < 8; q++) o.pig0[q] = cz_pack2(frac[2 * q], frac[2 * q + 1]);
    o.rheo[0] = cz_expf(cz_unpack_lo(o.pig0[7]) - 10.0f);
    o.rheo[1] = cz_unpack_hi(o.pig0[6]);
}

}  // namespace cz
