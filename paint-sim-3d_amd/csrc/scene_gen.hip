// scene_gen.hip - seeded blob canvases generated on the device
// (cz_film_seed_blobs) and the host twin of the per-cell generator used to
// check it (cz_scene_cells_host).  Canvases created by a serving loop or a
// large batch start from these instead of a host-built state.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "film_internal.h"
#include "scene_gen.h"

using namespace cz;
using namespace czi;

namespace {

// Writes the state parts and, per block, the FP64 sum of its relief cells
// into F.cg_part[block] (the weave volume of the mass report).
__global__ __launch_bounds__(256) void k_seed_blobs(FilmDev F, ScenePlan P) {
    __shared__ double part[4];
    const uint64_t n = (uint64_t)F.W * F.H;
    double msum = 0.0;
    for (uint64_t c = (uint64_t)blockIdx.x * 256u + threadIdx.x; c < n; c += (uint64_t)gridDim.x * 256u) {
        const uint32_t i = (uint32_t)(c % F.W), j = (uint32_t)(c / F.W);
        SceneCell o;
        scene_cell(P, i, j, o);
        F.h[c] = o.h;
        F.m[c] = o.m;
        set_rheo(F, (uint32_t)c, make_float2(o.rheo[0], o.rheo[1]));
        F.strata_h[c] = make_float4(o.h, 0.0f, 0.0f, 0.0f);
        F.strata_cure[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        F.strata_pig[c * 8u] = make_uint4(o.pig0[0], o.pig0[1], o.pig0[2], o.pig0[3]);
        F.strata_pig[c * 8u + 1u] = make_uint4(o.pig0[4], o.pig0[5], o.pig0[6], o.pig0[7]);
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t q = 2u; q < 8u; q++) F.strata_pig[c * 8u + q] = z;
        uint4* sr = (uint4*)(F.sub_refl + c * 8u);
        sr[0] = make_uint4(P.sub[0], P.sub[1], P.sub[2], P.sub[3]);
        sr[1] = make_uint4(P.sub[4], P.sub[5], P.sub[6], P.sub[7]);
        F.lay[c] = 0u;
        msum += (double)o.m;
    }
    for (int off = 32; off > 0; off >>= 1) msum += __shfl_down(msum, off, 64);
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = msum;
    __syncthreads();
    if (threadIdx.x == 0) F.cg_part[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

float unit(uint64_t seed, uint32_t k) {
    uint64_t u = cz_splitmix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1u));
    return ((float)(uint32_t)(u >> 40) + 0.5f) * 5.9604644775390625e-8f;
}

int make_plan(uint32_t w, uint32_t h, float dx, const cz_scene* s, ScenePlan& P) {
    if (s->nblobs > (uint32_t)SCENE_MAX_BLOBS) return set_err(CZ_EINVAL, "scene: at most %d blobs", SCENE_MAX_BLOBS);
    if (!(s->sigma_lo > 0.0f) || s->sigma_hi < s->sigma_lo || !(s->peak_lo > 0.0f) || s->peak_hi < s->peak_lo ||
        !(s->tau_lo > 0.0f) || s->tau_hi < s->tau_lo || !(s->mu_lo > 0.0f) || s->mu_hi < s->mu_lo || !(s->alpha > 0.0f))
        return set_err(CZ_EINVAL, "scene: bad ranges");
    if (s->n_hi != 0.0f && !(s->n_lo > 0.0f && s->n_hi >= s->n_lo && s->n_hi <= 4.0f))
        return set_err(CZ_EINVAL, "scene: power-law index range must lie in (0, 4]");
    memset(&P, 0, sizeof(P));
    P.seed = s->seed;
    P.nblobs = s->nblobs;
    P.w = w;
    P.h = h;
    P.ltau0 = cz_logf(s->tau_lo);
    P.ltau1 = cz_logf(s->tau_hi);
    P.lmu0 = cz_logf(s->mu_lo);
    P.lmu1 = cz_logf(s->mu_hi);
    P.alpha = s->alpha;
    P.n0 = s->n_hi != 0.0f ? s->n_lo : 0.0f;
    P.n1 = s->n_hi;
    P.weave_amp = s->weave_um * 1e-6f;
    P.dx = dx;
    P.wk = (2.0f * 3.14159265358979323846f) / 1.4e-3f;
    cz_substrate_refl(1, 1, P.sub);
    const float lp0 = cz_logf(s->peak_lo), lp1 = cz_logf(s->peak_hi);
    for (uint32_t k = 0; k < s->nblobs; k++) {
        SceneBlob& b = P.blobs[k];
        b.cx = unit(s->seed, 5u * k) * (float)w;
        b.cy = unit(s->seed, 5u * k + 1u) * (float)h;
        const float sg = s->sigma_lo + (s->sigma_hi - s->sigma_lo) * unit(s->seed, 5u * k + 2u);
        b.nhalf = -0.5f / (sg * sg);
        b.r2 = (3.0f * sg) * (3.0f * sg);
        b.pk = cz_expf(lp0 + (lp1 - lp0) * unit(s->seed, 5u * k + 3u));
    }
    return CZ_OK;
}

}  // namespace

extern "C" {

void cz_scene_default(cz_scene* s) {
    memset(s, 0, sizeof(*s));
    s->seed = 1;
    s->nblobs = 64;
    s->sigma_lo = 2.0f;
    s->sigma_hi = 40.0f;
    s->peak_lo = 20e-6f;
    s->peak_hi = 2e-3f;
    s->tau_lo = 0.1f;
    s->tau_hi = 50.0f;
    s->mu_lo = 0.5f;
    s->mu_hi = 50.0f;
    s->alpha = 0.3f;
    s->weave_um = 25.0f;
}

int cz_film_seed_blobs(cz_film* f, const cz_scene* s) {
    if (!f || !s) return set_err(CZ_EINVAL, "null argument");
    if (f->batch) return set_err(CZ_EINVAL, "seed_blobs: the film is stepped by a batch (cz_batch_remove it first)");
    ScenePlan P;
    int rc = make_plan(f->W, f->H, f->dx, s, P);
    if (rc) return rc;
    rc = wait_idle(f);
    if (rc) return rc;
    const uint64_t n = f->n;
    // one partial per block in cg_part (nt16 >= n / 256 entries)
    const uint32_t grid = (uint32_t)std::min<uint64_t>(std::min<uint64_t>((n + 255u) / 256u, f->nt16), 8192u);
    drop_stale_error();
    hipLaunchKernelGGL(k_seed_blobs, dim3(grid), dim3(256), 0, f->stream, f->dev(), P);
    HIP_TRY(hipGetLastError());
    std::vector<double> parts(grid);
    HIP_TRY(hipMemcpyAsync(parts.data(), f->cg_part, grid * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    rc = reset_solver_state(f);
    if (rc) return rc;
    double wsum = 0.0;
    for (double v : parts) wsum += v;
    f->weave_volume = wsum * (double)f->dx * (double)f->dx;
    f->total_deposited = 0.0;
    return CZ_OK;
}

int cz_scene_cells_host(uint32_t w, uint32_t h, float dx, const cz_scene* s, uint32_t j0, uint32_t j1, float* h_out,
                        float* m_out, float* rheo_out, uint32_t* pig0_out) {
    if (!s || !h_out || !m_out || !rheo_out || !pig0_out || j1 > h || j0 > j1)
        return set_err(CZ_EINVAL, "scene_cells_host: bad arguments");
    ScenePlan P;
    int rc = make_plan(w, h, dx, s, P);
    if (rc) return rc;
    for (uint32_t j = j0; j < j1; j++) {
        for (uint32_t i = 0; i < w; i++) {
            const size_t c = (size_t)(j - j0) * w + i;
            SceneCell o;
            scene_cell(P, i, j, o);
            h_out[c] = o.h;
            m_out[c] = o.m;
            rheo_out[2 * c] = o.rheo[0];
            rheo_out[2 * c + 1] = o.rheo[1];
            memcpy(pig0_out + 8 * c, o.pig0, sizeof(o.pig0));
        }
    }
    return CZ_OK;
}

}  // extern "C"
