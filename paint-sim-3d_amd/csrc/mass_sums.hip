// mass_sums.hip - the FP64 sums of the mass report on the device.
//
// compute_stats (src/main.rs:1679-1726) copies h and m to the host and sums
// them there in f64, skipping (and counting) cells whose height is not
// finite.  Here the same sums run on the GPU:
//  - k_mass_parts: one 256-thread block per chunk of MASS_CHUNK cells writes
//    the chunk's partial record (sum h, sum m in FP64, wet and non-finite
//    counts, max h, max h + m) into a workspace;
//  - k_mass_fold: one block folds the partials in chunk order into the
//    result, which stays in device memory (the caller copies it back or
//    reads it from a later kernel).
// The chunking and both tree shapes are fixed, so the sums do not depend on
// the launch or on the device: equal inputs give equal bits.
//
// Two entry points share the kernels.  cz_mass_sums takes the partials'
// workspace from the caller (cz_mass_sums_workspace bytes); the film's mass
// report passes a workspace its handle allocated at creation.
// cz_mass_sums_noalloc keeps the partials in a file-scope device array
// (`scratch`, MASS_NOALLOC_PARTS records = 192 KiB per GPU, held until the
// process ends) and calls no allocator, for callers that cannot allocate
// (code under graph capture, serving loops that forbid it).  Every call on
// one GPU uses that GPU's one array: calls that overlap on a GPU (other
// streams, other threads) race.  Every record a call reads is written by
// the same call first, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>
#include <math.h>

#include <mutex>

#include "film_internal.h"

namespace {

constexpr uint64_t MASS_CHUNK = 65536;          // cells per partial record
constexpr uint64_t MASS_NOALLOC_PARTS = 4096;   // records the static scratch holds
constexpr int kMaxDevices = 64;

struct __align__(16) MassPart {
    double sh, sm;
    unsigned long long wet, nan;
    float hmax, smax;
    uint32_t pad[2];
};
static_assert(sizeof(MassPart) == 48, "partial record layout");
static_assert(sizeof(cz_mass_totals) == 48, "cz_mass_totals layout");

__device__ __forceinline__ void mass_add(MassPart& a, const MassPart& b) {
    a.sh += b.sh;
    a.sm += b.sm;
    a.wet += b.wet;
    a.nan += b.nan;
    a.hmax = fmaxf(a.hmax, b.hmax);
    a.smax = fmaxf(a.smax, b.smax);
}

// Sum over the 256 threads of the block in a fixed order: a shuffle tree in
// each wave64, then the four waves as (w0 + w1) + (w2 + w3).  Thread 0 gets
// the result.
__device__ __forceinline__ MassPart block_sum(MassPart v) {
    __shared__ MassPart wv[4];
    for (int off = 32; off > 0; off >>= 1) {
        MassPart o;
        o.sh = __shfl_down(v.sh, off, 64);
        o.sm = __shfl_down(v.sm, off, 64);
        o.wet = __shfl_down(v.wet, off, 64);
        o.nan = __shfl_down(v.nan, off, 64);
        o.hmax = __shfl_down(v.hmax, off, 64);
        o.smax = __shfl_down(v.smax, off, 64);
        mass_add(v, o);
    }
    if ((threadIdx.x & 63u) == 0u) wv[threadIdx.x >> 6] = v;
    __syncthreads();
    MassPart r = wv[0];
    if (threadIdx.x == 0) {
        MassPart b = wv[2];
        mass_add(r, wv[1]);
        mass_add(b, wv[3]);
        mass_add(r, b);
    }
    return r;
}

// One block per chunk; thread t takes cells c0 + t, c0 + t + 256, ... in
// order (coalesced rows of 256 words), as compute_stats's loop body does:
// a non-finite height is counted and its cell skipped.
__global__ __launch_bounds__(256) void k_mass_parts(const float* __restrict__ h, const float* __restrict__ m,
                                                    uint64_t n, MassPart* __restrict__ parts) {
    const uint64_t c0 = (uint64_t)blockIdx.x * MASS_CHUNK;
    const uint64_t c1 = (c0 + MASS_CHUNK < n) ? c0 + MASS_CHUNK : n;
    MassPart v = {0.0, 0.0, 0ull, 0ull, 0.0f, 0.0f, {0u, 0u}};
    for (uint64_t c = c0 + threadIdx.x; c < c1; c += 256u) {
        const float a = h[c];
        if (!isfinite(a)) {
            v.nan += 1ull;
            continue;
        }
        const float b = m[c];
        v.sh += (double)a;
        v.sm += (double)b;
        if (a > 0.0f) v.wet += 1ull;
        v.hmax = fmaxf(v.hmax, a);
        v.smax = fmaxf(v.smax, a + b);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) parts[blockIdx.x] = v;
}

// One block: thread t folds records t, t + 256, ... in order, then the
// block tree; nparts = 0 gives the zero record.
__global__ __launch_bounds__(256) void k_mass_fold(const MassPart* __restrict__ parts, uint32_t nparts,
                                                   cz_mass_totals* __restrict__ out) {
    MassPart v = {0.0, 0.0, 0ull, 0ull, 0.0f, 0.0f, {0u, 0u}};
    for (uint32_t k = threadIdx.x; k < nparts; k += 256u) mass_add(v, parts[k]);
    v = block_sum(v);
    if (threadIdx.x == 0) {
        cz_mass_totals r;
        r.sum_h = v.sh;
        r.sum_m = v.sm;
        r.wet_cells = v.wet;
        r.nan_cells = v.nan;
        r.h_max = v.hmax;
        r.s_max = v.smax;
        r.pad[0] = 0u;
        r.pad[1] = 0u;
        *out = r;
    }
}

uint64_t parts_for(uint64_t n) { return (n + MASS_CHUNK - 1) / MASS_CHUNK; }

int launch_mass(const float* h, const float* m, uint64_t n, MassPart* parts, cz_mass_totals* out, hipStream_t st) {
    const uint64_t np = parts_for(n);
    drop_stale_error();
    if (np > 0) {
        hipLaunchKernelGGL(k_mass_parts, dim3((uint32_t)np), dim3(256), 0, st, h, m, n, parts);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mass_fold, dim3(1), dim3(256), 0, st, parts, (uint32_t)np, out);
    HIP_TRY(hipGetLastError());
    return CZ_OK;
}

int check_args(const float* h, const float* m, uint64_t n, cz_mass_totals* out) {
    if (!out || (n > 0 && (!h || !m))) return set_err(CZ_EINVAL, "mass sums: null argument");
    if (((uintptr_t)out & 15u) != 0u) return set_err(CZ_EINVAL, "mass sums: out not 16-byte aligned");
    if (parts_for(n) > 0x7fffffffull) return set_err(CZ_EINVAL, "mass sums: %llu cells is too many", (unsigned long long)n);
    return CZ_OK;
}

// The allocation-free entry point's partials.  In host code the name is a
// stand-in, not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device (each GPU holds its
// own copy of the array).
__device__ __align__(256) MassPart scratch[MASS_NOALLOC_PARTS];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once
// per device (with that device current) and kept.
int scratch_of_current_device(MassPart** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_err(CZ_EINVAL, "mass sums: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(CZ_EHIP, "mass sums: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = (MassPart*)g_scratch_ptr[dev];
    return CZ_OK;
}

}  // namespace

extern "C" uint64_t cz_mass_sums_workspace(uint64_t n) { return parts_for(n) * sizeof(MassPart); }

extern "C" int cz_mass_sums(const float* h, const float* m, uint64_t n, void* workspace, uint64_t workspace_bytes,
                            cz_mass_totals* out, void* stream) {
    int rc = check_args(h, m, n, out);
    if (rc) return rc;
    const uint64_t need = cz_mass_sums_workspace(n);
    if (need > 0 && (!workspace || workspace_bytes < need))
        return set_err(CZ_EINVAL, "mass sums: %llu cells need %llu bytes of workspace, got %llu", (unsigned long long)n,
                       (unsigned long long)need, (unsigned long long)workspace_bytes);
    if (((uintptr_t)workspace & 15u) != 0u) return set_err(CZ_EINVAL, "mass sums: workspace not 16-byte aligned");
    return launch_mass(h, m, n, (MassPart*)workspace, out, (hipStream_t)stream);
}

extern "C" int cz_mass_sums_noalloc(const float* h, const float* m, uint64_t n, cz_mass_totals* out, void* stream) {
    int rc = check_args(h, m, n, out);
    if (rc) return rc;
    if (cz_mass_sums_workspace(n) > MASS_NOALLOC_PARTS * sizeof(MassPart))
        return set_err(CZ_EINVAL, "mass sums (no-alloc): %llu cells exceed the static scratch's maximum of %llu cells",
                       (unsigned long long)n, (unsigned long long)(MASS_NOALLOC_PARTS * MASS_CHUNK));
    MassPart* parts = nullptr;
    rc = scratch_of_current_device(&parts);
    if (rc) return rc;
    return launch_mass(h, m, n, parts, out, (hipStream_t)stream);
}
