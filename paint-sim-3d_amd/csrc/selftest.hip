// selftest.hip - device check of the f16 conversions the kernels use
// (cz_f16_selftest of include/canvaz_film.h), and the stream probe that
// measures what HBM delivers to a kernel on this device (cz_stream_probe).
//
// The kernels pack and unpack pigment slots and substrate bands with the
// hardware conversions (v_cvt_f16_f32 / v_cvt_f32_f16) while the host
// restatement uses the software forms of czmath.h.  This pass runs both forms
// on the device over every f32 bit pattern and every f16 bit pattern (NaN
// inputs excluded) and counts the inputs whose results differ.
#include <hip/hip_runtime.h>

#include "film_internal.h"

namespace {

constexpr uint32_t kSelfGrid = 8192u, kSelfBlock = 256u;

__global__ __launch_bounds__(256) void k_f16_selftest(unsigned long long* bad) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t start = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nf = 0u, nh = 0u;
    for (uint64_t v = start; v < (1ull << 32); v += stride) {
        const float f = cz_bits2f((uint32_t)v);
        if (f != f) continue;
        if (cz_f2h(f) != cz_f2h_sw(f)) nf++;
    }
    for (uint64_t v = start; v < 65536u; v += stride) {
        const uint32_t h = (uint32_t)v;
        if (((h >> 10) & 0x1fu) == 31u && (h & 0x3ffu) != 0u) continue;  // NaN
        if (cz_f2bits(cz_h2f(h)) != cz_f2bits(cz_h2f_sw(h))) nh++;
    }
    if (nf) atomicAdd(&bad[0], (unsigned long long)nf);
    if (nh) atomicAdd(&bad[1], (unsigned long long)nh);
}

// Grid-stride streams of 16-byte words with non-temporal accesses: a copy
// (one read and one write per word) or a fill (one write).  The rates they
// reach are the measured ceiling the bandwidth-bound kernels are compared
// with, beside the datasheet's.
typedef float f32x4s __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_copy(const f32x4s* __restrict__ a, f32x4s* __restrict__ b, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ __launch_bounds__(256) void k_stream_fill(f32x4s* __restrict__ b, uint64_t n, float v) {
    const f32x4s w = {v, v, v, v};
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        __builtin_nontemporal_store(w, b + i);
}
constexpr uint32_t kStreamGrid = 32768u;  // 128 blocks of 256 threads per CU

}  // namespace

extern "C" int cz_stream_probe(void* dst, const void* src, uint64_t bytes, int mode, void* stream) {
    if (!dst || (mode == 0 && !src) || (mode != 0 && mode != 1))
        return set_err(CZ_EINVAL, "stream_probe: mode 0 (copy) needs src and dst, mode 1 (fill) dst");
    if ((bytes & 15u) != 0u || ((uintptr_t)dst & 15u) != 0u || ((uintptr_t)src & 15u) != 0u)
        return set_err(CZ_EINVAL, "stream_probe: bytes and pointers must be multiples of 16");
    const uint64_t n = bytes / 16u;
    if (n == 0) return CZ_OK;
    drop_stale_error();
    if (mode == 0)
        hipLaunchKernelGGL(k_stream_copy, dim3(kStreamGrid), dim3(256), 0, (hipStream_t)stream, (const f32x4s*)src,
                           (f32x4s*)dst, n);
    else
        hipLaunchKernelGGL(k_stream_fill, dim3(kStreamGrid), dim3(256), 0, (hipStream_t)stream, (f32x4s*)dst, n, 1.0f);
    HIP_TRY(hipGetLastError());
    return CZ_OK;
}

extern "C" int cz_f16_selftest(uint64_t* bad_f2h, uint64_t* bad_h2f) {
    if (!bad_f2h || !bad_h2f) return set_err(CZ_EINVAL, "f16_selftest: null argument");
    unsigned long long* d = nullptr;
    HIP_TRY(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0ull, 0ull};
    hipError_t e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        drop_stale_error();
        hipLaunchKernelGGL(k_f16_selftest, dim3(kSelfGrid), dim3(kSelfBlock), 0, nullptr, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return set_err(CZ_EHIP, "f16_selftest: %s", hipGetErrorString(e));
    *bad_f2h = h[0];
    *bad_h2f = h[1];
    return CZ_OK;
}
