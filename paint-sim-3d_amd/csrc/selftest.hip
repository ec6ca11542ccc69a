// selftest.hip - device check of the f16 conversions the kernels use
// (cz_f16_selftest of include/canvaz_film.h).
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

}  // namespace

extern "C" int cz_f16_selftest(uint64_t* bad_f2h, uint64_t* bad_h2f) {
    if (!bad_f2h || !bad_h2f) return set_err(CZ_EINVAL, "f16_selftest: null argument");
    unsigned long long* d = nullptr;
    HIP_TRY(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0ull, 0ull};
    hipError_t e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
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
