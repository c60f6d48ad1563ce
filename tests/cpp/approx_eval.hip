// Test-only device evaluation of the reference-arithmetic helpers in
// glimpsw_amd/csrc/device_math.h (VRCP14PS, VRSQRT14PS and the fast-math
// division built on them), so that a GPU test can compare the device code
// with the CPU oracle input by input. Not part of the library.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <vector>

#include "device_math.h"

namespace {

const uint32_t kPairs[] = {
#include "approx14_pwl.inc"
};

__global__ void k_eval(glimp::ApproxTables T, int op, const float* a, const float* b, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r;
    if (op == 0) r = glimp::rcp14(T, a[i]);
    else if (op == 1) r = glimp::rsqrt14(T, a[i]);
    else r = glimp::fdiv(T, a[i], b[i]);
    out[i] = r;
}

}  // namespace

// op 0: rcp14(a), 1: rsqrt14(a), 2: fdiv(a, b). Host arrays of n floats
// (b only for op 2). Returns 0 on success.
extern "C" int approx_eval(int op, const float* a, const float* b, float* out, uint32_t n) {
    if (n == 0) return 0;
    uint2* d_pairs = nullptr;
    float *d_a = nullptr, *d_b = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(&d_pairs, sizeof(kPairs));
    if (e == hipSuccess) e = hipMemcpy(d_pairs, kPairs, sizeof(kPairs), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&d_a, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&d_b, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpy(d_a, a, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_b, op == 2 ? b : a, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        glimp::ApproxTables T{ d_pairs };
        hipLaunchKernelGGL(k_eval, dim3((n + 255) / 256), dim3(256), 0, 0, T, op, d_a, d_b, d_out, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, (size_t)n * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d_pairs);
    (void)hipFree(d_a);
    (void)hipFree(d_b);
    (void)hipFree(d_out);
    if (e != hipSuccess) std::fprintf(stderr, "approx_eval: %s\n", hipGetErrorString(e));
    return e == hipSuccess ? 0 : 1;
}
