// nb_probe.hip -- FP32 issue-rate probe on gfx950.
//
// The library is built without packed FP32 instructions, so the published
// FP32 peak (which counts v_pk_fma_f32) is not the ceiling of its kernels.
// This probe measures the ceiling that applies: 256-lane blocks, 8 per CU,
// each lane running 16 independent v_fma_f32 chains whose operands are all
// VGPRs.  The benchmark reports the force kernels' FLOP rate beside it.
#include "nb_internal.hpp"

namespace nb {

constexpr int kProbeChains = 16;
constexpr int kProbeIters = 2048;
constexpr int kProbeBlocksPerCU = 8;

__global__ __launch_bounds__(256) void k_probe_fma(float* __restrict__ out, int iters) {
    float acc[kProbeChains];
    // Per-lane multiplier and addend keep every operand in a VGPR.
    const float a = 0.999f + threadIdx.x * 1e-7f, b = 1e-3f + threadIdx.x * 1e-9f;
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) acc[k] = threadIdx.x * 1e-3f + k;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kProbeChains; ++k) acc[k] = fmaf(acc[k], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;  // keeps the chains live
}

}  // namespace nb

using namespace nb;

extern "C" int nb_probe_fp32_fma(double* tflops, double* ms) {
    if (!tflops) {
        set_error("nb_probe_fp32_fma: null pointer");
        return NB_ERR_INVALID;
    }
    int dev = 0, cus = 0;
    NB_HIP(hipGetDevice(&dev));
    NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = std::max(1, cus) * kProbeBlocksPerCU;
    float* out = nullptr;
    NB_HIP(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NB_OK;
    float t = 0.f;
    auto run = [&]() -> int {
        NB_HIP(hipEventCreate(&e0));
        NB_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(256), 0, 0, out, kProbeIters);  // warm-up
        NB_HIP(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(256), 0, 0, out, kProbeIters);
        NB_HIP(hipEventRecord(e1, 0));
        NB_LAUNCH_CHECK();
        NB_HIP(hipEventSynchronize(e1));
        NB_HIP(hipEventElapsedTime(&t, e0, e1));
        return NB_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc != NB_OK) return rc;
    const double lane_fma = (double)blocks * 256 * kProbeIters * kProbeChains;
    *tflops = 2.0 * lane_fma / (t * 1e-3) / 1e12;
    if (ms) *ms = t;
    return NB_OK;
}
