// nb_probe.hip -- FP32 issue-rate probes on gfx950.
//
// The library is built without packed FP32 instructions, so the published
// FP32 peak (which counts v_pk_fma_f32) is not the ceiling of its kernels.
// nb_probe_fp32_fma measures the ceiling that applies: 256-lane blocks, 8 per
// CU, each lane running 16 independent v_fma_f32 chains whose operands are
// all VGPRs (the fastest of several launches).  nb_probe_fp32_rsq_cost measures what one v_rsq_f32 costs in
// that issue stream: the same chains with one rsqrt before every 12 FMAs
// (the mix of a softened particle-particle term), timed against FMA alone.
// The benchmark reports the force kernels' rates beside both.
#include "../../include/nbody_amd_diag.h"
#include "nb_internal.hpp"

namespace nb {

constexpr int kProbeChains = 16;
constexpr int kProbeIters = 2048;
constexpr int kProbeBlocksPerCU = 8;
constexpr int kProbeReps = 7;  // timed launches; the fastest is reported

constexpr int kProbeMixFma = 12;  // FMAs per rsqrt in the mixed chains

// kRsq = false: one FMA per chain step; true: one rsqrt and kProbeMixFma FMAs.
template <bool kRsq>
__global__ __launch_bounds__(256) void k_probe_fma(float* __restrict__ out, int iters) {
    float acc[kProbeChains];
    // Per-lane multiplier and addend keep every operand in a VGPR.
    const float a = 0.999f + threadIdx.x * 1e-7f, b = 1e-3f + threadIdx.x * 1e-9f;
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) acc[k] = 1.0f + threadIdx.x * 1e-3f + k;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kProbeChains; ++k) {
            if (kRsq) {
                float x = __builtin_amdgcn_rsqf(acc[k]);
#pragma unroll
                for (int j = 0; j < kProbeMixFma; ++j) x = fmaf(x, a, b);
                acc[k] = x;
            } else {
                acc[k] = fmaf(acc[k], a, b);
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;  // keeps the chains live
}

}  // namespace nb

using namespace nb;

// The fastest of kProbeReps timed launches of k_probe_fma<kRsq> (after a
// warm-up launch), in ms: a single sub-millisecond launch varies by several
// per cent with the clock.
template <bool kRsq>
static int probe_ms(int blocks, int iters, float* out, float* ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NB_OK;
    auto run = [&]() -> int {
        NB_HIP(hipEventCreate(&e0));
        NB_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_probe_fma<kRsq>, dim3(blocks), dim3(256), 0, 0, out, iters);
        float best = 0.f;
        for (int r = 0; r < kProbeReps; ++r) {
            NB_HIP(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(k_probe_fma<kRsq>, dim3(blocks), dim3(256), 0, 0, out, iters);
            NB_HIP(hipEventRecord(e1, 0));
            NB_LAUNCH_CHECK();
            NB_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            NB_HIP(hipEventElapsedTime(&t, e0, e1));
            if (r == 0 || t < best) best = t;
        }
        *ms = best;
        return NB_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

static int probe_blocks(int* blocks) {
    int dev = 0, cus = 0;
    NB_HIP(hipGetDevice(&dev));
    NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *blocks = std::max(1, cus) * kProbeBlocksPerCU;
    return NB_OK;
}

extern "C" int nb_probe_fp32_fma(double* tflops, double* ms) {
    if (!tflops) {
        set_error("nb_probe_fp32_fma: null pointer");
        return NB_ERR_INVALID;
    }
    int blocks = 0;
    NB_TRY(probe_blocks(&blocks));
    float* out = nullptr;
    NB_HIP(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
    float t = 0.f;
    const int rc = probe_ms<false>(blocks, kProbeIters, out, &t);
    (void)hipFree(out);
    if (rc != NB_OK) return rc;
    const double lane_fma = (double)blocks * 256 * kProbeIters * kProbeChains;
    *tflops = 2.0 * lane_fma / (t * 1e-3) / 1e12;
    if (ms) *ms = t;
    return NB_OK;
}

extern "C" int nb_probe_fp32_rsq_cost(double* fma_slots) {
    if (!fma_slots) {
        set_error("nb_probe_fp32_rsq_cost: null pointer");
        return NB_ERR_INVALID;
    }
    int blocks = 0;
    NB_TRY(probe_blocks(&blocks));
    float* out = nullptr;
    NB_HIP(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
    float tf = 0.f, tm = 0.f;
    const int iters_mix = kProbeIters / 8;  // 13 instructions per chain step
    int rc = probe_ms<false>(blocks, kProbeIters, out, &tf);
    if (rc == NB_OK) rc = probe_ms<true>(blocks, iters_mix, out, &tm);
    (void)hipFree(out);
    if (rc != NB_OK) return rc;
    // Time per FMA slot from the FMA-only run; the mixed run's time per chain
    // step minus its 12 FMA slots is the rsqrt's cost in FMA slots.
    const double per_fma = (double)tf / kProbeIters;
    const double per_step = (double)tm / iters_mix;
    *fma_slots = per_step / per_fma - kProbeMixFma;
    return NB_OK;
}
