// nb_direct.hip -- all-pairs softened gravity on gfx950.
//
// Three kernels share one structure: a 256-lane workgroup (four wave64s) owns
// 256*R targets held in registers; source particles stream through LDS in
// 256-particle tiles and are read back as wave-uniform broadcasts.
//
//  * k_direct_f64_exact: bit-identical to the reference CPU direct sum
//    (CPU/src/force_calculator_naive.cpp:20-30): per target, j ascending,
//    j != i skipped, ((dx*dx+dy*dy)+dz*dz)+eps2, IEEE sqrt and division,
//    (inv*inv)*inv, (G*m_j)*inv3, no FMA (the library is built with
//    -ffp-contract=off).  Up to 20480 particles k_direct_f64_exact_wave
//    gives the same bits with one wave per target.
//  * k_direct_f64_fast: FP64 with v_rsq_f64 + one Newton step and FMAs.
//  * k_direct_f32: FP32 with v_rsq_f32, 3 sub + 3 fma + rsq + 3 mul + 3 fma
//    per interaction, R targets per lane to amortise the LDS broadcast.
#include <climits>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "nb_internal.hpp"

namespace nb {

constexpr int kDirectBlock = 256;
constexpr int kExactWaveMaxDefault = 20480;  // n up to which k_direct_f64_exact_wave runs

__global__ __launch_bounds__(kDirectBlock) void k_direct_f64_exact(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double G, double soft2,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ double sx[kDirectBlock], sy[kDirectBlock], sz[kDirectBlock], sgm[kDirectBlock];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kDirectBlock + tid;
    const bool valid = i < n;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t0 = 0; t0 < n; t0 += kDirectBlock) {
        const int j = t0 + tid;
        if (j < n) {
            sx[tid] = x[j];
            sy[tid] = y[j];
            sz[tid] = z[j];
            sgm[tid] = G * m[j];
        }
        __syncthreads();
        const int cnt = min(kDirectBlock, n - t0);
        if (valid) {
            // Branch-free (its own j adds +0.0, an identity for a sum that
            // starts at +0.0), so the chains of consecutive sources overlap.
#pragma unroll 4
            for (int k = 0; k < cnt; ++k) {
                const double dx = sx[k] - xi;
                const double dy = sy[k] - yi;
                const double dz = sz[k] - zi;
                const double d2 = dx * dx + dy * dy + dz * dz + soft2;
                const double inv = 1.0 / sqrt(d2);
                const double inv3 = inv * inv * inv;
                const double s = sgm[k] * inv3;
                const bool self = t0 + k == i;
                fx = fx + (self ? 0.0 : dx * s);
                fy = fy + (self ? 0.0 : dy * s);
                fz = fz + (self ? 0.0 : dz * s);
            }
        }
        __syncthreads();
    }
    if (valid) {
        ax[i] = fx;
        ay[i] = fy;
        az[i] = fz;
    }
}

// The same sums for small n, where one target per lane leaves most of the GPU
// idle and each lane's chain of IEEE sqrt and division sets the time: one
// wave per target.  Lane k computes the terms of sources j0 + k (the same
// arithmetic as above: ((dx*dx+dy*dy)+dz*dz)+eps2, 1/sqrt, (inv*inv)*inv,
// (G*m_j)*inv3, dx*s); lanes 0, 1, 2 then add the x, y, z terms in ascending
// j, one add instruction per term for all three.  The skipped j == i
// (and j >= n) contribute +0.0, which leaves the running sum unchanged: it
// starts at +0.0 and so is never -0.0.
constexpr int kExactWaves = 4;  // targets per workgroup

__global__ __launch_bounds__(kExactWaves * 64) void k_direct_f64_exact_wave(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double G, double soft2,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ double tt[kExactWaves][3][64];  // the chunk's terms, per component
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = min(lane, 2);  // lanes 0, 1, 2 sum the x, y, z components
    const int i = blockIdx.x * kExactWaves + w;
    if (i >= n) return;  // the whole wave: no workgroup barrier follows
    const double xi = x[i], yi = y[i], zi = z[i];
    // Term of source j (branch-free: an excluded j reads a valid index and
    // its term is replaced by +0.0).
    auto term = [&](int j, double& px, double& py, double& pz) {
        const bool use = j < n && j != i;
        const int jc = min(j, n - 1);
        const double dx = x[jc] - xi;
        const double dy = y[jc] - yi;
        const double dz = z[jc] - zi;
        const double d2 = dx * dx + dy * dy + dz * dz + soft2;
        const double inv = 1.0 / sqrt(d2);
        const double inv3 = inv * inv * inv;
        const double s = (G * m[jc]) * inv3;
        px = use ? dx * s : 0.0;
        py = use ? dy * s : 0.0;
        pz = use ? dz * s : 0.0;
    };
    double f = 0.0;
    double px, py, pz;
    term(lane, px, py, pz);
    for (int j0 = 0; j0 < n; j0 += 64) {
        tt[w][0][lane] = px;
        tt[w][1][lane] = py;
        tt[w][2][lane] = pz;
        __builtin_amdgcn_wave_barrier();
        // The next chunk's terms do not depend on the sums below, so their
        // long division and square-root chains can overlap the adds.
        double qx, qy, qz;
        term(j0 + 64 + lane, qx, qy, qz);
        // One add per term advances all three sums (lanes 0, 1, 2).
        if (j0 + 64 <= n) {
#pragma unroll
            for (int k = 0; k < 64; ++k) f = f + tt[w][c][k];
        } else {
            for (int k = 0; k < n - j0; ++k) f = f + tt[w][c][k];
        }
        __builtin_amdgcn_wave_barrier();
        px = qx;
        py = qy;
        pz = qz;
    }
    if (lane < 3) (lane == 0 ? ax : lane == 1 ? ay : az)[i] = f;
}

template <bool kEps0>
__global__ __launch_bounds__(kDirectBlock) void k_direct_f64_fast(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double G, double soft2,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ double sx[kDirectBlock], sy[kDirectBlock], sz[kDirectBlock], sgm[kDirectBlock];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kDirectBlock + tid;
    const bool valid = i < n;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t0 = 0; t0 < n; t0 += kDirectBlock) {
        const int j = t0 + tid;
        if (j < n) {
            sx[tid] = x[j];
            sy[tid] = y[j];
            sz[tid] = z[j];
            sgm[tid] = G * m[j];
        } else {  // far away, massless padding: contributes exactly zero
            sx[tid] = 1e150;
            sy[tid] = 1e150;
            sz[tid] = 1e150;
            sgm[tid] = 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kDirectBlock; ++k) {
            const double dx = sx[k] - xi;
            const double dy = sy[k] - yi;
            const double dz = sz[k] - zi;
            const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, soft2)));
            double r = __builtin_amdgcn_rsq(r2);
            // One Newton-Raphson step for 1/sqrt: r *= 1.5 - 0.5*r2*r*r.
            const double hr = 0.5 * r2 * r;
            r = r * fma(-hr, r, 1.5);
            if (kEps0) r = r2 > 0.0 ? r : 0.0;
            const double c = sgm[k] * r * (r * r);
            fx = fma(dx, c, fx);
            fy = fma(dy, c, fy);
            fz = fma(dz, c, fz);
        }
        __syncthreads();
    }
    if (valid) {
        ax[i] = fx;
        ay[i] = fy;
        az[i] = fz;
    }
}

__global__ void k_pack_f32(int n, const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, const double* __restrict__ m, double G,
                           float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4((float)x[i], (float)y[i], (float)z[i], (float)(G * m[i]));
}

template <int R, bool kEps0>
__global__ __launch_bounds__(kDirectBlock) void k_direct_f32(
    int n, const float4* __restrict__ pm, float soft2, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az) {
    __shared__ float4 tile[kDirectBlock];
    const int tid = threadIdx.x;
    const int base = blockIdx.x * (kDirectBlock * R) + tid;
    float px[R], py[R], pz[R], fx[R], fy[R], fz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * kDirectBlock;
        const float4 q = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        px[r] = q.x;
        py[r] = q.y;
        pz[r] = q.z;
        fx[r] = fy[r] = fz[r] = 0.f;
    }
    for (int t0 = 0; t0 < n; t0 += kDirectBlock) {
        const int j = t0 + tid;
        // Padding sources sit far away with zero mass and add exactly zero.
        tile[tid] = j < n ? pm[j] : make_float4(1e18f, 1e18f, 1e18f, 0.f);
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < kDirectBlock; ++k) {
            const float4 s = tile[k];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = s.x - px[r];
                const float dy = s.y - py[r];
                const float dz = s.z - pz[r];
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                if (kEps0) inv = r2 > 0.f ? inv : 0.f;
                const float c = (s.w * inv) * (inv * inv);
                fx[r] = fmaf(dx, c, fx[r]);
                fy[r] = fmaf(dy, c, fy[r]);
                fz[r] = fmaf(dz, c, fz[r]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * kDirectBlock;
        if (i < n) {
            ax[i] = (double)fx[r];
            ay[i] = (double)fy[r];
            az[i] = (double)fz[r];
        }
    }
}

template <int R>
static void launch_f32(int n, const float4* pm, float soft2, double* ax, double* ay, double* az,
                       hipStream_t s) {
    const int grid = grid_for(n, kDirectBlock * R);
    if (soft2 > 0.f)
        hipLaunchKernelGGL((k_direct_f32<R, false>), dim3(grid), dim3(kDirectBlock), 0, s, n, pm,
                           soft2, ax, ay, az);
    else
        hipLaunchKernelGGL((k_direct_f32<R, true>), dim3(grid), dim3(kDirectBlock), 0, s, n, pm,
                           soft2, ax, ay, az);
}

// Packs (x, y, z, G*m) to FP32 into `pm` and launches the FP32 tiled kernel.
static int direct_f32(int n, const double* m, const double* x, const double* y, const double* z,
                      double G, double soft2, float4* pm, int rows, double* ax, double* ay,
                      double* az, hipStream_t s) {
    const int grid = grid_for(n, kDirectBlock);
    hipLaunchKernelGGL(k_pack_f32, dim3(grid), dim3(kDirectBlock), 0, s, n, x, y, z, m, G, pm);
    const float s2 = (float)soft2;
    if (rows >= 4)
        launch_f32<4>(n, pm, s2, ax, ay, az, s);
    else if (rows == 2)
        launch_f32<2>(n, pm, s2, ax, ay, az, s);
    else
        launch_f32<1>(n, pm, s2, ax, ay, az, s);
    NB_LAUNCH_CHECK();
    return NB_OK;
}

// Targets per lane of the FP32 kernel: NB_DIRECT_ROWS overrides the default 1.
// Every value gives the same bits (each target sums its sources in the same
// order); it only changes how the work is laid out.
static int direct_rows() {
    int rows = 1;
    if (const char* e = std::getenv("NB_DIRECT_ROWS")) rows = std::atoi(e);
    return rows;
}

// Largest n for the wave-per-target exact kernel (both kernels give the same
// bits); NB_DIRECT_EXACT_WAVE_MAX overrides it (0: never).
static int exact_wave_max() {
    if (const char* e = std::getenv("NB_DIRECT_EXACT_WAVE_MAX")) return std::atoi(e);
    return kExactWaveMaxDefault;
}

int direct_eval(int precision, double G, double soft2, int64_t n64, const double* m,
                const double* x, const double* y, const double* z, double* ax, double* ay,
                double* az, DirectWS& ws, hipStream_t s) {
    if (n64 <= 0) return NB_OK;
    if (n64 > INT_MAX / 2) {
        set_error("direct force: n=%lld exceeds the supported size", (long long)n64);
        return NB_ERR_INVALID;
    }
    const int n = (int)n64;
    const int grid = grid_for(n, kDirectBlock);
    switch (precision) {
        case NB_DIRECT_F64_EXACT:
            if (n <= exact_wave_max())
                hipLaunchKernelGGL(k_direct_f64_exact_wave, dim3(grid_for(n, kExactWaves)),
                                   dim3(kExactWaves * 64), 0, s, n, x, y, z, m, G, soft2, ax, ay,
                                   az);
            else
                hipLaunchKernelGGL(k_direct_f64_exact, dim3(grid), dim3(kDirectBlock), 0, s, n, x,
                                   y, z, m, G, soft2, ax, ay, az);
            break;
        case NB_DIRECT_F64:
            if (soft2 > 0.0)
                hipLaunchKernelGGL(k_direct_f64_fast<false>, dim3(grid), dim3(kDirectBlock), 0, s,
                                   n, x, y, z, m, G, soft2, ax, ay, az);
            else
                hipLaunchKernelGGL(k_direct_f64_fast<true>, dim3(grid), dim3(kDirectBlock), 0, s,
                                   n, x, y, z, m, G, soft2, ax, ay, az);
            break;
        case NB_DIRECT_F32: {
            if (ws.rows == 0) ws.rows = direct_rows();
            NB_TRY(ws.pack.reserve((size_t)n * sizeof(float4)));
            return direct_f32(n, m, x, y, z, G, soft2, ws.pack.as<float4>(), ws.rows, ax, ay, az,
                              s);
        }
        default:
            set_error("direct force: unknown precision %d", precision);
            return NB_ERR_INVALID;
    }
    NB_LAUNCH_CHECK();
    return NB_OK;
}

// ------------------------------------------------- allocation-free variant
// The packed FP32 sources of nb_force_direct_f32_noalloc live in this
// file-scope device array (one copy per GPU, held until the process ends)
// instead of a calculator workspace, so the call allocates nothing.  Host
// code never uses the array's name: its device address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch.
constexpr int64_t kNoAllocMaxN = NB_NOALLOC_MAX_N;  // 16 MiB of float4
static_assert(kNoAllocMaxN == ((int64_t)1 << 20), "scratch size changed: update the header");
static __device__ __align__(256) float4 scratch[kNoAllocMaxN];

constexpr int kNoAllocMaxDevices = 64;

namespace {
std::mutex g_scratch_mu;
float4* g_scratch[kNoAllocMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

static int scratch_for_current_device(float4** out) {
    int dev = 0;
    NB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("direct force (no-alloc): device %d beyond the %d supported", dev,
                  kNoAllocMaxDevices);
        return NB_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        NB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) {
            set_error("direct force (no-alloc): no device address for the scratch array");
            return NB_ERR_HIP;
        }
        g_scratch[dev] = static_cast<float4*>(ptr);
    }
    *out = g_scratch[dev];
    return NB_OK;
}

}  // namespace nb

using namespace nb;

extern "C" int nb_force_direct_f32_noalloc(nb_particles* p, double G, double softening,
                                           void* stream) {
    if (!p || !std::isfinite(G) || !std::isfinite(softening) || p->n < 0) {
        set_error("nb_force_direct_f32_noalloc: invalid arguments");
        return NB_ERR_INVALID;
    }
    if (p->n > kNoAllocMaxN) {
        set_error("nb_force_direct_f32_noalloc: n=%lld exceeds the scratch capacity of %lld "
                  "particles (use nb_force_compute with an NB_DIRECT_F32 calculator)",
                  (long long)p->n, (long long)kNoAllocMaxN);
        return NB_ERR_CAPACITY;
    }
    if (p->n == 0) return NB_OK;
    if (!p->x || !p->y || !p->z || !p->m || !p->ax || !p->ay || !p->az) {
        set_error("nb_force_direct_f32_noalloc: null particle array");
        return NB_ERR_INVALID;
    }
    float4* pm = nullptr;
    NB_TRY(scratch_for_current_device(&pm));
    return direct_f32((int)p->n, p->m, p->x, p->y, p->z, G, softening * softening, pm,
                      direct_rows(), p->ax, p->ay, p->az, static_cast<hipStream_t>(stream));
}
