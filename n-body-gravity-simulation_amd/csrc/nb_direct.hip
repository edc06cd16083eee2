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
//    -ffp-contract=off).
//  * k_direct_f64_fast: FP64 with v_rsq_f64 + one Newton step and FMAs.
//  * k_direct_f32: FP32 with v_rsq_f32, 3 sub + 3 fma + rsq + 3 mul + 3 fma
//    per interaction, R targets per lane to amortise the LDS broadcast.
#include <climits>
#include <cstdlib>

#include "nb_internal.hpp"

namespace nb {

constexpr int kDirectBlock = 256;

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
            for (int k = 0; k < cnt; ++k) {
                if (t0 + k == i) continue;
                const double dx = sx[k] - xi;
                const double dy = sy[k] - yi;
                const double dz = sz[k] - zi;
                const double d2 = dx * dx + dy * dy + dz * dz + soft2;
                const double inv = 1.0 / sqrt(d2);
                const double inv3 = inv * inv * inv;
                const double s = sgm[k] * inv3;
                fx = fx + dx * s;
                fy = fy + dy * s;
                fz = fz + dz * s;
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
            hipLaunchKernelGGL(k_direct_f64_exact, dim3(grid), dim3(kDirectBlock), 0, s, n, x, y,
                               z, m, G, soft2, ax, ay, az);
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
            if (ws.rows == 0) {  // targets per lane: NB_DIRECT_ROWS overrides the default
                ws.rows = 1;
                if (const char* e = std::getenv("NB_DIRECT_ROWS")) ws.rows = std::atoi(e);
            }
            NB_TRY(ws.pack.reserve((size_t)n * sizeof(float4)));
            float4* pm = ws.pack.as<float4>();
            hipLaunchKernelGGL(k_pack_f32, dim3(grid), dim3(kDirectBlock), 0, s, n, x, y, z, m, G,
                               pm);
            const float s2 = (float)soft2;
            if (ws.rows >= 4)
                launch_f32<4>(n, pm, s2, ax, ay, az, s);
            else if (ws.rows == 2)
                launch_f32<2>(n, pm, s2, ax, ay, az, s);
            else
                launch_f32<1>(n, pm, s2, ax, ay, az, s);
            break;
        }
        default:
            set_error("direct force: unknown precision %d", precision);
            return NB_ERR_INVALID;
    }
    NB_LAUNCH_CHECK();
    return NB_OK;
}

}  // namespace nb
