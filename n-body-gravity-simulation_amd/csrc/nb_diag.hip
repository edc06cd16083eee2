// nb_diag.hip -- energy / momentum diagnostics reduced on the GPU in FP64 with
// Neumaier-compensated sums (deterministic: fixed grid, fixed combine order).
// Quantities follow CPU/include/tests_common.hpp:33-75: KE = sum 0.5 m v.v,
// PE = -G sum_{i<j} m_i m_j / r_ij (reference: unsoftened), angular momentum
// sum m (r x v), centre of mass sum m r / sum m.  The softened potential
// -G sum m_i m_j / sqrt(r^2+eps^2) is returned too.
#include "nb_internal.hpp"

namespace nb {

constexpr int kDiagBlock = 256;
constexpr int kDiagGrid = 256;
constexpr int kNq = 11;  // ke, p3, l3, mr3, m

struct Neu {
    double s, c;
};

__device__ __forceinline__ void neu_add(Neu& a, double x) {
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ void neu_merge(Neu& a, const Neu& b) {
    neu_add(a, b.s);
    a.c += b.c;
}

__global__ __launch_bounds__(kDiagBlock) void k_diag_partial(
    int n, const double* __restrict__ m, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ vx,
    const double* __restrict__ vy, const double* __restrict__ vz, Neu* __restrict__ part) {
    __shared__ Neu sh[kDiagBlock];
    Neu acc[kNq];
    for (int q = 0; q < kNq; ++q) acc[q] = Neu{0.0, 0.0};
    for (int i = blockIdx.x * kDiagBlock + threadIdx.x; i < n; i += gridDim.x * kDiagBlock) {
        const double mi = m[i];
        const double px = x[i], py = y[i], pz = z[i], ux = vx[i], uy = vy[i], uz = vz[i];
        neu_add(acc[0], 0.5 * mi * (ux * ux + uy * uy + uz * uz));
        neu_add(acc[1], mi * ux);
        neu_add(acc[2], mi * uy);
        neu_add(acc[3], mi * uz);
        neu_add(acc[4], mi * (py * uz - pz * uy));
        neu_add(acc[5], mi * (pz * ux - px * uz));
        neu_add(acc[6], mi * (px * uy - py * ux));
        neu_add(acc[7], mi * px);
        neu_add(acc[8], mi * py);
        neu_add(acc[9], mi * pz);
        neu_add(acc[10], mi);
    }
    const int t = threadIdx.x;
    for (int q = 0; q < kNq; ++q) {
        sh[t] = acc[q];
        __syncthreads();
        for (int s = kDiagBlock / 2; s > 0; s >>= 1) {
            if (t < s) neu_merge(sh[t], sh[t + s]);
            __syncthreads();
        }
        if (t == 0) part[blockIdx.x * kNq + q] = sh[0];
        __syncthreads();
    }
}

// Per-target potential sums phi_i = sum_{j != i} m_j / r_ij (and softened),
// each Neumaier-compensated over its N-1 terms, then a compensated block
// reduction of m_i * phi_i.
__global__ __launch_bounds__(kDiagBlock) void k_diag_potential(
    int n, const double* __restrict__ m, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, double soft2,
    Neu* __restrict__ part) {
    __shared__ double sx[kDiagBlock], sy[kDiagBlock], sz[kDiagBlock], sm[kDiagBlock];
    __shared__ Neu sh[kDiagBlock];
    const int t = threadIdx.x;
    Neu pu{0.0, 0.0}, ps{0.0, 0.0};
    for (int i0 = blockIdx.x * kDiagBlock; i0 < n; i0 += gridDim.x * kDiagBlock) {
        const int i = i0 + t;
        const bool valid = i < n;
        const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
        Neu phu{0.0, 0.0}, phs{0.0, 0.0};
        for (int j0 = 0; j0 < n; j0 += kDiagBlock) {
            const int j = j0 + t;
            if (j < n) {
                sx[t] = x[j];
                sy[t] = y[j];
                sz[t] = z[j];
                sm[t] = m[j];
            }
            __syncthreads();
            const int cnt = min(kDiagBlock, n - j0);
            if (valid) {
                for (int k = 0; k < cnt; ++k) {
                    if (j0 + k == i) continue;
                    const double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    neu_add(phu, sm[k] / sqrt(r2));
                    neu_add(phs, sm[k] / sqrt(r2 + soft2));
                }
            }
            __syncthreads();
        }
        if (valid) {
            neu_add(pu, m[i] * (phu.s + phu.c));
            neu_add(ps, m[i] * (phs.s + phs.c));
        }
    }
    for (int q = 0; q < 2; ++q) {
        sh[t] = q == 0 ? pu : ps;
        __syncthreads();
        for (int s = kDiagBlock / 2; s > 0; s >>= 1) {
            if (t < s) neu_merge(sh[t], sh[t + s]);
            __syncthreads();
        }
        if (t == 0) part[blockIdx.x * 2 + q] = sh[0];
        __syncthreads();
    }
}

}  // namespace nb

using namespace nb;

extern "C" int nb_diagnostics(const nb_particles* p, double G, double softening,
                              int with_potential, nb_diag* out) {
    if (!p || !out) {
        set_error("nb_diagnostics: null argument");
        return NB_ERR_INVALID;
    }
    *out = nb_diag{};
    const int64_t n = p->n;
    if (n == 0) return NB_OK;
    if (n > (int64_t)1 << 30) {
        set_error("nb_diagnostics: n too large");
        return NB_ERR_INVALID;
    }
    DevBuf part;
    NB_TRY(part.reserve(sizeof(Neu) * kDiagGrid * (kNq + 2)));
    Neu* pq = part.as<Neu>();
    Neu* pp = pq + kDiagGrid * kNq;
    NB_HIP(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_diag_partial, dim3(kDiagGrid), dim3(kDiagBlock), 0, 0, (int)n, p->m,
                       p->x, p->y, p->z, p->vx, p->vy, p->vz, pq);
    NB_LAUNCH_CHECK();
    if (with_potential) {
        hipLaunchKernelGGL(k_diag_potential, dim3(kDiagGrid), dim3(kDiagBlock), 0, 0, (int)n,
                           p->m, p->x, p->y, p->z, softening * softening, pp);
        NB_LAUNCH_CHECK();
    }
    std::vector<Neu> h((size_t)kDiagGrid * (kNq + 2));
    NB_HIP(hipMemcpy(h.data(), pq, h.size() * sizeof(Neu), hipMemcpyDeviceToHost));
    double r[kNq + 2];
    for (int q = 0; q < kNq; ++q) {
        double s = 0.0, c = 0.0;
        for (int b = 0; b < kDiagGrid; ++b) {
            const Neu v = h[(size_t)b * kNq + q];
            const double t = s + v.s;
            c += std::fabs(s) >= std::fabs(v.s) ? (s - t) + v.s : (v.s - t) + s;
            s = t;
            c += v.c;
        }
        r[q] = s + c;
    }
    for (int q = 0; q < 2; ++q) {
        double s = 0.0, c = 0.0;
        if (with_potential) {
            for (int b = 0; b < kDiagGrid; ++b) {
                const Neu v = h[(size_t)kDiagGrid * kNq + b * 2 + q];
                const double t = s + v.s;
                c += std::fabs(s) >= std::fabs(v.s) ? (s - t) + v.s : (v.s - t) + s;
                s = t;
                c += v.c;
            }
        }
        r[kNq + q] = s + c;
    }
    out->ke = r[0];
    for (int a = 0; a < 3; ++a) {
        out->p[a] = r[1 + a];
        out->l[a] = r[4 + a];
        out->com[a] = r[10] != 0.0 ? r[7 + a] / r[10] : 0.0;
    }
    out->mass = r[10];
    out->pe = -0.5 * G * r[kNq];
    out->pe_soft = -0.5 * G * r[kNq + 1];
    return NB_OK;
}
