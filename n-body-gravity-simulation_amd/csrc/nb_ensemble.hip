// nb_ensemble.hip -- many independent systems of uneven size advanced together.
//
// Force: one persistent launch over all systems.  Work items are
// (system, first target, target count, shape) sorted by cost, pulled by
// 256-lane workgroups from an atomic counter.  Shape s gives T = 64 << s
// targets per item and 256/T lane subsets: each subset takes a contiguous
// 1/(256/T) slice of every 256-source LDS tile and the subsets' partial sums
// are added in subset order at the end.  A target's arithmetic therefore
// depends only on its own system and shape -- the same bits alone or batched.
// The shape per size bucket comes from the autotuner.
//
// Integration: leapfrog kick-drift-kick with the acceleration cached, FP64
// state, per-system dt; the kick-drift kernel also refreshes the packed FP32
// copy (x, y, z, G*m) the force kernel reads.  A sequence of steps can be
// captured once into a HIP graph and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>

#include "../../include/nbody_amd_diag.h"
#include "nb_internal.hpp"

namespace nb {

constexpr int kEnsBlock = 256;
constexpr int kEnsBuckets = 8;  // bucket b: N in [2^(b+6), 2^(b+7)); last bucket open-ended
constexpr int kEnsBlocksPerCU = 8;

struct EnsItem {
    int sys, i0, cnt, shape;
};

__global__ void k_ens_reset(int* counter) { *counter = 0; }

__global__ __launch_bounds__(kEnsBlock) void k_ens_pack(int n, const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const double* __restrict__ m,
                                                        const int* __restrict__ sys,
                                                        const double* __restrict__ G,
                                                        float4* __restrict__ pack) {
    const int i = blockIdx.x * kEnsBlock + threadIdx.x;
    if (i < n)
        pack[i] = make_float4((float)x[i], (float)y[i], (float)z[i], (float)(G[sys[i]] * m[i]));
}

template <bool kEps0Possible>
__global__ __launch_bounds__(kEnsBlock) void k_ens_force(
    const EnsItem* __restrict__ items, int n_items, int* __restrict__ counter,
    const int* __restrict__ off, const float* __restrict__ soft2, const float4* __restrict__ pack,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ float4 tile[kEnsBlock];
    __shared__ float red[3][kEnsBlock];
    __shared__ int sitem;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) sitem = atomicAdd(counter, 1);
        __syncthreads();
        const int k = sitem;
        __syncthreads();
        if (k >= n_items) break;
        const EnsItem it = items[k];
        const int T = 64 << it.shape;
        const int S = kEnsBlock / T;
        const int chunk = kEnsBlock / S;
        const int t = tid & (T - 1);
        const int u = tid / T;
        const int base = off[it.sys];
        const int nsys = off[it.sys + 1] - base;
        const float e2 = soft2[it.sys];
        const bool valid = t < it.cnt;
        const int i = base + it.i0 + t;
        const float4 me = pack[valid ? i : base];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        for (int j0 = 0; j0 < nsys; j0 += kEnsBlock) {
            tile[tid] = j0 + tid < nsys ? pack[base + j0 + tid]
                                        : make_float4(1e18f, 1e18f, 1e18f, 0.f);
            __syncthreads();
            const int jb = u * chunk;
#pragma unroll 8
            for (int jj = 0; jj < chunk; ++jj) {
                const float4 s = tile[jb + jj];
                const float dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, e2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                if (kEps0Possible) inv = r2 > 0.f ? inv : 0.f;
                const float c = (s.w * inv) * (inv * inv);
                fx = fmaf(dx, c, fx);
                fy = fmaf(dy, c, fy);
                fz = fmaf(dz, c, fz);
            }
            __syncthreads();
        }
        if (S > 1) {
            red[0][tid] = fx;
            red[1][tid] = fy;
            red[2][tid] = fz;
            __syncthreads();
            if (u == 0) {
                for (int q = 1; q < S; ++q) {
                    fx += red[0][q * T + t];
                    fy += red[1][q * T + t];
                    fz += red[2][q * T + t];
                }
            }
            __syncthreads();
        }
        if (u == 0 && valid) {
            ax[i] = (double)fx;
            ay[i] = (double)fy;
            az[i] = (double)fz;
        }
    }
}

// Leapfrog updates with a per-system dt; `phase` 0: v += a*h, x += v*dt
// (first half of a step); 1: v += a*h (closing kick); 2: closing kick of one
// step and opening kick-drift of the next.  Drifting kernels refresh the pack.
__global__ __launch_bounds__(kEnsBlock) void k_ens_kick(
    int n, int phase, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
    const double* __restrict__ m, const int* __restrict__ sys, const double* __restrict__ dt,
    const double* __restrict__ G, float4* __restrict__ pack) {
    const int i = blockIdx.x * kEnsBlock + threadIdx.x;
    if (i >= n) return;
    const int s = sys[i];
    const double d = dt[s];
    const double h = d / 2.0;
    double ux = vx[i] + ax[i] * h, uy = vy[i] + ay[i] * h, uz = vz[i] + az[i] * h;
    if (phase == 1) {
        vx[i] = ux;
        vy[i] = uy;
        vz[i] = uz;
        return;
    }
    if (phase == 2) {
        ux = ux + ax[i] * h;
        uy = uy + ay[i] * h;
        uz = uz + az[i] * h;
    }
    const double px = x[i] + ux * d, py = y[i] + uy * d, pz = z[i] + uz * d;
    x[i] = px;
    y[i] = py;
    z[i] = pz;
    vx[i] = ux;
    vy[i] = uy;
    vz[i] = uz;
    pack[i] = make_float4((float)px, (float)py, (float)pz, (float)(G[s] * m[i]));
}

inline int bucket_of(int64_t n) {
    int b = 0;
    while (b < kEnsBuckets - 1 && n >= (int64_t)128 << b) ++b;
    return b;
}

}  // namespace nb

using namespace nb;

struct nb_ensemble {
    int S = 0;
    int64_t N = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int64_t> offsets;
    std::vector<double> G, soft, dt;
    bool eps0 = false;
    int shape[kEnsBuckets] = {0, 0, 1, 1, 2, 2, 2, 2};  // default shapes until tuned
    DevBuf d_state, d_off, d_sys, d_G, d_soft2, d_dt, d_pack, d_items, d_counter;
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr,
           *ax = nullptr, *ay = nullptr, *az = nullptr, *m = nullptr;
    int n_items = 0;
    std::vector<EnsItem> items_host;        // batched order (cost-sorted)
    std::vector<int> sys_item_begin;        // per-system ranges in per-system order
    std::vector<EnsItem> items_by_sys;
    DevBuf d_items_by_sys;
    bool have_acc = false;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
};

namespace {

int ens_build_items(nb_ensemble* e) {
    e->items_host.clear();
    e->items_by_sys.clear();
    e->sys_item_begin.assign(e->S + 1, 0);
    for (int s = 0; s < e->S; ++s) {
        const int64_t n = e->offsets[s + 1] - e->offsets[s];
        const int sh = e->shape[bucket_of(n)];
        const int T = 64 << sh;
        e->sys_item_begin[s] = (int)e->items_by_sys.size();
        for (int64_t i0 = 0; i0 < n; i0 += T)
            e->items_by_sys.push_back(EnsItem{s, (int)i0, (int)std::min<int64_t>(T, n - i0), sh});
    }
    e->sys_item_begin[e->S] = (int)e->items_by_sys.size();
    e->items_host = e->items_by_sys;
    // Longest first: cost ~ (targets) x (system size).
    std::stable_sort(e->items_host.begin(), e->items_host.end(),
                     [&](const EnsItem& a, const EnsItem& b) {
                         const int64_t ca = (int64_t)a.cnt * (e->offsets[a.sys + 1] - e->offsets[a.sys]);
                         const int64_t cb = (int64_t)b.cnt * (e->offsets[b.sys + 1] - e->offsets[b.sys]);
                         return ca > cb;
                     });
    e->n_items = (int)e->items_host.size();
    NB_TRY(e->d_items.reserve(sizeof(EnsItem) * std::max(1, e->n_items)));
    NB_TRY(e->d_items_by_sys.reserve(sizeof(EnsItem) * std::max(1, e->n_items)));
    NB_HIP(hipMemcpy(e->d_items.ptr, e->items_host.data(), sizeof(EnsItem) * e->n_items,
                     hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(e->d_items_by_sys.ptr, e->items_by_sys.data(), sizeof(EnsItem) * e->n_items,
                     hipMemcpyHostToDevice));
    if (e->graph_exec) {
        (void)hipGraphExecDestroy(e->graph_exec);
        e->graph_exec = nullptr;
    }
    if (e->graph) {
        (void)hipGraphDestroy(e->graph);
        e->graph = nullptr;
    }
    e->graph_steps = 0;
    return NB_OK;
}

int ens_force(nb_ensemble* e, const EnsItem* items, int n_items, hipStream_t s) {
    if (n_items <= 0) return NB_OK;
    hipLaunchKernelGGL(k_ens_reset, dim3(1), dim3(1), 0, s, e->d_counter.as<int>());
    const int grid = std::min(e->grid, n_items);
    if (e->eps0)
        hipLaunchKernelGGL(k_ens_force<true>, dim3(grid), dim3(kEnsBlock), 0, s, items, n_items,
                           e->d_counter.as<int>(), e->d_off.as<int>(), e->d_soft2.as<float>(),
                           e->d_pack.as<float4>(), e->ax, e->ay, e->az);
    else
        hipLaunchKernelGGL(k_ens_force<false>, dim3(grid), dim3(kEnsBlock), 0, s, items, n_items,
                           e->d_counter.as<int>(), e->d_off.as<int>(), e->d_soft2.as<float>(),
                           e->d_pack.as<float4>(), e->ax, e->ay, e->az);
    NB_LAUNCH_CHECK();
    return NB_OK;
}

int ens_kick(nb_ensemble* e, int phase, int64_t first, int64_t count, hipStream_t s) {
    if (count <= 0) return NB_OK;
    // Kernel indices are relative to `first`: offset every pointer.
    hipLaunchKernelGGL(k_ens_kick, dim3(grid_for(count, kEnsBlock)), dim3(kEnsBlock), 0, s,
                       (int)count, phase, e->x + first, e->y + first, e->z + first, e->vx + first,
                       e->vy + first, e->vz + first, e->ax + first, e->ay + first, e->az + first,
                       e->m + first, e->d_sys.as<int>() + first, e->d_dt.as<double>(),
                       e->d_G.as<double>(), e->d_pack.as<float4>() + first);
    NB_LAUNCH_CHECK();
    return NB_OK;
}

// The launch sequence of `steps` batched leapfrog steps (acceleration cached).
int ens_enqueue_steps(nb_ensemble* e, int steps, hipStream_t s) {
    for (int k = 0; k < steps; ++k) {
        NB_TRY(ens_kick(e, k == 0 ? 0 : 2, 0, e->N, s));
        NB_TRY(ens_force(e, e->d_items.as<EnsItem>(), e->n_items, s));
    }
    NB_TRY(ens_kick(e, 1, 0, e->N, s));
    return NB_OK;
}

int ens_prime(nb_ensemble* e) {
    if (e->have_acc) return NB_OK;
    hipLaunchKernelGGL(k_ens_pack, dim3(grid_for(e->N, kEnsBlock)), dim3(kEnsBlock), 0, e->stream,
                       (int)e->N, e->x, e->y, e->z, e->m, e->d_sys.as<int>(), e->d_G.as<double>(),
                       e->d_pack.as<float4>());
    NB_LAUNCH_CHECK();
    NB_TRY(ens_force(e, e->d_items.as<EnsItem>(), e->n_items, e->stream));
    e->have_acc = true;
    return NB_OK;
}

}  // namespace

extern "C" {

int nb_ensemble_create(nb_ensemble** out, int n_systems, const int64_t* offsets, const double* G,
                       const double* softening, const double* dt) {
    if (!out || n_systems <= 0 || !offsets || !G || !softening || !dt || offsets[0] != 0) {
        set_error("nb_ensemble_create: invalid arguments");
        return NB_ERR_INVALID;
    }
    for (int s = 0; s < n_systems; ++s)
        if (offsets[s + 1] <= offsets[s] || !std::isfinite(G[s]) || !std::isfinite(softening[s]) ||
            !std::isfinite(dt[s])) {
            set_error("nb_ensemble_create: system %d is empty or has invalid parameters", s);
            return NB_ERR_INVALID;
        }
    if (offsets[n_systems] > ((int64_t)1 << 30)) {
        set_error("nb_ensemble_create: too many particles");
        return NB_ERR_INVALID;
    }
    nb_ensemble* e = new nb_ensemble();
    e->S = n_systems;
    e->N = offsets[n_systems];
    e->offsets.assign(offsets, offsets + n_systems + 1);
    e->G.assign(G, G + n_systems);
    e->soft.assign(softening, softening + n_systems);
    e->dt.assign(dt, dt + n_systems);
    auto fail = [&](int st) {
        nb_ensemble_destroy(e);
        return st;
    };
    int st;
    if ((st = [&]() -> int {
             NB_HIP(hipGetDevice(&e->device));
             NB_HIP(hipStreamCreate(&e->stream));
             int cus = 0;
             NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
             e->grid = std::max(1, cus) * kEnsBlocksPerCU;
             const size_t stride = ((size_t)e->N * 8 + 255) & ~(size_t)255;
             NB_TRY(e->d_state.reserve(stride * 10));
             char* b = e->d_state.as<char>();
             double** arr[10] = {&e->x, &e->y, &e->z, &e->vx, &e->vy, &e->vz, &e->ax, &e->ay, &e->az, &e->m};
             for (int k = 0; k < 10; ++k) *arr[k] = reinterpret_cast<double*>(b + k * stride);
             NB_HIP(hipMemset(b, 0, stride * 10));
             std::vector<int> off32(e->S + 1), sys(e->N);
             std::vector<float> s2(e->S);
             for (int s = 0; s <= e->S; ++s) off32[s] = (int)e->offsets[s];
             for (int s = 0; s < e->S; ++s) {
                 for (int64_t i = e->offsets[s]; i < e->offsets[s + 1]; ++i) sys[i] = s;
                 s2[s] = (float)(e->soft[s] * e->soft[s]);
                 if (s2[s] <= 0.f) e->eps0 = true;
             }
             NB_TRY(e->d_off.reserve(sizeof(int) * (e->S + 1)));
             NB_TRY(e->d_sys.reserve(sizeof(int) * e->N));
             NB_TRY(e->d_G.reserve(sizeof(double) * e->S));
             NB_TRY(e->d_soft2.reserve(sizeof(float) * e->S));
             NB_TRY(e->d_dt.reserve(sizeof(double) * e->S));
             NB_TRY(e->d_pack.reserve(sizeof(float4) * e->N));
             NB_TRY(e->d_counter.reserve(sizeof(int)));
             NB_HIP(hipMemcpy(e->d_off.ptr, off32.data(), sizeof(int) * (e->S + 1), hipMemcpyHostToDevice));
             NB_HIP(hipMemcpy(e->d_sys.ptr, sys.data(), sizeof(int) * e->N, hipMemcpyHostToDevice));
             NB_HIP(hipMemcpy(e->d_G.ptr, e->G.data(), sizeof(double) * e->S, hipMemcpyHostToDevice));
             NB_HIP(hipMemcpy(e->d_soft2.ptr, s2.data(), sizeof(float) * e->S, hipMemcpyHostToDevice));
             NB_HIP(hipMemcpy(e->d_dt.ptr, e->dt.data(), sizeof(double) * e->S, hipMemcpyHostToDevice));
             return ens_build_items(e);
         }()) != NB_OK)
        return fail(st);
    *out = e;
    return NB_OK;
}

int nb_ensemble_destroy(nb_ensemble* e) {
    if (!e) return NB_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return NB_OK;
}

int nb_ensemble_upload(nb_ensemble* e, const double* m, const double* x, const double* y,
                       const double* z, const double* vx, const double* vy, const double* vz) {
    if (!e || !m || !x || !y || !z || !vx || !vy || !vz) {
        set_error("nb_ensemble_upload: null argument");
        return NB_ERR_INVALID;
    }
    const size_t b = (size_t)e->N * 8;
    NB_HIP(hipStreamSynchronize(e->stream));
    const double* src[7] = {m, x, y, z, vx, vy, vz};
    double* dst[7] = {e->m, e->x, e->y, e->z, e->vx, e->vy, e->vz};
    for (int k = 0; k < 7; ++k) NB_HIP(hipMemcpy(dst[k], src[k], b, hipMemcpyHostToDevice));
    e->have_acc = false;
    return NB_OK;
}

int nb_ensemble_download(nb_ensemble* e, double* x, double* y, double* z, double* vx, double* vy,
                         double* vz, double* ax, double* ay, double* az) {
    if (!e) return NB_ERR_INVALID;
    NB_HIP(hipStreamSynchronize(e->stream));
    const size_t b = (size_t)e->N * 8;
    double* dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    const double* src[9] = {e->x, e->y, e->z, e->vx, e->vy, e->vz, e->ax, e->ay, e->az};
    for (int k = 0; k < 9; ++k)
        if (dst[k]) NB_HIP(hipMemcpy(dst[k], src[k], b, hipMemcpyDeviceToHost));
    return NB_OK;
}

int nb_ensemble_run(nb_ensemble* e, int steps, int use_graph) {
    if (!e || steps < 0) return NB_ERR_INVALID;
    if (steps == 0) return NB_OK;
    NB_TRY(ens_prime(e));
    if (!use_graph) return ens_enqueue_steps(e, steps, e->stream);
    if (!e->graph_exec || e->graph_steps != steps) {
        if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        e->graph_exec = nullptr;
        e->graph = nullptr;
        NB_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        const int st = ens_enqueue_steps(e, steps, e->stream);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(e->stream, &g);
        if (st != NB_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        NB_HIP(ce);
        e->graph = g;
        NB_HIP(hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0));
        e->graph_steps = steps;
    }
    NB_HIP(hipGraphLaunch(e->graph_exec, e->stream));
    return NB_OK;
}

int nb_ensemble_run_unbatched(nb_ensemble* e, int steps) {
    if (!e || steps < 0) return NB_ERR_INVALID;
    if (steps == 0) return NB_OK;  // as nb_ensemble_run: no step, no closing kick
    NB_TRY(ens_prime(e));
    const EnsItem* items = e->d_items_by_sys.as<EnsItem>();
    for (int k = 0; k < steps; ++k) {
        for (int s = 0; s < e->S; ++s) {
            const int64_t a = e->offsets[s], n = e->offsets[s + 1] - a;
            NB_TRY(ens_kick(e, k == 0 ? 0 : 2, a, n, e->stream));
            const int i0 = e->sys_item_begin[s], i1 = e->sys_item_begin[s + 1];
            NB_TRY(ens_force(e, items + i0, i1 - i0, e->stream));
        }
    }
    for (int s = 0; s < e->S; ++s)
        NB_TRY(ens_kick(e, 1, e->offsets[s], e->offsets[s + 1] - e->offsets[s], e->stream));
    return NB_OK;
}

int nb_ensemble_tuning(const nb_ensemble* e, int32_t* codes, int n_buckets) {
    if (!e || !codes) return NB_ERR_INVALID;
    for (int b = 0; b < n_buckets && b < kEnsBuckets; ++b) codes[b] = e->shape[b];
    return NB_OK;
}

int nb_ensemble_load_tuning(nb_ensemble* e, const char* path) {
    if (!e || !path) return NB_ERR_INVALID;
    std::ifstream f(path);
    if (!f) {
        set_error("nb_ensemble_load_tuning: cannot open %s", path);
        return NB_ERR_INVALID;
    }
    int sh[kEnsBuckets];
    for (int b = 0; b < kEnsBuckets; ++b) {
        int bb = -1;
        if (!(f >> bb >> sh[b]) || bb != b || sh[b] < 0 || sh[b] > 2) {
            set_error("nb_ensemble_load_tuning: %s is malformed", path);
            return NB_ERR_INVALID;
        }
    }
    NB_HIP(hipStreamSynchronize(e->stream));
    std::copy(sh, sh + kEnsBuckets, e->shape);
    return ens_build_items(e);
}

// Mean duration of the batched force launch (k_ens_force over every system)
// from HIP events on the ensemble's stream.  The launches recompute the
// cached acceleration from the same positions with the same shapes, so the
// state keeps its bits.
int nb_ensemble_time_force(nb_ensemble* e, int reps, double* ms) {
    if (!e || reps < 1 || !ms) {
        set_error("nb_ensemble_time_force: invalid arguments");
        return NB_ERR_INVALID;
    }
    NB_TRY(ens_prime(e));
    hipEvent_t a, b;
    NB_HIP(hipEventCreate(&a));
    NB_HIP(hipEventCreate(&b));
    int st = NB_OK;
    float t = 0.f;
    if (hipEventRecord(a, e->stream) != hipSuccess) st = NB_ERR_HIP;
    for (int r = 0; r < reps && st == NB_OK; ++r)
        st = ens_force(e, e->d_items.as<EnsItem>(), e->n_items, e->stream);
    if (st == NB_OK && (hipEventRecord(b, e->stream) != hipSuccess ||
                        hipEventSynchronize(b) != hipSuccess ||
                        hipEventElapsedTime(&t, a, b) != hipSuccess)) {
        set_error("nb_ensemble_time_force: event timing failed");
        st = NB_ERR_HIP;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)t / reps;
    return st;
}

// Times the three shapes on each occupied size bucket of this ensemble (one
// batched force launch per shape over the bucket's systems) and keeps the
// fastest.  File format: one "bucket shape" pair per line.
int nb_ensemble_autotune(nb_ensemble* e, const char* path) {
    if (!e) return NB_ERR_INVALID;
    NB_TRY(ens_prime(e));
    hipEvent_t a, b;
    NB_HIP(hipEventCreate(&a));
    NB_HIP(hipEventCreate(&b));
    for (int bk = 0; bk < kEnsBuckets; ++bk) {
        float best = INFINITY;
        int best_sh = e->shape[bk];
        for (int sh = 0; sh <= 2; ++sh) {
            std::vector<EnsItem> it;
            for (int s = 0; s < e->S; ++s) {
                const int64_t n = e->offsets[s + 1] - e->offsets[s];
                if (bucket_of(n) != bk) continue;
                const int T = 64 << sh;
                for (int64_t i0 = 0; i0 < n; i0 += T)
                    it.push_back(EnsItem{s, (int)i0, (int)std::min<int64_t>(T, n - i0), sh});
            }
            if (it.empty()) break;
            DevBuf d;
            NB_TRY(d.reserve(sizeof(EnsItem) * it.size()));
            NB_HIP(hipMemcpy(d.ptr, it.data(), sizeof(EnsItem) * it.size(), hipMemcpyHostToDevice));
            float tot = 0.f;
            for (int rep = 0; rep < 4; ++rep) {
                NB_HIP(hipEventRecord(a, e->stream));
                NB_TRY(ens_force(e, d.as<EnsItem>(), (int)it.size(), e->stream));
                NB_HIP(hipEventRecord(b, e->stream));
                NB_HIP(hipEventSynchronize(b));
                float ms = 0.f;
                NB_HIP(hipEventElapsedTime(&ms, a, b));
                if (rep > 0) tot += ms;  // first launch warms up
            }
            if (tot < best) {
                best = tot;
                best_sh = sh;
            }
        }
        e->shape[bk] = best_sh;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->have_acc = false;  // the tuning launches overwrote the cached acceleration
    NB_TRY(ens_build_items(e));
    if (path) {
        std::ofstream f(path);
        for (int bk = 0; bk < kEnsBuckets; ++bk) f << bk << " " << e->shape[bk] << "\n";
        if (!f) {
            set_error("nb_ensemble_autotune: cannot write %s", path);
            return NB_ERR_INVALID;
        }
    }
    return NB_OK;
}

}  // extern "C"
