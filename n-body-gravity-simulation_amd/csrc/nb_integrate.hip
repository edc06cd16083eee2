// nb_integrate.hip -- leapfrog (kick-drift-kick, cached acceleration), RK4,
// Euler, Euler-Cromer and velocity Verlet (cached acceleration) on FP64
// structure-of-arrays state.  The library is built with -ffp-contract=off, so
// every update rounds exactly like the reference CPU loops
// (CPU/src/integrator_leapfrog.cpp:8-21, CPU/src/integrator_rk4.cpp:18-42,
// CPU/src/integrator_euler.cpp:9-15, CPU/src/integrator_euler-cromer.cpp:9-15,
// CPU/src/integrator_velocity_verlet.cpp:5-17).
#include "nb_internal.hpp"

struct nb_integrator {
    int kind = NB_LEAPFROG;
    bool have_acc = false;           // leapfrog: acceleration of the current positions is cached
    const double* cached_x = nullptr; // state the cache belongs to
    int64_t cached_n = -1;
    nb::DevBuf rk;                   // RK4 stage buffers (15 arrays of n doubles);
                                     // velocity Verlet: the previous acceleration (3 arrays)
    std::string name;
};

namespace nb {

constexpr int kIntBlock = 256;

struct Vec3Ptr {
    double *x, *y, *z;
};
struct CVec3Ptr {
    const double *x, *y, *z;
};

// v += a*h; x += v*dt    (CPU/src/integrator_leapfrog.cpp:8-15, CUDA/src/integrator_leafrog.cu:16-26)
__global__ void k_kick_drift(int n, Vec3Ptr x, Vec3Ptr v, CVec3Ptr a, double h, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vx = v.x[i] + a.x[i] * h;
    const double vy = v.y[i] + a.y[i] * h;
    const double vz = v.z[i] + a.z[i] * h;
    x.x[i] = x.x[i] + vx * dt;
    x.y[i] = x.y[i] + vy * dt;
    x.z[i] = x.z[i] + vz * dt;
    v.x[i] = vx;
    v.y[i] = vy;
    v.z[i] = vz;
}

// Closing kick of one step followed by the opening kick-drift of the next;
// two separately rounded kicks, so identical to the step-by-step sequence.
__global__ void k_kick_kick_drift(int n, Vec3Ptr x, Vec3Ptr v, CVec3Ptr a, double h, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double vx = v.x[i] + a.x[i] * h;
    double vy = v.y[i] + a.y[i] * h;
    double vz = v.z[i] + a.z[i] * h;
    vx = vx + a.x[i] * h;
    vy = vy + a.y[i] * h;
    vz = vz + a.z[i] * h;
    x.x[i] = x.x[i] + vx * dt;
    x.y[i] = x.y[i] + vy * dt;
    x.z[i] = x.z[i] + vz * dt;
    v.x[i] = vx;
    v.y[i] = vy;
    v.z[i] = vz;
}

// v += a*h    (CPU/src/integrator_leapfrog.cpp:19-21, CUDA/src/integrator_leafrog.cu:41-43)
__global__ void k_kick(int n, Vec3Ptr v, CVec3Ptr a, double h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v.x[i] = v.x[i] + a.x[i] * h;
    v.y[i] = v.y[i] + a.y[i] * h;
    v.z[i] = v.z[i] + a.z[i] * h;
}

// RK4 stage 1 (after a1 = f(x)): xs = x + v*h, vs = v + a1*h, sv = v + 2*vs, sa = a1.
__global__ void k_rk4_s1(int n, CVec3Ptr x, CVec3Ptr v, CVec3Ptr a, double h, Vec3Ptr xs,
                         Vec3Ptr vs, Vec3Ptr sv, Vec3Ptr sa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_RK_S1(c)                       \
    {                                     \
        const double v2 = v.c[i] + a.c[i] * h; \
        xs.c[i] = x.c[i] + v.c[i] * h;    \
        vs.c[i] = v2;                     \
        sv.c[i] = v.c[i] + 2.0 * v2;      \
        sa.c[i] = a.c[i];                 \
    }
    NB_RK_S1(x) NB_RK_S1(y) NB_RK_S1(z)
#undef NB_RK_S1
}

// RK4 stage 2 (after a2 = f(x2)): xs = x + v2*h, v3 = v + a2*h, sv += 2*v3, sa += 2*a2.
__global__ void k_rk4_s2(int n, CVec3Ptr x, CVec3Ptr v, CVec3Ptr a, double h, Vec3Ptr xs,
                         Vec3Ptr vs, Vec3Ptr sv, Vec3Ptr sa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_RK_S2(c)                          \
    {                                        \
        const double v2 = vs.c[i];           \
        const double v3 = v.c[i] + a.c[i] * h; \
        xs.c[i] = x.c[i] + v2 * h;           \
        vs.c[i] = v3;                        \
        sv.c[i] = sv.c[i] + 2.0 * v3;        \
        sa.c[i] = sa.c[i] + 2.0 * a.c[i];    \
    }
    NB_RK_S2(x) NB_RK_S2(y) NB_RK_S2(z)
#undef NB_RK_S2
}

// RK4 stage 3 (after a3 = f(x3)): xs = x + v3*dt, v4 = v + a3*dt, sv += v4, sa += 2*a3.
__global__ void k_rk4_s3(int n, CVec3Ptr x, CVec3Ptr v, CVec3Ptr a, double dt, Vec3Ptr xs,
                         Vec3Ptr vs, Vec3Ptr sv, Vec3Ptr sa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_RK_S3(c)                           \
    {                                         \
        const double v3 = vs.c[i];            \
        const double v4 = v.c[i] + a.c[i] * dt; \
        xs.c[i] = x.c[i] + v3 * dt;           \
        sv.c[i] = sv.c[i] + v4;               \
        sa.c[i] = sa.c[i] + 2.0 * a.c[i];     \
    }
    NB_RK_S3(x) NB_RK_S3(y) NB_RK_S3(z)
#undef NB_RK_S3
}

// RK4 update (after a4 = f(x4)): sa += a4; x += sv*(dt/6); v += sa*(dt/6).
__global__ void k_rk4_fin(int n, Vec3Ptr x, Vec3Ptr v, CVec3Ptr a, double sixth, CVec3Ptr sv,
                          Vec3Ptr sa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_RK_F(c)                                  \
    {                                               \
        const double s = sa.c[i] + a.c[i];          \
        x.c[i] = x.c[i] + sv.c[i] * sixth;          \
        v.c[i] = v.c[i] + s * sixth;                \
    }
    NB_RK_F(x) NB_RK_F(y) NB_RK_F(z)
#undef NB_RK_F
}

// Euler: x = x + v*dt; v = v + a*dt  (CPU/src/integrator_euler.cpp:9-15,
// CUDA/src/integrator_euler.cu:15-20).
__global__ void k_euler(int n, Vec3Ptr x, Vec3Ptr v, CVec3Ptr a, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_EU(c)                          \
    {                                     \
        x.c[i] = x.c[i] + v.c[i] * dt;    \
        v.c[i] = v.c[i] + a.c[i] * dt;    \
    }
    NB_EU(x) NB_EU(y) NB_EU(z)
#undef NB_EU
}

// Euler-Cromer: v = v + a*dt; x = x + v*dt  (CPU/src/integrator_euler-cromer.cpp:9-15).
__global__ void k_euler_cromer(int n, Vec3Ptr x, Vec3Ptr v, CVec3Ptr a, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_EC(c)                                  \
    {                                             \
        const double vn = v.c[i] + a.c[i] * dt;   \
        v.c[i] = vn;                              \
        x.c[i] = x.c[i] + vn * dt;                \
    }
    NB_EC(x) NB_EC(y) NB_EC(z)
#undef NB_EC
}

// Velocity Verlet, first half (CPU/src/integrator_velocity_verlet.cpp:5-12):
// x = x + ((v*dt) + (((a*0.5)*dt)*dt)); a_old = a.
__global__ void k_vv_drift(int n, Vec3Ptr x, CVec3Ptr v, CVec3Ptr a, Vec3Ptr aold, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#define NB_VD(c)                                                        \
    {                                                                   \
        const double ac = a.c[i];                                       \
        x.c[i] = x.c[i] + ((v.c[i] * dt) + (((ac * 0.5) * dt) * dt));   \
        aold.c[i] = ac;                                                 \
    }
    NB_VD(x) NB_VD(y) NB_VD(z)
#undef NB_VD
}

// Velocity Verlet, second half (after a = f(x)): v = v + ((a_old + a)*(0.5*dt))
// (CPU/src/integrator_velocity_verlet.cpp:14-17).
__global__ void k_vv_kick(int n, Vec3Ptr v, CVec3Ptr a, CVec3Ptr aold, double hdt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v.x[i] = v.x[i] + ((aold.x[i] + a.x[i]) * hdt);
    v.y[i] = v.y[i] + ((aold.y[i] + a.y[i]) * hdt);
    v.z[i] = v.z[i] + ((aold.z[i] + a.z[i]) * hdt);
}

static Vec3Ptr pos(nb_particles* p) { return Vec3Ptr{p->x, p->y, p->z}; }
static Vec3Ptr vel(nb_particles* p) { return Vec3Ptr{p->vx, p->vy, p->vz}; }
static CVec3Ptr cacc(nb_particles* p) { return CVec3Ptr{p->ax, p->ay, p->az}; }

static int check_args(nb_integrator* it, nb_force* f, nb_particles* p, double dt) {
    if (!it || !f || !p) {
        set_error("integrator: null argument");
        return NB_ERR_INVALID;
    }
    if (!(dt == dt)) {
        set_error("integrator: dt is NaN");
        return NB_ERR_INVALID;
    }
    if (p->n > (int64_t)1 << 30) {
        set_error("integrator: n=%lld too large", (long long)p->n);
        return NB_ERR_INVALID;
    }
    return NB_OK;
}

static int ensure_acc(nb_integrator* it, nb_force* f, nb_particles* p) {
    if (!it->have_acc || it->cached_x != p->x || it->cached_n != p->n) {
        NB_TRY(nb_force_compute(f, p));
        it->have_acc = true;
        it->cached_x = p->x;
        it->cached_n = p->n;
    }
    return NB_OK;
}

static int leapfrog(nb_integrator* it, nb_force* f, nb_particles* p, double dt, int steps) {
    const int n = (int)p->n;
    if (n == 0 || steps <= 0) return NB_OK;
    const double h = dt / 2.0;
    const int grid = grid_for(n, kIntBlock);
    NB_TRY(ensure_acc(it, f, p));
    hipLaunchKernelGGL(k_kick_drift, dim3(grid), dim3(kIntBlock), 0, f->stream, n, pos(p), vel(p),
                       cacc(p), h, dt);
    NB_LAUNCH_CHECK();
    NB_TRY(nb_force_compute(f, p));
    for (int k = 1; k < steps; ++k) {
        hipLaunchKernelGGL(k_kick_kick_drift, dim3(grid), dim3(kIntBlock), 0, f->stream, n, pos(p),
                           vel(p), cacc(p), h, dt);
        NB_LAUNCH_CHECK();
        NB_TRY(nb_force_compute(f, p));
    }
    hipLaunchKernelGGL(k_kick, dim3(grid), dim3(kIntBlock), 0, f->stream, n, vel(p), cacc(p), h);
    NB_LAUNCH_CHECK();
    return NB_OK;
}

static int rk4(nb_integrator* it, nb_force* f, nb_particles* p, double dt, int steps) {
    const int64_t n = p->n;
    if (n == 0 || steps <= 0) return NB_OK;
    const size_t stride = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
    NB_TRY(it->rk.reserve(stride * 15));
    char* b = it->rk.as<char>();
    auto arr = [&](int k) { return reinterpret_cast<double*>(b + k * stride); };
    Vec3Ptr xs{arr(0), arr(1), arr(2)}, vs{arr(3), arr(4), arr(5)}, sv{arr(6), arr(7), arr(8)},
        sa{arr(9), arr(10), arr(11)}, as{arr(12), arr(13), arr(14)};
    CVec3Ptr cx{p->x, p->y, p->z}, cv{p->vx, p->vy, p->vz}, ca{p->ax, p->ay, p->az};
    CVec3Ptr cas{as.x, as.y, as.z}, csv{sv.x, sv.y, sv.z};
    const double h = dt / 2.0, sixth = dt / 6.0;
    const int grid = grid_for(n, kIntBlock);
    hipStream_t s = f->stream;
    for (int k = 0; k < steps; ++k) {
        NB_TRY(nb_force_compute(f, p));  // a1 -> p->a (the reference leaves a1 there)
        hipLaunchKernelGGL(k_rk4_s1, dim3(grid), dim3(kIntBlock), 0, s, (int)n, cx, cv, ca, h, xs,
                           vs, sv, sa);
        NB_LAUNCH_CHECK();
        NB_TRY(force_eval(f, n, p->m, xs.x, xs.y, xs.z, as.x, as.y, as.z));
        hipLaunchKernelGGL(k_rk4_s2, dim3(grid), dim3(kIntBlock), 0, s, (int)n, cx, cv, cas, h,
                           xs, vs, sv, sa);
        NB_LAUNCH_CHECK();
        NB_TRY(force_eval(f, n, p->m, xs.x, xs.y, xs.z, as.x, as.y, as.z));
        hipLaunchKernelGGL(k_rk4_s3, dim3(grid), dim3(kIntBlock), 0, s, (int)n, cx, cv, cas, dt,
                           xs, vs, sv, sa);
        NB_LAUNCH_CHECK();
        NB_TRY(force_eval(f, n, p->m, xs.x, xs.y, xs.z, as.x, as.y, as.z));
        hipLaunchKernelGGL(k_rk4_fin, dim3(grid), dim3(kIntBlock), 0, s, (int)n, pos(p), vel(p),
                           cas, sixth, csv, sa);
        NB_LAUNCH_CHECK();
    }
    it->have_acc = false;
    return NB_OK;
}

// Euler and Euler-Cromer evaluate the force at the start of every step, as
// the reference does; velocity Verlet reuses the force of the previous step's
// end (same positions, same bits as the reference's two evaluations per step).
static int step1(nb_integrator* it, nb_force* f, nb_particles* p, double dt, int steps) {
    const int64_t n = p->n;
    if (n == 0 || steps <= 0) return NB_OK;
    const int grid = grid_for(n, kIntBlock);
    hipStream_t s = f->stream;
    if (it->kind == NB_EULER || it->kind == NB_EULER_CROMER) {
        for (int k = 0; k < steps; ++k) {
            NB_TRY(nb_force_compute(f, p));
            if (it->kind == NB_EULER)
                hipLaunchKernelGGL(k_euler, dim3(grid), dim3(kIntBlock), 0, s, (int)n, pos(p),
                                   vel(p), cacc(p), dt);
            else
                hipLaunchKernelGGL(k_euler_cromer, dim3(grid), dim3(kIntBlock), 0, s, (int)n,
                                   pos(p), vel(p), cacc(p), dt);
            NB_LAUNCH_CHECK();
        }
        it->have_acc = false;
        return NB_OK;
    }
    const size_t stride = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
    NB_TRY(it->rk.reserve(stride * 3));
    char* b = it->rk.as<char>();
    Vec3Ptr old{reinterpret_cast<double*>(b), reinterpret_cast<double*>(b + stride),
                reinterpret_cast<double*>(b + 2 * stride)};
    CVec3Ptr cold{old.x, old.y, old.z};
    CVec3Ptr cv{p->vx, p->vy, p->vz};
    const double hdt = 0.5 * dt;
    NB_TRY(ensure_acc(it, f, p));
    for (int k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(k_vv_drift, dim3(grid), dim3(kIntBlock), 0, s, (int)n, pos(p), cv,
                           cacc(p), old, dt);
        NB_LAUNCH_CHECK();
        NB_TRY(nb_force_compute(f, p));
        hipLaunchKernelGGL(k_vv_kick, dim3(grid), dim3(kIntBlock), 0, s, (int)n, vel(p), cacc(p),
                           cold, hdt);
        NB_LAUNCH_CHECK();
    }
    return NB_OK;
}

}  // namespace nb

using namespace nb;

extern "C" {

int nb_integrator_create(nb_integrator** out, int kind) {
    static const char* const names[] = {"Leapfrog", "RK4", "Euler", "Euler-Cromer",
                                        "Velocity-Verlet"};
    if (!out || kind < NB_LEAPFROG || kind > NB_VELOCITY_VERLET) {
        set_error("nb_integrator_create: unknown kind %d", kind);
        return NB_ERR_INVALID;
    }
    nb_integrator* it = new nb_integrator();
    it->kind = kind;
    it->name = names[kind];
    *out = it;
    return NB_OK;
}

int nb_integrator_destroy(nb_integrator* it) {
    delete it;
    return NB_OK;
}

int nb_integrator_reset(nb_integrator* it) {
    if (!it) return NB_ERR_INVALID;
    it->have_acc = false;
    it->cached_x = nullptr;
    return NB_OK;
}

int nb_integrator_step(nb_integrator* it, nb_force* f, nb_particles* p, double dt) {
    return nb_integrator_run(it, f, p, dt, 1);
}

int nb_integrator_run(nb_integrator* it, nb_force* f, nb_particles* p, double dt, int steps) {
    NB_TRY(check_args(it, f, p, dt));
    if (steps < 0) {
        set_error("nb_integrator_run: negative step count");
        return NB_ERR_INVALID;
    }
    switch (it->kind) {
        case NB_LEAPFROG: return leapfrog(it, f, p, dt, steps);
        case NB_RK4: return rk4(it, f, p, dt, steps);
        default: return step1(it, f, p, dt, steps);
    }
}

const char* nb_integrator_name(const nb_integrator* it) { return it ? it->name.c_str() : ""; }

}  // extern "C"
