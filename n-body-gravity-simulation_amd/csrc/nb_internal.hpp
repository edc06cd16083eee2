// nb_internal.hpp -- shared definitions of the gfx950 N-body library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nbody_amd.h"

namespace nb {

// Thread-local message of the last failing call (nb_last_error()).
void set_error(const char* fmt, ...);
void clear_error();

// Evaluates a HIP call; on failure records the message and returns NB_ERR_HIP
// (or NB_ERR_NOMEM for allocation failures) from the enclosing function.
#define NB_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t nb_e_ = (call);                                                     \
        if (nb_e_ != hipSuccess) {                                                     \
            ::nb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(nb_e_), \
                            __FILE__, __LINE__);                                       \
            return nb_e_ == hipErrorOutOfMemory ? NB_ERR_NOMEM : NB_ERR_HIP;           \
        }                                                                              \
    } while (0)

// Propagates a non-zero status.
#define NB_TRY(expr)                \
    do {                            \
        int nb_s_ = (expr);         \
        if (nb_s_ != NB_OK) return nb_s_; \
    } while (0)

#define NB_LAUNCH_CHECK() NB_HIP(hipGetLastError())

// Device buffer that grows on demand (never shrinks); freed on destruction.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int reserve(size_t want);
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

// ------------------------------------------------------------------ direct sum
struct DirectWS {
    DevBuf pack;  // float4 (x, y, z, G*m) for the FP32 kernel
    int rows = 0; // targets per lane of the FP32 kernel (0: not chosen yet)
};

// Launches the direct-sum force for positions (x,y,z) and masses m, writing
// FP64 accelerations.  Asynchronous on `s`.
int direct_eval(int precision, double G, double soft2, int64_t n, const double* m,
                const double* x, const double* y, const double* z, double* ax, double* ay,
                double* az, DirectWS& ws, hipStream_t s);

// ------------------------------------------------------------------ Barnes-Hut
struct BHWorkspace;
int bh_create(BHWorkspace** w);
// Walk frontier entries per block per level for opening angle theta, and the
// scratch bytes a calculator reserves for it on a GPU with `cus` CUs.
int walk_fcap(double theta);
int64_t walk_scratch_bytes(double theta, int cus);
// Device bytes a Barnes-Hut calculator reserves for n particles (every
// workspace buffer, walk scratch included).
int64_t bh_workspace_bytes(int64_t n, int leaf_cap, double theta, int cus);
void bh_destroy(BHWorkspace* w);
int bh_eval(BHWorkspace* w, double G, double soft2, double theta, int leaf_cap, int multipole,
            int64_t n, const double* m, const double* x, const double* y, const double* z,
            double* ax, double* ay, double* az, hipStream_t s, const hipEvent_t* stage_ev);
int bh_stats(BHWorkspace* w, hipStream_t s, nb_bh_stats* out);
// Launches a one-thread kernel on `s` that ORs the error flags of the
// evaluation just enqueued into the device int *dst.
int bh_flags_to(BHWorkspace* w, int* dst, hipStream_t s);
// Node pool size for n particles at a leaf capacity (at most 2^28).
int64_t node_capacity(int64_t n, int cap);
int bh_topology(BHWorkspace* w, hipStream_t s, int64_t n, int32_t* depth, uint64_t* path,
                uint64_t* keys, uint32_t* perm);
int bh_nodes(BHWorkspace* w, hipStream_t s, int64_t capacity, int64_t* count, int32_t* level,
             uint64_t* path, int32_t* first, int32_t* last, double* mass, double* com,
             double* quad);

}  // namespace nb

// Force calculator handle.
struct nb_force {
    int kind = 0;  // 0 = direct, 1 = Barnes-Hut
    int precision = NB_DIRECT_F32;
    double G = 1.0, softening = 0.0, soft2 = 0.0, theta = 0.5;
    int leaf_cap = 8;
    int multipole = NB_QUADRUPOLE;
    int device = 0;
    hipStream_t stream = nullptr;
    nb::DirectWS direct;
    nb::BHWorkspace* bh = nullptr;
    std::string name;
    // Timing (no synchronisation; resolved when the timings are read): every
    // call records HIP events taken from a pool on the calculator's stream.
    // timing == 1: one event per Barnes-Hut stage boundary; timing == 2: only
    // the start, the start of the walk and the end (each recorded event costs
    // a few microseconds of idle GPU between kernels).  A direct sum records
    // only the start and end of its one stage.
    int timing = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pending;  // (first pool index, events recorded) per call
    size_t ev_used = 0;
    double t_stage[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double t_build_coarse = 0;  // tree-build time of timing == 2 calls
    int64_t timed_calls = 0;
    int64_t calls = 0;
    // Barnes-Hut only: when set, every evaluation ORs its error flags into
    // this device int (the server points it at the request being stepped).
    int* err_sink = nullptr;
};

namespace nb {
// Barnes-Hut stages timed separately (a direct sum is one "walk" stage):
// bbox, keys, sort, gather, build, up-sweep, walk.
constexpr int kStages = 7;
// Resolves pending timing events into f->t_stage (synchronizes f->stream).
int resolve_timing(nb_force* f);
// Leaf capacity of an NB_LEAF_AUTO Barnes-Hut evaluation of n particles.
int auto_leaf_capacity(int64_t n, int* cap);
// Force evaluation on arbitrary position arrays (used by the integrators for
// intermediate stages).
int force_eval(nb_force* f, int64_t n, const double* m, const double* x, const double* y,
               const double* z, double* ax, double* ay, double* az);
}  // namespace nb
