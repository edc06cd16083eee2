// nb_runtime.hip -- device selection, particle state and the force-calculator
// handles of the C ABI.
#include <dlfcn.h>

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <utility>

#include "nb_internal.hpp"

namespace nb {

int DevBuf::reserve(size_t want) {
    if (want <= bytes && ptr) return NB_OK;
    release();
    if (want == 0) return NB_OK;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? NB_ERR_NOMEM : NB_ERR_HIP;
    }
    bytes = want;
    return NB_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

namespace {
// Which events a timed call records.
enum { kEvAll = 0, kEvCoarse = 1, kEvWalkOnly = 2 };
}  // namespace

int resolve_timing(nb_force* f) {
    if (f->ev_pending.empty()) return NB_OK;
    NB_HIP(hipStreamSynchronize(f->stream));
    for (const auto& pe : f->ev_pending) {
        const hipEvent_t* e = &f->ev_pool[pe.first];
        auto elapsed = [&](int a, int b, double& acc) -> int {
            float ms = 0.f;
            NB_HIP(hipEventElapsedTime(&ms, e[a], e[b]));
            acc += ms;
            return NB_OK;
        };
        if (pe.second == kEvAll) {
            for (int k = 0; k < kStages; ++k) NB_TRY(elapsed(k, k + 1, f->t_stage[k]));
        } else {
            if (pe.second == kEvCoarse) NB_TRY(elapsed(0, kStages - 1, f->t_build_coarse));
            NB_TRY(elapsed(kStages - 1, kStages, f->t_stage[kStages - 1]));
        }
    }
    f->timed_calls += (int64_t)f->ev_pending.size();
    f->ev_pending.clear();
    f->ev_used = 0;
    return NB_OK;
}

// ------------------------------------------------------- leaf-capacity table
// (n_min, leaf capacity) rows, n_min ascending: an NB_LEAF_AUTO calculator
// evaluating n particles uses the capacity of the last row with n_min <= n.
// The table is read from the file named by nb_bh_set_tuning_file, else from
// $NB_BH_TUNING, else from tuning/bh_leaf_capacity.txt beside the library's
// lib/ folder (written by the leaf_capacity_sweep tool); without any of them
// every n uses capacity 8, the reference's default.
namespace {
std::mutex g_tune_mu;
std::vector<std::pair<int64_t, int>> g_tune;
bool g_tune_loaded = false;
std::string g_tune_source = "built-in";

int parse_tuning(const std::string& path, std::vector<std::pair<int64_t, int>>& out) {
    std::ifstream f(path);
    if (!f) {
        set_error("leaf-capacity table: cannot open %s", path.c_str());
        return NB_ERR_INVALID;
    }
    out.clear();
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        long long nmin;
        int cap;
        if (!(ss >> nmin)) continue;  // blank or comment line
        std::string rest;
        if (!(ss >> cap) || (ss >> rest) || nmin < 0 || cap < 1 || cap > 64 ||
            (!out.empty() && nmin <= out.back().first)) {
            set_error("leaf-capacity table %s: line %d is malformed (want \"n_min cap\", n_min "
                      "ascending, 1 <= cap <= 64)", path.c_str(), ln);
            return NB_ERR_INVALID;
        }
        out.emplace_back((int64_t)nmin, cap);
    }
    if (out.empty()) {
        set_error("leaf-capacity table %s has no rows", path.c_str());
        return NB_ERR_INVALID;
    }
    return NB_OK;
}

std::string default_tuning_path() {
    if (const char* e = std::getenv("NB_BH_TUNING")) return e;
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void*>(&nb_version), &info) || !info.dli_fname) return "";
    std::string lib = info.dli_fname;  // .../lib/libnbody_amd.so
    const size_t a = lib.rfind('/');
    if (a == std::string::npos) return "";
    const std::string dir = lib.substr(0, a);
    const size_t b = dir.rfind('/');
    return (b == std::string::npos ? std::string(".") : dir.substr(0, b)) +
           "/tuning/bh_leaf_capacity.txt";
}

// With g_tune_mu held.  A missing default file means the built-in table; a
// malformed one is an error (reported by the evaluation that needed it).
int load_default_tuning() {
    std::vector<std::pair<int64_t, int>> t;
    const std::string path = default_tuning_path();
    if (!path.empty() && std::ifstream(path)) {
        NB_TRY(parse_tuning(path, t));
        g_tune = std::move(t);
        g_tune_source = path;
    } else {
        g_tune = {{0, 8}};
        g_tune_source = "built-in";
    }
    g_tune_loaded = true;
    return NB_OK;
}
}  // namespace

int auto_leaf_capacity(int64_t n, int* cap) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_loaded) NB_TRY(load_default_tuning());
    int c = g_tune.front().second;
    for (const auto& r : g_tune)
        if (r.first <= n) c = r.second;
    *cap = c;
    return NB_OK;
}

int force_eval(nb_force* f, int64_t n, const double* m, const double* x, const double* y,
               const double* z, double* ax, double* ay, double* az) {
    if (n <= 0) return NB_OK;
    hipEvent_t evs[kStages + 1] = {};  // null entries are not recorded
    const hipEvent_t* ev = nullptr;
    if (f->timing) {
        if (f->ev_pending.size() >= 256) NB_TRY(resolve_timing(f));  // bound the event pool
        const size_t need = f->ev_used + kStages + 1;
        while (f->ev_pool.size() < need) {
            hipEvent_t e;
            NB_HIP(hipEventCreate(&e));
            f->ev_pool.push_back(e);
        }
        const int mode = f->kind == 0 ? kEvWalkOnly : f->timing == 1 ? kEvAll : kEvCoarse;
        f->ev_pending.emplace_back((int)f->ev_used, mode);
        const hipEvent_t* base = &f->ev_pool[f->ev_used];
        f->ev_used = need;
        for (int k = 0; k <= kStages; ++k)
            if (mode == kEvAll || k >= kStages - 1 || (mode == kEvCoarse && k == 0)) evs[k] = base[k];
        ev = evs;
        // A direct sum has no tree stages: its whole time is the last ("walk") stage.
        NB_HIP(hipEventRecord(f->kind == 0 ? ev[kStages - 1] : ev[0], f->stream));
    }
    if (f->kind == 0) {
        NB_TRY(direct_eval(f->precision, f->G, f->soft2, n, m, x, y, z, ax, ay, az, f->direct,
                           f->stream));
        if (ev) NB_HIP(hipEventRecord(ev[kStages], f->stream));
    } else {
        int cap = f->leaf_cap;
        if (cap == NB_LEAF_AUTO) NB_TRY(auto_leaf_capacity(n, &cap));
        NB_TRY(bh_eval(f->bh, f->G, f->soft2, f->theta, cap, f->multipole, n, m, x, y, z, ax, ay,
                       az, f->stream, ev));
        if (f->err_sink) NB_TRY(bh_flags_to(f->bh, f->err_sink, f->stream));
    }
    f->calls++;
    return NB_OK;
}

}  // namespace nb

using namespace nb;

extern "C" {

const char* nb_version(void) { return "nbody_amd 0.2 (gfx950)"; }

int nb_device_count(int* count) {
    if (!count) {
        set_error("nb_device_count: null pointer");
        return NB_ERR_INVALID;
    }
    NB_HIP(hipGetDeviceCount(count));
    return NB_OK;
}

int nb_set_device(int device) {
    NB_HIP(hipSetDevice(device));
    return NB_OK;
}

int nb_synchronize(void) {
    NB_HIP(hipDeviceSynchronize());
    return NB_OK;
}

// ------------------------------------------------------------------ particles
int nb_particles_alloc(nb_particles* p, int64_t n) {
    if (!p || n < 0) {
        set_error("nb_particles_alloc: invalid arguments (n=%lld)", (long long)n);
        return NB_ERR_INVALID;
    }
    std::memset(p, 0, sizeof(*p));
    p->n = n;
    if (n == 0) return NB_OK;
    double** arrs[10] = {&p->x, &p->y, &p->z, &p->vx, &p->vy, &p->vz, &p->ax, &p->ay, &p->az, &p->m};
    // One allocation, ten 256-byte-aligned arrays.
    const size_t stride = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
    char* base = nullptr;
    hipError_t e = hipMalloc((void**)&base, stride * 10);
    if (e != hipSuccess) {
        set_error("nb_particles_alloc: hipMalloc(%zu) failed: %s", stride * 10, hipGetErrorString(e));
        p->n = 0;
        return e == hipErrorOutOfMemory ? NB_ERR_NOMEM : NB_ERR_HIP;
    }
    for (int k = 0; k < 10; ++k) *arrs[k] = reinterpret_cast<double*>(base + k * stride);
    NB_HIP(hipMemset(base, 0, stride * 10));
    return NB_OK;
}

int nb_particles_free(nb_particles* p) {
    if (!p) return NB_OK;
    if (p->x) NB_HIP(hipFree(p->x));
    std::memset(p, 0, sizeof(*p));
    return NB_OK;
}

int nb_particles_upload(nb_particles* p, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz, const double* m) {
    if (!p || (p->n > 0 && (!x || !y || !z || !m))) {
        set_error("nb_particles_upload: null argument");
        return NB_ERR_INVALID;
    }
    const size_t b = (size_t)p->n * sizeof(double);
    if (b == 0) return NB_OK;
    NB_HIP(hipDeviceSynchronize());
    NB_HIP(hipMemcpy(p->x, x, b, hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(p->y, y, b, hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(p->z, z, b, hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(p->m, m, b, hipMemcpyHostToDevice));
    if (vx && vy && vz) {
        NB_HIP(hipMemcpy(p->vx, vx, b, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(p->vy, vy, b, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(p->vz, vz, b, hipMemcpyHostToDevice));
    } else {
        NB_HIP(hipMemset(p->vx, 0, b));
        NB_HIP(hipMemset(p->vy, 0, b));
        NB_HIP(hipMemset(p->vz, 0, b));
    }
    NB_HIP(hipDeviceSynchronize());
    return NB_OK;
}

int nb_particles_download(const nb_particles* p, double* x, double* y, double* z, double* vx,
                          double* vy, double* vz, double* ax, double* ay, double* az, double* m) {
    if (!p) {
        set_error("nb_particles_download: null particles");
        return NB_ERR_INVALID;
    }
    const size_t b = (size_t)p->n * sizeof(double);
    NB_HIP(hipDeviceSynchronize());
    if (b == 0) return NB_OK;
    const double* src[10] = {p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay, p->az, p->m};
    double* dst[10] = {x, y, z, vx, vy, vz, ax, ay, az, m};
    for (int k = 0; k < 10; ++k)
        if (dst[k]) NB_HIP(hipMemcpy(dst[k], src[k], b, hipMemcpyDeviceToHost));
    return NB_OK;
}

// ------------------------------------------------------------ force handles
static int force_common_init(nb_force* f) {
    NB_HIP(hipGetDevice(&f->device));
    NB_HIP(hipStreamCreate(&f->stream));
    return NB_OK;
}

int nb_force_create_direct(nb_force** out, int precision, double G, double softening) {
    if (!out || precision < NB_DIRECT_F64_EXACT || precision > NB_DIRECT_F32 || !std::isfinite(G) ||
        !std::isfinite(softening)) {
        set_error("nb_force_create_direct: invalid arguments");
        return NB_ERR_INVALID;
    }
    nb_force* f = new nb_force();
    f->kind = 0;
    f->precision = precision;
    f->G = G;
    f->softening = softening;
    f->soft2 = softening * softening;
    f->name = precision == NB_DIRECT_F64_EXACT ? "Naive-F64-exact"
              : precision == NB_DIRECT_F64     ? "Naive-F64"
                                               : "Naive-F32-tiled";
    int st = force_common_init(f);
    if (st != NB_OK) {
        nb_force_destroy(f);
        return st;
    }
    *out = f;
    return NB_OK;
}

int nb_force_create_bh(nb_force** out, double G, double softening, double theta, int leaf_capacity,
                       int multipole) {
    if (!out || !(theta >= 0.0) || leaf_capacity < NB_LEAF_AUTO || leaf_capacity > 64 ||
        (multipole != NB_MONOPOLE && multipole != NB_QUADRUPOLE) || !std::isfinite(G) ||
        !std::isfinite(softening)) {
        set_error("nb_force_create_bh: invalid arguments (theta=%g leaf_capacity=%d multipole=%d)",
                  theta, leaf_capacity, multipole);
        return NB_ERR_INVALID;
    }
    nb_force* f = new nb_force();
    f->kind = 1;
    f->G = G;
    f->softening = softening;
    f->soft2 = softening * softening;
    f->theta = theta;
    f->leaf_cap = leaf_capacity;
    f->multipole = multipole;
    f->name = multipole == NB_QUADRUPOLE ? "Barnes-Hut-quadrupole" : "Barnes-Hut-monopole";
    int st = force_common_init(f);
    if (st == NB_OK) st = bh_create(&f->bh);
    if (st != NB_OK) {
        nb_force_destroy(f);
        return st;
    }
    *out = f;
    return NB_OK;
}

int nb_bh_set_tuning_file(const char* path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!path) {
        g_tune_loaded = false;  // the default table is read again on next use
        return NB_OK;
    }
    std::vector<std::pair<int64_t, int>> t;
    NB_TRY(parse_tuning(path, t));
    g_tune = std::move(t);
    g_tune_source = path;
    g_tune_loaded = true;
    return NB_OK;
}

int nb_bh_auto_leaf_capacity(int64_t n, int* leaf_capacity) {
    if (!leaf_capacity || n < 0) {
        set_error("nb_bh_auto_leaf_capacity: invalid arguments");
        return NB_ERR_INVALID;
    }
    return auto_leaf_capacity(n, leaf_capacity);
}

int nb_force_destroy(nb_force* f) {
    if (!f) return NB_OK;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->bh) bh_destroy(f->bh);
    for (auto& e : f->ev_pool) (void)hipEventDestroy(e);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return NB_OK;
}

int nb_force_compute(nb_force* f, nb_particles* p) {
    if (!f || !p) {
        set_error("nb_force_compute: null argument");
        return NB_ERR_INVALID;
    }
    return force_eval(f, p->n, p->m, p->x, p->y, p->z, p->ax, p->ay, p->az);
}

const char* nb_force_name(const nb_force* f) { return f ? f->name.c_str() : ""; }

int nb_force_set_timing(nb_force* f, int enabled) {
    if (!f || enabled < 0 || enabled > 2) {
        set_error("nb_force_set_timing: mode must be 0, 1 or 2");
        return NB_ERR_INVALID;
    }
    if (!enabled) NB_TRY(resolve_timing(f));
    f->timing = enabled;
    return NB_OK;
}

int nb_force_timing(nb_force* f, double out[4]) {
    if (!f || !out) return NB_ERR_INVALID;
    NB_TRY(resolve_timing(f));
    double build = f->t_build_coarse;
    for (int k = 0; k < kStages - 1; ++k) build += f->t_stage[k];
    out[0] = build + f->t_stage[kStages - 1];
    out[1] = build;
    out[2] = f->t_stage[kStages - 1];
    out[3] = (double)f->timed_calls;
    return NB_OK;
}

int nb_force_stage_timing(nb_force* f, double* out, int n_out) {
    if (!f || !out || n_out < 1) return NB_ERR_INVALID;
    NB_TRY(resolve_timing(f));
    for (int k = 0; k < n_out && k < kStages; ++k) out[k] = f->t_stage[k];
    if (n_out > kStages) out[kStages] = (double)f->timed_calls;
    return NB_OK;
}

int nb_force_reset_timing(nb_force* f) {
    if (!f) return NB_ERR_INVALID;
    NB_TRY(resolve_timing(f));
    for (double& t : f->t_stage) t = 0;
    f->t_build_coarse = 0;
    f->timed_calls = 0;
    return NB_OK;
}

int nb_force_bh_stats(nb_force* f, nb_bh_stats* out) {
    if (!f || !out || f->kind != 1) {
        set_error("nb_force_bh_stats: not a Barnes-Hut calculator");
        return NB_ERR_INVALID;
    }
    return bh_stats(f->bh, f->stream, out);
}

int nb_force_bh_topology(nb_force* f, int64_t n, int32_t* leaf_depth, uint64_t* leaf_path,
                         uint64_t* keys, uint32_t* perm) {
    if (!f || f->kind != 1) {
        set_error("nb_force_bh_topology: not a Barnes-Hut calculator");
        return NB_ERR_INVALID;
    }
    return bh_topology(f->bh, f->stream, n, leaf_depth, leaf_path, keys, perm);
}

int nb_force_bh_nodes(nb_force* f, int64_t capacity, int64_t* count, int32_t* level,
                      uint64_t* path, int32_t* first, int32_t* last, double* mass, double* com,
                      double* quad) {
    if (!f || f->kind != 1 || !count) {
        set_error("nb_force_bh_nodes: not a Barnes-Hut calculator");
        return NB_ERR_INVALID;
    }
    return bh_nodes(f->bh, f->stream, capacity, count, level, path, first, last, mass, com, quad);
}

}  // extern "C"
