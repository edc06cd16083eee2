// nb_server.hip -- continuous-batching simulation server.
//
// Requests (one N-body system each, with a step count) are admitted against
// an HBM budget and keep their state resident on the device until they
// finish.  Every tick advances the Barnes-Hut requests shortest remaining
// work first within a budget of particle-steps (4e7 unless set; with a budget
// of 0 every request takes up to B steps, B = 1 unless set) and each direct
// request by K steps, K = min(the direct steps per tick, 64 unless set, the
// fewest steps any direct request has left), so that cheap requests are not
// paced by the tree steps in the same tick:
//   * direct-method requests (N below a threshold) are stepped together: per
//     step one batched kick-drift launch, one persistent force launch over all
//     of them (work items (request, target tile, shape) pulled from an atomic
//     counter), one closing-kick launch -- all on the server stream;
//   * Barnes-Hut requests are spread over a fixed pool of BH lanes, each a
//     force calculator with its own stream and workspace, so tree builds and
//     walks of different requests overlap on the device with each other and
//     with the direct batch.
// Finished requests are downloaded, their memory is released, and pending
// requests are admitted as the budget allows.  A request whose step fails
// (a launch or allocation error, or a Barnes-Hut evaluation whose device
// checks failed -- each BH evaluation ORs its error flags into the request's
// own flag word, read at the tick's synchronisation) finishes at once with
// that status; the other requests of the tick are not affected.  A request's arithmetic does
// not depend on what else is running: a direct request's force uses the same
// kernel and tile shape whatever the batch, a BH request runs its own
// deterministic pipeline.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "nb_internal.hpp"

namespace nb {

constexpr int kSrvBlock = 256;
constexpr int kDirectStepsPerTick = 64;  // default direct-request steps per tick (fewer near their end)
constexpr int kDirectStepsPerTickMax = 1024;
constexpr int kBHStepsPerTick = 1;  // default Barnes-Hut request steps per tick
constexpr double kTickBudget = 4e7;  // default Barnes-Hut particle-steps per tick, shortest first

struct SrvSys {
    const float4* pack;
    double *ax, *ay, *az;
    int n;
    float soft2;
};

struct SrvItem {
    int sys, i0, cnt, shape;
};

struct SrvKick {
    double *x, *y, *z, *vx, *vy, *vz;
    const double *ax, *ay, *az, *m;
    float4* pack;
    double G, dt;
};

struct SrvKickItem {
    int sys, i0, cnt, pad;
};

__global__ void k_srv_reset(int* c) { *c = 0; }

__global__ __launch_bounds__(kSrvBlock) void k_srv_force(const SrvItem* __restrict__ items,
                                                         int n_items, int* __restrict__ counter,
                                                         const SrvSys* __restrict__ sys) {
    __shared__ float4 tile[kSrvBlock];
    __shared__ float red[3][kSrvBlock];
    __shared__ int sitem;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) sitem = atomicAdd(counter, 1);
        __syncthreads();
        const int k = sitem;
        __syncthreads();
        if (k >= n_items) break;
        const SrvItem it = items[k];
        const SrvSys sd = sys[it.sys];
        const int T = 64 << it.shape;
        const int S = kSrvBlock / T;
        const int chunk = kSrvBlock / S;
        const int t = tid & (T - 1);
        const int u = tid / T;
        const bool valid = t < it.cnt;
        const int i = it.i0 + t;
        const float4 me = sd.pack[valid ? i : 0];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        for (int j0 = 0; j0 < sd.n; j0 += kSrvBlock) {
            tile[tid] = j0 + tid < sd.n ? sd.pack[j0 + tid] : make_float4(1e18f, 1e18f, 1e18f, 0.f);
            __syncthreads();
            const int jb = u * chunk;
#pragma unroll 8
            for (int jj = 0; jj < chunk; ++jj) {
                const float4 s = tile[jb + jj];
                const float dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, sd.soft2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                inv = r2 > 0.f ? inv : 0.f;
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
            if (u == 0)
                for (int q = 1; q < S; ++q) {
                    fx += red[0][q * T + t];
                    fy += red[1][q * T + t];
                    fz += red[2][q * T + t];
                }
            __syncthreads();
        }
        if (u == 0 && valid) {
            sd.ax[i] = (double)fx;
            sd.ay[i] = (double)fy;
            sd.az[i] = (double)fz;
        }
    }
}

// phase 0: v += a*h, x += v*dt, repack; 1: v += a*h; 2: repack only.
__global__ __launch_bounds__(kSrvBlock) void k_srv_kick(const SrvKickItem* __restrict__ items,
                                                        const SrvKick* __restrict__ sys,
                                                        int phase) {
    const SrvKickItem it = items[blockIdx.x];
    const int t = threadIdx.x;
    if (t >= it.cnt) return;
    const SrvKick s = sys[it.sys];
    const int i = it.i0 + t;
    if (phase == 2) {
        s.pack[i] = make_float4((float)s.x[i], (float)s.y[i], (float)s.z[i], (float)(s.G * s.m[i]));
        return;
    }
    const double h = s.dt / 2.0;
    const double ux = s.vx[i] + s.ax[i] * h, uy = s.vy[i] + s.ay[i] * h, uz = s.vz[i] + s.az[i] * h;
    s.vx[i] = ux;
    s.vy[i] = uy;
    s.vz[i] = uz;
    if (phase == 1) return;
    const double px = s.x[i] + ux * s.dt, py = s.y[i] + uy * s.dt, pz = s.z[i] + uz * s.dt;
    s.x[i] = px;
    s.y[i] = py;
    s.z[i] = pz;
    s.pack[i] = make_float4((float)px, (float)py, (float)pz, (float)(s.G * s.m[i]));
}

}  // namespace nb

using namespace nb;

namespace {

struct Request {
    int64_t id = 0;
    nb_request desc{};
    std::vector<double> host[7];  // m x y z vx vy vz (pending) / final x y z vx vy vz (done)
    bool bh = false;
    size_t bytes = 0;
    // device state
    nb_particles p{};
    DevBuf pack;
    nb_integrator* integ = nullptr;
    bool have_acc = false;
    int lane = -1;
    int done = 0;
    int stepped = 0;      // steps launched in the current tick
    int status = NB_OK;   // a failed request finishes early with this status
    std::string message;  // and this message (nb_server_result reports both)
};

// Device bytes a request holds while it is active: its FP64 state (10
// doubles per particle) and, for a direct request, the packed FP32 copy.  A
// Barnes-Hut request's tree workspace lives in the lanes (see nb_server).
int64_t request_bytes(int64_t n, bool bh) { return n * (bh ? 80 : 80 + 16); }

}  // namespace

struct nb_server {
    double budget = 0;
    int64_t direct_threshold = 32768;
    int leaf_cap = 8;
    int n_lanes = 0;
    hipStream_t stream = nullptr;
    std::vector<nb_force*> lanes;
    int64_t next_id = 1;
    int64_t bytes_used = 0;
    // Barnes-Hut lanes run any BH request (round robin), and a calculator keeps
    // its workspace at the largest size it has needed.  The budget therefore
    // holds n_lanes workspaces sized for the largest n and smallest theta of
    // any BH request admitted so far; that charge is never released.
    int64_t lane_n = 0;
    double lane_theta = 0.5;
    int64_t lane_ws = 0;  // bytes charged per lane
    std::deque<std::unique_ptr<Request>> pending;
    std::vector<std::unique_ptr<Request>> active;
    std::map<int64_t, std::unique_ptr<Request>> done;
    std::vector<int64_t> completed_since_poll;
    std::vector<int64_t> admitted_since_poll;
    int direct_steps = kDirectStepsPerTick;  // direct-request steps per tick
    int bh_steps = kBHStepsPerTick;          // Barnes-Hut request steps per tick
    // Barnes-Hut particle-steps per tick; > 0 selects shortest remaining work
    // first (see nb_server_tick), 0 every request each tick.
    double tick_budget = kTickBudget;
    DevBuf d_flags;  // one error-flag word per Barnes-Hut request stepped in a tick
    // Two launch slots (priming batch, stepping batch) per tick, each with its
    // own device tables and the host staging they are copied from (kept alive
    // until the tick's synchronisation).
    struct Slot {
        DevBuf d_sys, d_items, d_kick, d_kitems, d_counter;
        std::vector<SrvSys> sys;
        std::vector<SrvKick> kick;
        std::vector<SrvItem> items;
        std::vector<SrvKickItem> kitems;
    } slot[2];
    int64_t steps_total = 0, particle_steps = 0;
    int cus = 0;
};

namespace {

// Extra budget needed to grow every lane to hold a BH request of n particles
// at opening angle theta (0 if the lanes are already large enough).
int64_t lane_growth(const nb_server* s, int64_t n, double theta, int64_t* ws_out) {
    const int64_t ws = bh_workspace_bytes(std::max(n, s->lane_n), s->leaf_cap,
                                          std::min(theta, s->lane_theta), s->cus);
    if (ws_out) *ws_out = ws;
    return (int64_t)s->n_lanes * std::max<int64_t>(0, ws - s->lane_ws);
}

int admit(nb_server* s) {
    while (!s->pending.empty()) {
        Request* r = s->pending.front().get();
        int64_t ws = 0;
        const int64_t grow = r->bh ? lane_growth(s, r->desc.n, r->desc.theta, &ws) : 0;
        if ((double)(s->bytes_used + (int64_t)r->bytes + grow) > s->budget) break;  // wait for memory
        NB_TRY(nb_particles_alloc(&r->p, r->desc.n));
        const double* h[7];
        for (int k = 0; k < 7; ++k) h[k] = r->host[k].data();
        NB_TRY(nb_particles_upload(&r->p, h[1], h[2], h[3], h[4], h[5], h[6], h[0]));
        for (auto& v : r->host) std::vector<double>().swap(v);
        if (r->bh) {
            NB_TRY(nb_integrator_create(&r->integ, NB_LEAPFROG));
            s->bytes_used += grow;
            s->lane_ws = std::max(s->lane_ws, ws);
            s->lane_n = std::max(s->lane_n, r->desc.n);
            s->lane_theta = std::min(s->lane_theta, r->desc.theta);
        } else {
            NB_TRY(r->pack.reserve((size_t)r->desc.n * sizeof(float4)));
        }
        s->bytes_used += (int64_t)r->bytes;
        s->admitted_since_poll.push_back(r->id);
        s->active.push_back(std::move(s->pending.front()));
        s->pending.pop_front();
    }
    return NB_OK;
}

int shape_for(int n) { return n >= 2048 ? 2 : n >= 512 ? 1 : 0; }

// `nsteps` batched leapfrog steps (or only the priming force) of the direct
// requests; the launch tables are built once and serve every step.
int direct_step(nb_server* s, const std::vector<Request*>& reqs, bool prime_only, int nsteps = 1) {
    if (reqs.empty()) return NB_OK;
    nb_server::Slot& sl = s->slot[prime_only ? 0 : 1];
    std::vector<SrvSys>& sys = sl.sys;
    std::vector<SrvKick>& kick = sl.kick;
    std::vector<SrvItem>& items = sl.items;
    std::vector<SrvKickItem>& kitems = sl.kitems;
    sys.clear();
    kick.clear();
    items.clear();
    kitems.clear();
    for (size_t q = 0; q < reqs.size(); ++q) {
        Request* r = reqs[q];
        const int n = (int)r->desc.n;
        sys.push_back(SrvSys{r->pack.as<float4>(), r->p.ax, r->p.ay, r->p.az, n,
                             (float)(r->desc.softening * r->desc.softening)});
        kick.push_back(SrvKick{r->p.x, r->p.y, r->p.z, r->p.vx, r->p.vy, r->p.vz, r->p.ax, r->p.ay,
                               r->p.az, r->p.m, r->pack.as<float4>(), r->desc.G, r->desc.dt});
        const int sh = shape_for(n), T = 64 << sh;
        for (int i0 = 0; i0 < n; i0 += T) items.push_back(SrvItem{(int)q, i0, std::min(T, n - i0), sh});
        for (int i0 = 0; i0 < n; i0 += kSrvBlock)
            kitems.push_back(SrvKickItem{(int)q, i0, std::min(kSrvBlock, n - i0), 0});
    }
    // Largest systems first for load balance.
    std::stable_sort(items.begin(), items.end(), [&](const SrvItem& a, const SrvItem& b) {
        return (int64_t)a.cnt * sys[a.sys].n > (int64_t)b.cnt * sys[b.sys].n;
    });
    NB_TRY(sl.d_sys.reserve(sys.size() * sizeof(SrvSys)));
    NB_TRY(sl.d_kick.reserve(kick.size() * sizeof(SrvKick)));
    NB_TRY(sl.d_items.reserve(items.size() * sizeof(SrvItem)));
    NB_TRY(sl.d_kitems.reserve(kitems.size() * sizeof(SrvKickItem)));
    NB_HIP(hipMemcpyAsync(sl.d_sys.ptr, sys.data(), sys.size() * sizeof(SrvSys), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemcpyAsync(sl.d_kick.ptr, kick.data(), kick.size() * sizeof(SrvKick), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemcpyAsync(sl.d_items.ptr, items.data(), items.size() * sizeof(SrvItem), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemcpyAsync(sl.d_kitems.ptr, kitems.data(), kitems.size() * sizeof(SrvKickItem), hipMemcpyHostToDevice, s->stream));
    const int grid = std::min((int)items.size(), std::max(1, s->cus) * 8);
    auto force = [&]() -> int {
        hipLaunchKernelGGL(k_srv_reset, dim3(1), dim3(1), 0, s->stream, sl.d_counter.as<int>());
        hipLaunchKernelGGL(k_srv_force, dim3(grid), dim3(kSrvBlock), 0, s->stream,
                           sl.d_items.as<SrvItem>(), (int)items.size(), sl.d_counter.as<int>(),
                           sl.d_sys.as<SrvSys>());
        NB_LAUNCH_CHECK();
        return NB_OK;
    };
    const dim3 kg((unsigned)kitems.size());
    if (prime_only) {
        hipLaunchKernelGGL(k_srv_kick, kg, dim3(kSrvBlock), 0, s->stream, sl.d_kitems.as<SrvKickItem>(),
                           sl.d_kick.as<SrvKick>(), 2);
        return force();
    }
    for (int k = 0; k < nsteps; ++k) {
        hipLaunchKernelGGL(k_srv_kick, kg, dim3(kSrvBlock), 0, s->stream, sl.d_kitems.as<SrvKickItem>(),
                           sl.d_kick.as<SrvKick>(), 0);
        NB_TRY(force());
        hipLaunchKernelGGL(k_srv_kick, kg, dim3(kSrvBlock), 0, s->stream, sl.d_kitems.as<SrvKickItem>(),
                           sl.d_kick.as<SrvKick>(), 1);
        NB_LAUNCH_CHECK();
    }
    return NB_OK;
}

void release(nb_server* s, Request* r) {
    nb_particles_free(&r->p);
    r->pack.release();
    if (r->integ) nb_integrator_destroy(r->integ);
    r->integ = nullptr;
    s->bytes_used -= (int64_t)r->bytes;
}

}  // namespace

extern "C" {

int nb_server_create(nb_server** out, double hbm_budget_bytes, int64_t direct_threshold,
                     int bh_lanes, int leaf_capacity) {
    if (!out || !(hbm_budget_bytes > 0) || direct_threshold < 0 || bh_lanes < 1 ||
        bh_lanes > 64 || leaf_capacity < 1 || leaf_capacity > 64) {
        set_error("nb_server_create: invalid arguments");
        return NB_ERR_INVALID;
    }
    nb_server* s = new nb_server();
    s->budget = hbm_budget_bytes;
    s->direct_threshold = direct_threshold;
    s->leaf_cap = leaf_capacity;
    s->n_lanes = bh_lanes;
    int st = [&]() -> int {
        NB_HIP(hipStreamCreate(&s->stream));
        int dev = 0;
        NB_HIP(hipGetDevice(&dev));
        NB_HIP(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, dev));
        for (auto& sl : s->slot) NB_TRY(sl.d_counter.reserve(sizeof(int)));
        return NB_OK;
    }();
    if (st != NB_OK) {
        nb_server_destroy(s);
        return st;
    }
    *out = s;
    return NB_OK;
}

int nb_server_destroy(nb_server* s) {
    if (!s) return NB_OK;
    (void)hipDeviceSynchronize();
    for (auto& r : s->active) release(s, r.get());
    for (nb_force* f : s->lanes) nb_force_destroy(f);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return NB_OK;
}

int nb_server_submit(nb_server* s, const nb_request* req, int64_t* id) {
    if (!s || !req || !id || req->n <= 0 || req->steps < 0 || !req->m || !req->x || !req->y ||
        !req->z || !req->vx || !req->vy || !req->vz || !(req->dt == req->dt) ||
        !(req->theta >= 0.0) || !(req->softening >= 0.0)) {
        set_error("nb_server_submit: invalid request");
        return NB_ERR_INVALID;
    }
    auto r = std::make_unique<Request>();
    r->desc = *req;
    r->bh = req->n >= s->direct_threshold;
    r->bytes = (size_t)request_bytes(req->n, r->bh);
    // Alone on the server it still needs its state plus every lane's workspace.
    const int64_t need = (int64_t)r->bytes +
                         (r->bh ? (int64_t)s->n_lanes * bh_workspace_bytes(req->n, s->leaf_cap,
                                                                           req->theta, s->cus)
                                : 0);
    if ((double)need > s->budget) {
        set_error("nb_server_submit: request needs %lld bytes, budget is %.0f", (long long)need,
                  s->budget);
        return NB_ERR_BUDGET;
    }
    const double* src[7] = {req->m, req->x, req->y, req->z, req->vx, req->vy, req->vz};
    for (int k = 0; k < 7; ++k) r->host[k].assign(src[k], src[k] + req->n);
    r->desc.m = r->desc.x = r->desc.y = r->desc.z = r->desc.vx = r->desc.vy = r->desc.vz = nullptr;
    r->id = s->next_id++;
    *id = r->id;
    s->pending.push_back(std::move(r));
    return NB_OK;
}

namespace {
void fail_request(Request* r, int status) {
    r->status = status;
    r->message = nb_last_error();
}
}  // namespace

int nb_server_tick(nb_server* s, int* n_completed) {
    if (!s) return NB_ERR_INVALID;
    NB_TRY(admit(s));
    std::vector<Request*> direct, fresh, bh;
    for (auto& up : s->active) {
        Request* r = up.get();
        r->stepped = 0;
        if (r->done >= r->desc.steps || r->status != NB_OK) continue;
        if (r->bh) {
            bh.push_back(r);
        } else {
            direct.push_back(r);
            if (!r->have_acc) fresh.push_back(r);
        }
    }
    // Barnes-Hut steps of this tick.  Without a budget every request takes
    // up to bh_steps steps.  With one, requests go in order of remaining
    // work (particles x steps left, then id); each takes as many steps as
    // its share of the budget (budget / lanes) holds, at least one, and
    // requests are added until the budget is spent and every lane has one.
    std::vector<std::pair<Request*, int>> plan;
    if (s->tick_budget <= 0.0) {
        for (Request* r : bh) plan.emplace_back(r, std::min(s->bh_steps, r->desc.steps - r->done));
    } else {
        auto rem = [](const Request* r) { return r->desc.n * (int64_t)(r->desc.steps - r->done); };
        std::stable_sort(bh.begin(), bh.end(), [&](const Request* a, const Request* b) {
            return rem(a) < rem(b) || (rem(a) == rem(b) && a->id < b->id);
        });
        const double share = s->tick_budget / s->n_lanes;
        double spent = 0.0;
        for (Request* r : bh) {
            if (spent >= s->tick_budget && (int)plan.size() >= s->n_lanes) break;
            const int left = r->desc.steps - r->done;
            const int k = (int)std::max<double>(1.0, std::min<double>(left, std::floor(share / (double)r->desc.n)));
            plan.emplace_back(r, k);
            spent += (double)r->desc.n * k;
        }
    }
    bh.clear();
    for (const auto& pk : plan) bh.push_back(pk.first);
    // Every Barnes-Hut evaluation of request k ORs its device error flags
    // into word k, read back at the tick's synchronisation.
    if (!bh.empty()) {
        NB_TRY(s->d_flags.reserve(bh.size() * sizeof(int)));
        NB_HIP(hipMemset(s->d_flags.ptr, 0, bh.size() * sizeof(int)));
    }
    for (size_t k = 0; k < bh.size(); ++k) {
        Request* r = bh[k];
        const int lane = (int)(k % (size_t)s->n_lanes);
        while ((int)s->lanes.size() <= lane) {
            nb_force* f = nullptr;
            NB_TRY(nb_force_create_bh(&f, r->desc.G, r->desc.softening, r->desc.theta,
                                      s->leaf_cap, NB_QUADRUPOLE));
            s->lanes.push_back(f);  // its workspace is charged at admission
        }
        nb_force* f = s->lanes[lane];
        // A lane is a workspace + stream; the physics parameters are the request's.
        f->G = r->desc.G;
        f->softening = r->desc.softening;
        f->soft2 = r->desc.softening * r->desc.softening;
        f->theta = r->desc.theta;
        f->err_sink = s->d_flags.as<int>() + k;
        const int nst = plan[k].second;
        const int rc = nb_integrator_run(r->integ, f, &r->p, r->desc.dt, nst);
        f->err_sink = nullptr;
        if (rc == NB_OK) r->stepped = nst;
        else fail_request(r, rc);  // only this request fails; the tick goes on
    }
    int kdir = s->direct_steps;
    for (Request* r : direct) kdir = std::min(kdir, r->desc.steps - r->done);
    int rc = direct_step(s, fresh, true);
    if (rc == NB_OK) {
        for (Request* r : fresh) r->have_acc = true;
        rc = direct_step(s, direct, false, kdir);
    }
    for (Request* r : direct) {
        if (rc == NB_OK) r->stepped = kdir;
        else fail_request(r, rc);
    }
    NB_HIP(hipDeviceSynchronize());
    if (!bh.empty()) {
        std::vector<int> flags(bh.size());
        NB_HIP(hipMemcpy(flags.data(), s->d_flags.ptr, flags.size() * sizeof(int),
                         hipMemcpyDeviceToHost));
        for (size_t k = 0; k < bh.size(); ++k) {
            if (flags[k] == 0 || bh[k]->status != NB_OK) continue;
            set_error("Barnes-Hut request %lld: invalid evaluation (flags %d: 1=walk frontier, "
                      "2=node pool, 4=tree consistency, 8=sort look-back wait, 16=non-finite "
                      "position or mass)", (long long)bh[k]->id, flags[k]);
            fail_request(bh[k], (flags[k] & ~16) ? NB_ERR_CAPACITY : NB_ERR_INVALID);
        }
    }
    int nc = 0;
    std::vector<std::unique_ptr<Request>> still;
    for (auto& up : s->active) {
        Request* r = up.get();
        r->done += r->stepped;
        s->particle_steps += r->desc.n * r->stepped;
        if (r->done >= r->desc.steps || r->status != NB_OK) {
            for (int k = 1; k < 7; ++k) r->host[k].resize(r->desc.n);
            const int drc = nb_particles_download(&r->p, r->host[1].data(), r->host[2].data(),
                                                  r->host[3].data(), r->host[4].data(),
                                                  r->host[5].data(), r->host[6].data(), nullptr,
                                                  nullptr, nullptr, nullptr);
            if (drc != NB_OK && r->status == NB_OK) fail_request(r, drc);
            release(s, r);
            s->completed_since_poll.push_back(r->id);
            s->done[r->id] = std::move(up);
            ++nc;
        } else {
            still.push_back(std::move(up));
        }
    }
    s->active.swap(still);
    s->steps_total++;
    NB_TRY(admit(s));
    if (n_completed) *n_completed = nc;
    return NB_OK;
}

int nb_server_poll(nb_server* s, int64_t* ids, int capacity, int* count) {
    if (!s || !count || (!ids && capacity > 0)) return NB_ERR_INVALID;
    const int n = (int)std::min<size_t>((size_t)std::max(0, capacity), s->completed_since_poll.size());
    for (int k = 0; k < n; ++k) ids[k] = s->completed_since_poll[k];
    s->completed_since_poll.erase(s->completed_since_poll.begin(), s->completed_since_poll.begin() + n);
    *count = n;
    return NB_OK;
}

int nb_server_poll_admitted(nb_server* s, int64_t* ids, int capacity, int* count) {
    if (!s || !count || (!ids && capacity > 0)) return NB_ERR_INVALID;
    const int n = (int)std::min<size_t>((size_t)std::max(0, capacity), s->admitted_since_poll.size());
    for (int k = 0; k < n; ++k) ids[k] = s->admitted_since_poll[k];
    s->admitted_since_poll.erase(s->admitted_since_poll.begin(), s->admitted_since_poll.begin() + n);
    *count = n;
    return NB_OK;
}

int nb_server_result(nb_server* s, int64_t id, double* x, double* y, double* z, double* vx,
                     double* vy, double* vz) {
    if (!s) return NB_ERR_INVALID;
    auto it = s->done.find(id);
    if (it == s->done.end()) {
        set_error("nb_server_result: request %lld is not finished (or already collected)", (long long)id);
        return NB_ERR_STATE;
    }
    Request* r = it->second.get();
    double* dst[6] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) std::memcpy(dst[k], r->host[k + 1].data(), sizeof(double) * r->desc.n);
    const int status = r->status;
    if (status != NB_OK)
        set_error("nb_server_result: request %lld failed after %d of %d steps: %s", (long long)id,
                  r->done, r->desc.steps, r->message.c_str());
    s->done.erase(it);
    return status;
}

int nb_server_set_direct_steps(nb_server* s, int steps_per_tick) {
    if (!s || steps_per_tick < 1 || steps_per_tick > kDirectStepsPerTickMax) {
        set_error("nb_server_set_direct_steps: steps_per_tick=%d outside [1, %d]", steps_per_tick,
                  kDirectStepsPerTickMax);
        return NB_ERR_INVALID;
    }
    s->direct_steps = steps_per_tick;
    return NB_OK;
}

int nb_server_set_bh_steps(nb_server* s, int steps_per_tick) {
    if (!s || steps_per_tick < 1 || steps_per_tick > kDirectStepsPerTickMax) {
        set_error("nb_server_set_bh_steps: steps_per_tick=%d outside [1, %d]", steps_per_tick,
                  kDirectStepsPerTickMax);
        return NB_ERR_INVALID;
    }
    s->bh_steps = steps_per_tick;
    return NB_OK;
}

int nb_server_set_tick_budget(nb_server* s, double particle_steps) {
    if (!s || !(particle_steps >= 0.0) || !std::isfinite(particle_steps)) {
        set_error("nb_server_set_tick_budget: budget must be a finite number >= 0");
        return NB_ERR_INVALID;
    }
    s->tick_budget = particle_steps;
    return NB_OK;
}

int nb_server_get_status(const nb_server* s, nb_server_status* out) {
    if (!s || !out) return NB_ERR_INVALID;
    out->pending = (int64_t)s->pending.size();
    out->active = (int64_t)s->active.size();
    out->finished_uncollected = (int64_t)s->done.size();
    out->bytes_in_use = s->bytes_used;
    out->ticks = s->steps_total;
    out->particle_steps = s->particle_steps;
    return NB_OK;
}

}  // extern "C"
