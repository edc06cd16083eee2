/*
 * nbody_amd.h -- C ABI of the MI355X (gfx950) N-body force and integration
 * engine.  Plain pointers and sizes only; every call returns a status code
 * (NB_OK == 0) and never exits the process.  The message of the last failing
 * call on the calling thread is returned by nb_last_error().
 *
 * Ownership: device buffers inside nb_particles are owned by the library
 * (allocated by nb_particles_alloc); host arrays passed to upload/download are
 * owned by the caller.  Force calculators, integrators, ensembles and servers
 * own their device workspaces.
 *
 * Ordering: compute calls are asynchronous and ordered on the calculator's
 * HIP stream; nb_particles_download() and nb_synchronize() wait for all work
 * on the device.  One calculator/integrator must not be used from two host
 * threads at once.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository).
 */
#ifndef NBODY_AMD_H
#define NBODY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum nb_status {
    NB_OK = 0,
    NB_ERR_INVALID = 1,  /* bad argument or shape */
    NB_ERR_HIP = 2,      /* HIP runtime / launch error */
    NB_ERR_NOMEM = 3,    /* device allocation failed */
    NB_ERR_CAPACITY = 4, /* tree depth / walk frontier capacity exceeded (reported, never silent) */
    NB_ERR_BUDGET = 5,   /* HBM budget exceeded (server admission) */
    NB_ERR_STATE = 6     /* call not valid in the object's current state */
};

/* Message of the last failing call on this thread ("" if none).
 * Replaces the reference's exit()/assert() error paths
 * (CUDA/src/force_calculator_naive.cu:86-90, CPU/src/force_calculator_barnes_hut_optimized.cpp:189). */
const char* nb_last_error(void);

/* Number of visible devices and selection of the current one. */
int nb_device_count(int* count);
int nb_set_device(int device);
/* Waits for all outstanding work on the current device. */
int nb_synchronize(void);
/* Library version string.  0.2: nb_bh_stats gained far_cells and
 * far_target_terms (callers built against 0.1's struct must be rebuilt). */
const char* nb_version(void);

/* ---------------------------------------------------------- particle state */
/* Structure-of-arrays device state, FP64 throughout.  Replaces ParticleArrays
 * (CUDA/include/particle_data.hpp:6-43); the reference's Morton/sorted scratch
 * arrays live in the Barnes-Hut calculator's workspace instead. */
typedef struct nb_particles {
    int64_t n;
    double *x, *y, *z;    /* positions */
    double *vx, *vy, *vz; /* velocities */
    double *ax, *ay, *az; /* accelerations written by nb_force_compute */
    double *m;            /* masses */
} nb_particles;

/* allocateParticles (CUDA/include/particle_data.hpp:45-47, CUDA/src/particle_data.cu:6-43).
 * Accelerations are zeroed. */
int nb_particles_alloc(nb_particles* p, int64_t n);
/* freeParticles (CUDA/include/particle_data.hpp:49-50). */
int nb_particles_free(nb_particles* p);
/* copyParticlesToDevice (CUDA/include/particle_data.hpp:52-61, CUDA/src/particle_data.cu:95-119).
 * Velocity pointers may be NULL (zero velocities). */
int nb_particles_upload(nb_particles* p, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz, const double* m);
/* copyParticlesToHost (CUDA/include/particle_data.hpp:63-73, CUDA/src/particle_data.cu:122-143).
 * Any pointer may be NULL to skip that array.  Synchronizes the device first. */
int nb_particles_download(const nb_particles* p, double* x, double* y, double* z, double* vx,
                          double* vy, double* vz, double* ax, double* ay, double* az, double* m);

/* -------------------------------------------------------- force calculators */
typedef struct nb_force nb_force;

enum nb_direct_precision {
    /* FP64, bit-identical to CPU/src/force_calculator_naive.cpp:20-30
     * (IEEE sqrt and division, no FMA contraction, j ascending, j != i). */
    NB_DIRECT_F64_EXACT = 0,
    /* FP64 with rsqrt + one Newton step and FMA (fast FP64 reference). */
    NB_DIRECT_F64 = 1,
    /* FP32 LDS-tiled wave64 kernel with v_rsq_f32; FP64 results. */
    NB_DIRECT_F32 = 2
};

/* ForceCalculatorNaive(double softening) (CPU/include/force_calculator_naive.hpp:12,
 * CUDA/include/force_calculator_naive.hpp:13).  G is explicit here; the reference
 * fixes it at 4*pi^2 (CPU/include/common.hpp:28).  softening is the length eps
 * (squared internally, as CPU/include/force_calculator.hpp:19). */
int nb_force_create_direct(nb_force** f, int precision, double G, double softening);

enum nb_multipole { NB_MONOPOLE = 0, NB_QUADRUPOLE = 1 };

/* BarnesHutForceCalculator(double softening, double theta, int leaf_capacity)
 * (CUDA/include/force_calculator_barnes_hut.hpp:56) /
 * ForceCalculatorBarnesHutOptimized(double, double, int)
 * (CPU/include/force_calculator_barnes_hut_optimized.hpp:40).
 * Tree topology follows the CPU reference (root cube centred on the bbox
 * midpoint, octant = x>=c | y>=c<<1 | z>=c<<2, split when count > leaf_capacity).
 * leaf_capacity = NB_LEAF_AUTO picks the capacity per evaluation from the
 * tuning table (nb_bh_set_tuning_file), the way the reference's leaf-capacity
 * sweep recommends one capacity per N (CPU/tests/test3_computeLeafCapacity.cpp,
 * CPU/documentation.md:406-418). */
#define NB_LEAF_AUTO 0
int nb_force_create_bh(nb_force** f, double G, double softening, double theta, int leaf_capacity,
                       int multipole);
/* Leaf-capacity table of NB_LEAF_AUTO calculators (process-wide), as written
 * by the leaf_capacity_sweep tool (--write-tuning): text lines
 * "n_min leaf_capacity", n_min ascending, '#' starts a comment.  An
 * evaluation of n particles uses the capacity of the last line with
 * n_min <= n.  By default the table is read on first use from $NB_BH_TUNING,
 * else from tuning/bh_leaf_capacity.txt beside the library's lib/ folder;
 * without either file every n uses capacity 8 (the reference's default).
 * path = NULL returns to that default.  Applies to the evaluations that
 * follow. */
int nb_bh_set_tuning_file(const char* path);
/* Leaf capacity an NB_LEAF_AUTO calculator uses for n particles. */
int nb_bh_auto_leaf_capacity(int64_t n, int* leaf_capacity);
/* Node pool a Barnes-Hut calculator reserves for n particles at this leaf
 * capacity: an upper bound on the tree's nodes, never more than 2^28 (the
 * walk's node-record addressing).  A tree that needs more nodes is reported
 * by nb_force_bh_stats (flag 2), never walked.  Host arithmetic only. */
int64_t nb_bh_node_capacity(int64_t n, int leaf_capacity);
int nb_force_destroy(nb_force* f);

/* ForceCalculator::computeAccelerations(ParticleArrays&)
 * (CUDA/include/force_calculator.hpp:25-27; CPU form CPU/include/force_calculator.hpp:24-26).
 * Writes p->ax, p->ay, p->az.  Asynchronous: for Barnes-Hut, the input
 * (finite positions and masses), capacity and tree-consistency checks run on
 * the device, and nb_force_bh_stats() reports their outcome, including the
 * failures of every evaluation since its last call.  An evaluation that failed
 * one is invalid and its accelerations are NaN (after a walk-frontier
 * overflow, those of the targets whose walk group overflowed), so a failure
 * also shows in every later state that depends on it. */
int nb_force_compute(nb_force* f, nb_particles* p);

/* No counterpart in the reference: the NB_DIRECT_F32 force of
 * nb_force_compute for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It allocates nothing: the packed FP32 copy of the sources lives in a
 * library-owned static device array of NB_NOALLOC_MAX_N float4 (16 MiB per
 * GPU, held until the process ends), so n > NB_NOALLOC_MAX_N is rejected with
 * NB_ERR_CAPACITY.  Writes p->ax, p->ay, p->az with the same bits as an
 * NB_DIRECT_F32 calculator.  Asynchronous on `stream` (a hipStream_t; NULL =
 * the null stream).  All calls on one GPU share its array: calls that
 * overlap on one GPU (different streams or host threads) race. */
#define NB_NOALLOC_MAX_N (1 << 20)
int nb_force_direct_f32_noalloc(nb_particles* p, double G, double softening, void* stream);

/* ForceCalculator::name() (CPU/include/force_calculator.hpp:28). */
const char* nb_force_name(const nb_force* f);

/* Timing of the calls made while timing was enabled, in ms measured with HIP
 * events on the calculator's stream (getTimeStatistics,
 * CPU/include/force_calculator.hpp:30; CUDA BarnesHutForceCalculator getters,
 * CUDA/include/force_calculator_barnes_hut.hpp:120-125).  Recording adds no
 * synchronisation; reading the timings synchronizes the stream.
 * nb_force_set_timing: 0 = off, 1 = every Barnes-Hut stage, 2 = only tree
 * build and walk (fewer events: each recorded event leaves the GPU idle for a
 * few microseconds between kernels).
 * nb_force_timing: out[0]=total, out[1]=tree build (bbox..up-sweep),
 * out[2]=walk (the whole direct sum for direct calculators), out[3]=calls.
 * nb_force_stage_timing: out[0..6] = bbox, keys, sort, gather, build,
 * up-sweep, walk; out[7] = calls (if n_out > 7).  Calls timed in mode 2
 * contribute to the walk entry only. */
int nb_force_set_timing(nb_force* f, int enabled);
int nb_force_timing(nb_force* f, double out[4]);
int nb_force_stage_timing(nb_force* f, double* out, int n_out);
int nb_force_reset_timing(nb_force* f);

/* Statistics of the last Barnes-Hut evaluation (synchronizes). */
typedef struct nb_bh_stats {
    int64_t n_nodes;        /* non-empty tree nodes */
    int64_t n_leaves;       /* non-empty leaves */
    int32_t max_level;      /* deepest occupied level */
    int32_t level_limit;    /* deepest level the 63-bit keys / 1e-9 split rule allow */
    int64_t oversize_cells; /* leaves at the key-depth limit holding > leaf_capacity
                               (the CPU reference would keep splitting them) */
    int64_t pp_interactions;   /* particle-particle interactions in the walk */
    int64_t pc_interactions;   /* particle-cell interactions in the walk */
    int32_t max_frontier;      /* largest walk frontier seen */
    int32_t frontier_overflow; /* error flags, nonzero = result invalid: 1 walk frontier full,
                                  2 node pool full, 4 malformed tree, 8 sort look-back wait limit,
                                  16 non-finite position or mass.  With 2, 4, 8 or 16 the tree is
                                  not walked and the accelerations are set to NaN. */
    double root_center[3];
    double root_size;
    int64_t n_groups;          /* walk groups (tree cells of <= 64 particles) */
    int32_t flags_since_read;  /* OR of the error flags of every evaluation since the previous
                                  nb_force_bh_stats call (the last one's included); reading
                                  clears it */
    int32_t leaf_capacity;     /* leaf capacity of the last evaluation */
    int64_t far_cells;         /* (walk group, cell) pairs whose field entered the group's local
                                  expansion instead of the per-target sum (quadrupole walk) */
    int64_t far_target_terms;  /* (target, cell) pairs those expansions stand for:
                                  pc_interactions + far_target_terms is the number of accepted
                                  (target, cell) pairs of the walk */
} nb_bh_stats;
/* Fills *out and returns NB_OK, or (with *out filled) NB_ERR_CAPACITY if an
 * evaluation since the previous call set flag 1, 2, 4 or 8 (its NaN
 * accelerations can make later inputs non-finite), else NB_ERR_INVALID if
 * one had a non-finite input (flag 16); NB_ERR_STATE before any evaluation.
 * Synchronizes. */
int nb_force_bh_stats(nb_force* f, nb_bh_stats* out);

/* Topology of the last Barnes-Hut tree, per particle in ORIGINAL order: depth
 * of the leaf holding it and its octant path (3 bits per level, root first).
 * Also the 63-bit keys (original order) and the sort permutation (sorted
 * position -> original index).  Any output pointer may be NULL.  Synchronizes. */
int nb_force_bh_topology(nb_force* f, int64_t n, int32_t* leaf_depth, uint64_t* leaf_path,
                         uint64_t* keys, uint32_t* perm);

/* Multipole moments of every non-empty node of the last tree (synchronizes).
 * Nodes are in level order, key order inside a level.  For each node:
 * level, path, particle range [first,last) in sorted order, mass, centre of
 * mass (3), traceless quadrupole about the COM (xx,yy,zz,xy,xz,yz).  Nodes of
 * levels 0..7 report their FP64 moments; deeper nodes keep (and report) the
 * FP32 moments the walk uses, accumulated in FP64 from their children's FP32
 * records.  Pass capacity = number of slots in the arrays; *count receives
 * the node count. */
int nb_force_bh_nodes(nb_force* f, int64_t capacity, int64_t* count, int32_t* level,
                      uint64_t* path, int32_t* first, int32_t* last, double* mass, double* com,
                      double* quad);

/* -------------------------------------------------------------- integrators */
typedef struct nb_integrator nb_integrator;

enum nb_integrator_kind {
    /* Kick-drift-kick leapfrog with the acceleration cached between steps
     * (CUDA/src/integrator_leafrog.cu:47-103); bit-identical to the two-eval
     * CPU form CPU/src/integrator_leapfrog.cpp:3-26 for deterministic forces. */
    NB_LEAPFROG = 0,
    /* Classic RK4 (CPU/src/integrator_rk4.cpp:3-47). */
    NB_RK4 = 1,
    /* Explicit Euler (CPU/src/integrator_euler.cpp:3-20, CUDA/src/integrator_euler.cu:24-50). */
    NB_EULER = 2,
    /* Euler-Cromer (CPU/src/integrator_euler-cromer.cpp:3-20,
     * CUDA/src/integrator_euler-cromer.cu). */
    NB_EULER_CROMER = 3,
    /* Velocity Verlet with the acceleration cached between steps, as
     * CUDA/src/integrator_velocity_verlet.cu:46-102 does; bit-identical to the
     * two-eval CPU form CPU/src/integrator_velocity_verlet.cpp:3-23 for
     * deterministic forces. */
    NB_VELOCITY_VERLET = 4
};

int nb_integrator_create(nb_integrator** it, int kind);
int nb_integrator_destroy(nb_integrator* it);
/* TimeIntegrator::step(ForceCalculator&, ParticleArrays&, double dt)
 * (CUDA/include/integrator.hpp:10-14; CPU form CPU/include/integrator.hpp:16-20). */
int nb_integrator_step(nb_integrator* it, nb_force* f, nb_particles* p, double dt);
/* `steps` calls of nb_integrator_step with the kick of one step fused into the
 * next step's kick-drift kernel; same bits as the step-by-step loop. */
int nb_integrator_run(nb_integrator* it, nb_force* f, nb_particles* p, double dt, int steps);
/* Forget the cached acceleration (next leapfrog step evaluates the force first). */
int nb_integrator_reset(nb_integrator* it);
/* TimeIntegrator::name() (CPU/include/integrator.hpp:22). */
const char* nb_integrator_name(const nb_integrator* it);

/* -------------------------------------------------------------- diagnostics */
/* Energy / momentum diagnostics reduced on the GPU in FP64 with compensated
 * (Neumaier) sums.  Replaces kineticEnergy / potentialEnergy / angularMomentum /
 * centerOfMass (CPU/include/tests_common.hpp:33-75, CUDA/src/test_common.cpp:63-182).
 * The potential is the O(N^2) pair sum; pe uses the reference's unsoftened
 * 1/r (tests_common.hpp:47-49), pe_soft uses 1/sqrt(r^2+eps^2). */
typedef struct nb_diag {
    double ke, pe, pe_soft;
    double p[3];   /* linear momentum */
    double l[3];   /* angular momentum */
    double com[3]; /* centre of mass */
    double mass;
} nb_diag;
int nb_diagnostics(const nb_particles* p, double G, double softening, int with_potential,
                   nb_diag* out);

/* ------------------------------------------------- initial conditions (host) */
/* Counter-based Philox4x32-10 (Salmon et al., SC'11) streams: identical on
 * every host.  All arrays are caller-owned host arrays of length n. */
/* ctr = {c0,c1,c2,c3}, key = {k0,k1} -> out[4]. */
void nb_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Plummer sphere: scale radius a, total mass M (m_i = M/n), G, isotropic
 * velocities from the distribution function (Aarseth, Henon & Wielen 1974);
 * radii truncated at rmax*a (rmax <= 0: no truncation); centre-of-mass frame. */
int nb_ic_plummer(int64_t n, uint64_t seed, double a, double M, double G, double rmax, double* m,
                  double* x, double* y, double* z, double* vx, double* vy, double* vz);
/* Uniform positions in [-1,1]^3, Gaussian velocities (sigma per axis), m = 1/n. */
int nb_ic_uniform_gaussian(int64_t n, uint64_t seed, double sigma, double* m, double* x, double* y,
                           double* z, double* vx, double* vy, double* vz);
/* The reference's fixture generator: one std::mt19937(seed) stream with
 * std::uniform_real_distribution<double>, masses in [0.5, 2) and coordinates
 * in [-100, 100).  order 0 draws m, x, y, z per particle -- the stream that
 * wrote particles_{100,500,1000,5000,10000}.txt from seed 42, one file after
 * the other (skip = particles drawn for the earlier files: 0, 100, 600, 1600,
 * 6600); order 1 draws z, y, x, m (the g++ build of
 * CPU/tests/test2_ForceBenchmark.cpp:48-57); order 2 draws x, y, z, m
 * (CUDA/tests/test2_ForceBenchmark.cu:63-66).  Velocities of the fixtures
 * are zero.  The first `skip` particles of the stream are drawn and dropped. */
int nb_ic_reference_fixture(int64_t n, uint32_t seed, int order, int64_t skip, double* m,
                            double* x, double* y, double* z);
/* Plummer sphere (fraction 1-disk_frac of M, scale a) plus a cold exponential
 * disk (disk_frac of M, scale length rd, sech^2 thickness zd) on circular
 * orbits in the enclosed mass. */
int nb_ic_plummer_disk(int64_t n, uint64_t seed, double a, double M, double G, double disk_frac,
                       double rd, double zd, double* m, double* x, double* y, double* z,
                       double* vx, double* vy, double* vz);

/* ------------------------------------------------------ particle-file layout */
/* Text layout of particles_*.txt: line 1 "N", then N lines "m x y z vx vy vz"
 * (reader: CUDA/tests/test3_correctness.cu:16-38).  nb_file_read with NULL
 * arrays only returns the count.  nb_file_write prints %.17g (exact round trip). */
int nb_file_read(const char* path, int64_t* n, double* m, double* x, double* y, double* z,
                 double* vx, double* vy, double* vz, int64_t capacity);
int nb_file_write(const char* path, int64_t n, const double* m, const double* x, const double* y,
                  const double* z, const double* vx, const double* vy, const double* vz);

/* ------------------------------------------------------- batched ensembles */
/* Many independent systems of uneven size advanced together: one persistent
 * direct-sum launch per force evaluation over all systems, leapfrog kick/drift
 * over all particles, step sequences captured in a HIP graph.  No reference
 * counterpart (the reference advances one system per process). */
typedef struct nb_ensemble nb_ensemble;
/* offsets[k] = first particle of system k (offsets[n_systems] = total);
 * G, softening, dt per system.  State arrays are host arrays of the total. */
int nb_ensemble_create(nb_ensemble** e, int n_systems, const int64_t* offsets, const double* G,
                       const double* softening, const double* dt);
int nb_ensemble_destroy(nb_ensemble* e);
int nb_ensemble_upload(nb_ensemble* e, const double* m, const double* x, const double* y,
                       const double* z, const double* vx, const double* vy, const double* vz);
int nb_ensemble_download(nb_ensemble* e, double* x, double* y, double* z, double* vx, double* vy,
                         double* vz, double* ax, double* ay, double* az);
/* Advance every system by `steps` leapfrog steps.  use_graph=1 replays a
 * captured HIP graph of the step sequence. */
int nb_ensemble_run(nb_ensemble* e, int steps, int use_graph);
/* The same steps with one launch sequence per system (baseline for the
 * batched path; same per-system arithmetic, hence the same bits). */
int nb_ensemble_run_unbatched(nb_ensemble* e, int steps);
/* Autotuner: time the tile shapes (0: 64 targets x 4 source slices, 1: 128 x 2,
 * 2: 256 x 1 per 256-lane workgroup) on every occupied size bucket of this
 * ensemble and keep the fastest (also written to `path` if not NULL).
 * nb_ensemble_tuning returns the 8 per-bucket shape codes; bucket b covers
 * N in [2^(b+6), 2^(b+7)), bucket 0 everything below 128, bucket 7 from 8192. */
int nb_ensemble_autotune(nb_ensemble* e, const char* path);
int nb_ensemble_load_tuning(nb_ensemble* e, const char* path);
int nb_ensemble_tuning(const nb_ensemble* e, int32_t* shape_codes, int n_buckets);

/* ---------------------------------------------------------------- serving */
/* Continuous-batching server: requests (one system each) are admitted against
 * an HBM budget, keep their state resident on the device, and every tick
 * advances every active request -- direct-method requests (n <
 * direct_threshold) together in batched launches, K leapfrog steps per tick
 * (K = 64 unless set by nb_server_set_direct_steps, and no more than any of
 * them has left), Barnes-Hut requests shortest remaining work first within a
 * per-tick budget (nb_server_set_tick_budget) on a pool of bh_lanes
 * calculators with their own streams, overlapping on the device.  A request that fails finishes early
 * with its error status (nb_server_result) without stopping the others.
 * Finished requests are downloaded and released; pending ones are admitted as
 * memory frees up.  Budget accounting: every active request is charged its
 * device state (80 B/particle; direct requests 16 B/particle more for the FP32
 * copy); the Barnes-Hut lanes are charged their workspaces at the size of the
 * largest n and smallest theta admitted so far, a charge that is never
 * released (the lanes keep that memory).  No reference counterpart (the
 * reference runs one system per process). */
typedef struct nb_server nb_server;
typedef struct nb_request {
    int64_t n;                  /* particles */
    int32_t steps;              /* leapfrog steps to run */
    double G, softening, dt;
    double theta;               /* opening angle (Barnes-Hut requests) */
    const double *m, *x, *y, *z, *vx, *vy, *vz; /* host arrays, copied at submission */
} nb_request;
typedef struct nb_server_status {
    int64_t pending, active, finished_uncollected;
    int64_t bytes_in_use;       /* budget charged to admitted requests and lanes */
    int64_t ticks, particle_steps;
} nb_server_status;
int nb_server_create(nb_server** s, double hbm_budget_bytes, int64_t direct_threshold,
                     int bh_lanes, int leaf_capacity);
int nb_server_destroy(nb_server* s);
/* Queues a request (NB_ERR_BUDGET if it can never fit). */
int nb_server_submit(nb_server* s, const nb_request* r, int64_t* id);
/* Admits what fits, advances every active request by one step, completes
 * the finished ones.  *n_completed receives how many finished in this tick. */
int nb_server_tick(nb_server* s, int* n_completed);
/* Ids of requests finished since the last poll. */
int nb_server_poll(nb_server* s, int64_t* ids, int capacity, int* count);
/* Ids of requests admitted (moved from pending to active, device state
 * allocated) since the last call, in admission order, so that a caller can
 * split a request's latency into queueing and service time.  The server
 * keeps one id per admitted request until it is polled here. */
int nb_server_poll_admitted(nb_server* s, int64_t* ids, int capacity, int* count);
/* Final positions/velocities of a finished request (then forgotten).
 * Returns the request's status: NB_OK, or the error that ended it early (a
 * failed launch or allocation, NB_ERR_CAPACITY / NB_ERR_INVALID for a
 * Barnes-Hut evaluation whose device checks failed, see nb_force_bh_stats);
 * the arrays then hold the state at that point. */
int nb_server_result(nb_server* s, int64_t id, double* x, double* y, double* z, double* vx,
                     double* vy, double* vz);
int nb_server_get_status(const nb_server* s, nb_server_status* out);
/* Direct-request steps per tick, 1..1024 (default 64).  More steps per tick
 * finish small requests sooner while tree steps of large ones share the
 * device; results do not depend on it (each step is the same arithmetic). */
int nb_server_set_direct_steps(nb_server* s, int steps_per_tick);
/* Barnes-Hut request steps per tick when the tick budget is 0, 1..1024
 * (default 1).  Results do not depend on it. */
int nb_server_set_bh_steps(nb_server* s, int steps_per_tick);
/* Barnes-Hut scheduling by remaining work.  particle_steps > 0: each tick
 * takes the Barnes-Hut requests in order of remaining work (n x steps left,
 * shortest first); each gets as many steps as its share of this budget
 * (particle_steps / bh_lanes) holds, at least one, and requests are added
 * until the budget is spent and every lane has one (bh_steps is then not
 * used), so short requests finish early.  Default 4e7 (about 0.15 s of
 * device time per tick at 2.7e8 particle-steps/s).  0: every Barnes-Hut
 * request advances up to bh_steps steps every tick.  Results do not depend
 * on it. */
int nb_server_set_tick_budget(nb_server* s, double particle_steps);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_AMD_H */
