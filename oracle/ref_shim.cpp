// ref_shim.cpp -- C entry points over the reference's own CPU classes, compiled
// together with the reference sources where they lie (oracle/Makefile, target
// `ref`) into oracle/_ref/libnbref.so.  Test infrastructure only: it lets the
// tests compare the oracle restatement against the unmodified reference code,
// and lets bench.py time the reference CPU path as a baseline.  No reference
// source is copied; this file only calls the reference classes.
//
// The reference hard-codes G = 4*pi^2 (CPU/include/common.hpp:28) and takes the
// softening length (not its square) in its constructors
// (CPU/include/force_calculator.hpp:19).
#include <chrono>
#include <cstdint>
#include <vector>

#include "force_calculator_barnes_hut.hpp"
#include "force_calculator_barnes_hut_optimized.hpp"
#include "force_calculator_naive.hpp"
#include "integrator_leapfrog.hpp"
#include "integrator_rk4.hpp"
#include "integrator_euler.hpp"
#include "integrator_euler-cromer.hpp"
#include "integrator_velocity_verlet.hpp"

namespace {

std::vector<Particle> make(int64_t n, const double* m, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy, const double* vz) {
    std::vector<Particle> p;
    p.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        p.emplace_back(m[i], Vector3D(x[i], y[i], z[i]),
                       vx ? Vector3D(vx[i], vy[i], vz[i]) : Vector3D(0, 0, 0));
    return p;
}

void store(const std::vector<Particle>& p, double* x, double* y, double* z, double* vx, double* vy,
           double* vz, double* ax, double* ay, double* az) {
    for (size_t i = 0; i < p.size(); ++i) {
        if (x) {
            x[i] = p[i].position.x;
            y[i] = p[i].position.y;
            z[i] = p[i].position.z;
        }
        if (vx) {
            vx[i] = p[i].velocity.x;
            vy[i] = p[i].velocity.y;
            vz[i] = p[i].velocity.z;
        }
        if (ax) {
            ax[i] = p[i].acceleration.x;
            ay[i] = p[i].acceleration.y;
            az[i] = p[i].acceleration.z;
        }
    }
}

// Exposes the optimized tree's node pool for topology dumps.
struct PoolProbe : ForceCalculatorBarnesHutOptimized {
    using ForceCalculatorBarnesHutOptimized::ForceCalculatorBarnesHutOptimized;
    // buildTree() reads the particles through particles_, which only
    // computeAccelerations() sets (force_calculator_barnes_hut_optimized.cpp:7).
    void build(std::vector<Particle>& p) {
        particles_ = &p;
        buildTree(p);
    }
    void leaves(int32_t* depth, uint64_t* path, int64_t* n_internal, int64_t* n_nodes) {
        struct It {
            int node;
            uint64_t path;
            int depth;
        };
        std::vector<It> st{{rootIndex, 0, 0}};
        int64_t internal = 0;
        while (!st.empty()) {
            It it = st.back();
            st.pop_back();
            const Node& nd = nodePool[it.node];
            if (nd.firstChild == -1) {
                for (int q : nd.particleIndices) {
                    depth[q] = it.depth;
                    path[q] = it.path;
                }
                continue;
            }
            ++internal;
            for (int k = 0; k < 8; ++k)
                st.push_back(It{nd.firstChild + k, (it.path << 3) | (uint64_t)k, it.depth + 1});
        }
        *n_internal = internal;
        *n_nodes = (int64_t)nextFreeNode;
    }
};

// Exposes the recursive tree root so a sample of targets can be walked.
struct RecProbe : ForceCalculatorBarnesHut {
    using ForceCalculatorBarnesHut::ForceCalculatorBarnesHut;
    Node_* tree_root() { return root; }
    void drop() {
        destroyTree(root);
        root = nullptr;
    }
};

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

void ref_direct(int64_t n, double softening, const double* m, const double* x, const double* y,
                const double* z, double* ax, double* ay, double* az) {
    auto p = make(n, m, x, y, z, nullptr, nullptr, nullptr);
    ForceCalculatorNaive f(softening);
    f.computeAccelerations(p);
    store(p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ax, ay, az);
}

// integrator: 0 = LeapfrogIntegrator, 1 = RK4Integrator, 2 = EulerIntegrator,
// 3 = EulerCromerIntegrator, 4 = VelocityVerletIntegrator; force: ForceCalculatorNaive.
void ref_integrate_direct(int64_t n, int integrator, double softening, double dt, int steps,
                          const double* m, double* x, double* y, double* z, double* vx, double* vy,
                          double* vz, double* ax, double* ay, double* az) {
    auto p = make(n, m, x, y, z, vx, vy, vz);
    ForceCalculatorNaive f(softening);
    auto run = [&](TimeIntegrator& it) {
        for (int k = 0; k < steps; ++k) it.step(p, f, dt);
    };
    if (integrator == 0) {
        LeapfrogIntegrator it;
        run(it);
    } else if (integrator == 1) {
        RK4Integrator it;
        run(it);
    } else if (integrator == 2) {
        EulerIntegrator it;
        run(it);
    } else if (integrator == 3) {
        EulerCromerIntegrator it;
        run(it);
    } else {
        VelocityVerletIntegrator it;
        run(it);
    }
    store(p, x, y, z, vx, vy, vz, ax, ay, az);
}

void ref_bhopt(int64_t n, double softening, double theta, int leaf_cap, const double* m,
               const double* x, const double* y, const double* z, double* ax, double* ay,
               double* az) {
    auto p = make(n, m, x, y, z, nullptr, nullptr, nullptr);
    ForceCalculatorBarnesHutOptimized f(softening, theta, leaf_cap);
    f.computeAccelerations(p);
    store(p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ax, ay, az);
}

void ref_bhopt_leaves(int64_t n, int leaf_cap, const double* m, const double* x, const double* y,
                      const double* z, int32_t* depth, uint64_t* path, int64_t* n_internal,
                      int64_t* n_nodes) {
    auto p = make(n, m, x, y, z, nullptr, nullptr, nullptr);
    PoolProbe f(0.0, 0.5, leaf_cap);
    f.build(p);
    f.leaves(depth, path, n_internal, n_nodes);
}

void ref_bhrec(int64_t n, double softening, double theta, const double* m, const double* x,
               const double* y, const double* z, double* ax, double* ay, double* az) {
    auto p = make(n, m, x, y, z, nullptr, nullptr, nullptr);
    ForceCalculatorBarnesHut f(softening, theta);
    f.computeAccelerations(p);
    store(p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ax, ay, az);
}

// Times the recursive reference Barnes-Hut on a bounded sample: full tree build,
// walk of `n_sample` targets (indices stride apart, OpenMP over targets exactly
// like computeAccelerations), and tree destruction.  ms = {build, walk, destroy}.
void ref_bhrec_sample(int64_t n, double softening, double theta, const double* m, const double* x,
                      const double* y, const double* z, int64_t n_sample, double* ax, double* ay,
                      double* az, double* ms) {
    auto p = make(n, m, x, y, z, nullptr, nullptr, nullptr);
    RecProbe f(softening, theta);
    auto t0 = std::chrono::steady_clock::now();
    f.buildTree(p);
    ms[0] = ms_since(t0);
    const int64_t stride = n_sample > 0 ? (n / n_sample > 0 ? n / n_sample : 1) : 1;
    t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n_sample; ++k) {
        Particle& q = p[(k * stride) % n];
        q.acceleration = Vector3D(0, 0, 0);
        f.computeAccelerationForParticle(f.tree_root(), &q);
    }
    ms[1] = ms_since(t0);
    for (int64_t k = 0; k < n_sample; ++k) {
        const Particle& q = p[(k * stride) % n];
        ax[k] = q.acceleration.x;
        ay[k] = q.acceleration.y;
        az[k] = q.acceleration.z;
    }
    t0 = std::chrono::steady_clock::now();
    f.drop();
    ms[2] = ms_since(t0);
}

}  // extern "C"
