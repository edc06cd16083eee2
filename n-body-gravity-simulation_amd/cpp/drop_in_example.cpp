// drop_in_example.cpp -- a reference-style test main written only against the
// drop-in classes (nbody.hpp), the way CPU/tests/test2_ForceBenchmark.cpp and
// CUDA/tests/test3_correctness.cu use the reference classes.
//
// usage: drop_in_example <particles.txt> <out.bin>
// Writes, as raw little-endian doubles, in order:
//   1. accelerations (ax, ay, az) of ForceCalculatorNaive(0.01) through the
//      CPU-style vector API;
//   2. state (x, y, z, vx, vy, vz, ax, ay, az) after 10 LeapfrogIntegrator steps
//      (dt = 1e-3) through the CPU-style vector API with ForceCalculatorNaive(0.01);
//   3. the same state after 5 RK4Integrator steps through the CUDA-style
//      ParticleArrays API;
//   4. accelerations of BarnesHutForceCalculator(0.01, 0.5, 8) through the
//      CUDA-style API;
//   5. accelerations of the recursive-build ForceCalculatorBarnesHut(0.01, 0.5)
//      through the CPU-style vector API.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <vector>

#include "nbody.hpp"

static void put(std::ofstream& o, const std::vector<double>& v) {
    o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

// Host copy of device-resident state, sized from particles.num_particles the
// way the reference's CUDA test helpers size theirs (CUDA/src/test_common.cpp:32).
struct HostState {
    std::vector<double> x, y, z, vx, vy, vz, ax, ay, az, m;
};
static HostState hostState(const ParticleArrays& particles) {
    const size_t n = particles.num_particles;
    HostState h;
    for (auto* v : {&h.x, &h.y, &h.z, &h.vx, &h.vy, &h.vz, &h.ax, &h.ay, &h.az, &h.m}) v->resize(n);
    copyParticlesToHost(particles, h.x.data(), h.y.data(), h.z.data(), h.vx.data(), h.vy.data(),
                        h.vz.data(), h.ax.data(), h.ay.data(), h.az.data(), h.m.data());
    return h;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: " << argv[0] << " particles.txt out.bin\n";
        return 2;
    }
    try {
        // Particle-file layout: N, then m x y z vx vy vz per line
        // (CUDA/tests/test3_correctness.cu:27-38).
        std::ifstream in(argv[1]);
        int N = 0;
        in >> N;
        std::vector<Particle> particles;
        particles.reserve(N);
        std::vector<double> hm(N), hx(N), hy(N), hz(N), hvx(N), hvy(N), hvz(N);
        for (int i = 0; i < N; ++i) {
            in >> hm[i] >> hx[i] >> hy[i] >> hz[i] >> hvx[i] >> hvy[i] >> hvz[i];
            particles.emplace_back(hm[i], Vector3D(hx[i], hy[i], hz[i]),
                                   Vector3D(hvx[i], hvy[i], hvz[i]), "Particle " + std::to_string(i));
        }
        std::ofstream out(argv[2], std::ios::binary);

        // 1. CPU-style force call.
        {
            std::vector<Particle> p = particles;
            ForceCalculatorNaive naive(0.01);
            naive.computeAccelerations(p);
            std::vector<double> a;
            for (const auto& q : p) a.push_back(q.acceleration.x);
            for (const auto& q : p) a.push_back(q.acceleration.y);
            for (const auto& q : p) a.push_back(q.acceleration.z);
            put(out, a);
            std::cout << naive.name() << " calls=" << naive.getTimeStatistics().calls << "\n";
        }
        // 2. CPU-style integration (runSimulation, CPU/include/tests_common.hpp:19-31).
        {
            std::vector<Particle> p = particles;
            ForceCalculatorNaive naive(0.01);
            LeapfrogIntegrator lf;
            for (int s = 0; s < 10; ++s) lf.step(p, naive, 1e-3);
            std::vector<double> st;
            for (int k = 0; k < 9; ++k)
                for (const auto& q : p) {
                    const Vector3D& v = k < 3 ? q.position : k < 6 ? q.velocity : q.acceleration;
                    st.push_back(k % 3 == 0 ? v.x : k % 3 == 1 ? v.y : v.z);
                }
            put(out, st);
            std::cout << lf.name() << " 10 steps done\n";
        }
        // 3. CUDA-style integration on device-resident state.
        {
            ParticleArrays d{};
            allocateParticles(d, N);
            copyParticlesToDevice(d, hx.data(), hy.data(), hz.data(), hvx.data(), hvy.data(),
                                  hvz.data(), hm.data());
            ForceCalculatorNaive naive(0.01);
            RK4Integrator rk;
            for (int s = 0; s < 5; ++s) rk.step(naive, d, 1e-3);
            const HostState h = hostState(d);
            if (d.num_particles != N || h.m != hm) {
                std::cerr << "device state does not round-trip\n";
                return 1;
            }
            for (auto* v : {&h.x, &h.y, &h.z, &h.vx, &h.vy, &h.vz, &h.ax, &h.ay, &h.az}) put(out, *v);
            freeParticles(d);
            if (d.num_particles != 0 || d.pos_x != nullptr) {
                std::cerr << "freeParticles left the arrays set\n";
                return 1;
            }
            std::cout << rk.name() << " 5 steps done\n";
        }
        // 4. CUDA-style Barnes-Hut force.
        {
            ParticleArrays d{};
            allocateParticles(d, N);
            copyParticlesToDevice(d, hx.data(), hy.data(), hz.data(), nullptr, nullptr, nullptr,
                                  hm.data());
            BarnesHutForceCalculator bh(0.01, 0.5, 8);
            bh.computeAccelerations(d);
            std::vector<double> ax(N), ay(N), az(N);
            copyParticlesToHost(d, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ax.data(),
                                ay.data(), az.data(), nullptr);
            put(out, ax);
            put(out, ay);
            put(out, az);
            const nb_bh_stats s = bh.stats();
            std::cout << bh.name() << " nodes=" << s.n_nodes << " build_ms=" << bh.getBuildTreeTime()
                      << " walk_ms=" << bh.getTraversalTime() << "\n";
            freeParticles(d);
        }
        // 5. CPU-style recursive Barnes-Hut (leaf capacity 1), vector API.
        {
            std::vector<Particle> ps = particles;
            ForceCalculatorBarnesHut rec(0.01, 0.5);
            rec.computeAccelerations(ps);
            std::vector<double> ax(N), ay(N), az(N);
            for (int i = 0; i < N; ++i) {
                ax[i] = ps[i].acceleration.x;
                ay[i] = ps[i].acceleration.y;
                az[i] = ps[i].acceleration.z;
            }
            put(out, ax);
            put(out, ay);
            put(out, az);
            std::cout << rec.name() << " (recursive) build_ms="
                      << rec.getBuildingTreeStatistics().total_time
                      << " walk_ms=" << rec.getTraversalStatistics().total_time
                      << " calls=" << rec.getTraversalStatistics().calls << "\n";
        }
        // Error path: an invalid theta is reported as an exception, not exit().
        try {
            BarnesHutForceCalculator bad(0.0, -1.0, 8);
            std::cerr << "expected an exception\n";
            return 1;
        } catch (const nbody::Error& e) {
            std::cout << "invalid argument reported: code " << e.code << "\n";
        }
        // Error path: a non-finite position is found on the device during the
        // evaluation; the vector API's device check turns it into an exception.
        try {
            std::vector<Particle> ps = particles;
            ps[7].position.x = std::numeric_limits<double>::infinity();
            ForceCalculatorBarnesHutOptimized opt(0.01, 0.5, 8);
            opt.computeAccelerations(ps);
            std::cerr << "expected an exception for a non-finite position\n";
            return 1;
        } catch (const nbody::Error& e) {
            std::cout << "non-finite input reported: code " << e.code << "\n";
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
