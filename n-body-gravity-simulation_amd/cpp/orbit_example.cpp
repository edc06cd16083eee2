// orbit_example.cpp -- the reference's Earth-Sun program (CPU/src/n-body.cpp)
// written against the drop-in classes: velocity Verlet, dt = 1/365 yr,
// 366 steps, one trajectory row per step in the orbit.csv column layout.
//
// usage: orbit_example <out.csv>
//
// The reference uses ForceCalculatorBarnesHut(0.0, 0.3); with two bodies its
// walk opens the root and meets the other body's leaf, so the force is the
// direct pair force.  This program uses the FP64 direct calculator, which
// computes that pair force in FP64 (the GPU Barnes-Hut walk is FP32).
#include <fstream>
#include <iostream>
#include <vector>

#include "nbody.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " out.csv\n";
        return 2;
    }
    try {
        std::ofstream out(argv[1]);
        const double earth_mass = 3.003e-6;  // solar masses
        // Sun (index 0) at rest at the origin, Earth (index 1) at 1 AU with 2*pi AU/yr.
        const double m[2] = {1.0, earth_mass};
        const double x[2] = {0.0, 1.0}, y[2] = {0.0, 0.0}, z[2] = {0.0, 0.0};
        const double vx[2] = {0.0, 0.0}, vy[2] = {0.0, 2 * constant::PI}, vz[2] = {0.0, 0.0};
        ParticleArrays d{};
        allocateParticles(d, 2);
        copyParticlesToDevice(d, x, y, z, vx, vy, vz, m);

        ForceCalculatorNaive force_calculator(0.0);
        VelocityVerletIntegrator integrator;
        nbody::OrbitCsvWriter csv(out, /*body=*/1, /*center=*/0);
        csv.row(d);
        const double dt = 1.0 / 365.0;
        const int steps = static_cast<int>(1.0 / dt);
        for (int step = 0; step <= steps; step++) {
            integrator.step(force_calculator, d, dt);
            csv.row(d);
        }
        freeParticles(d);
        std::cout << integrator.name() << ": " << steps + 1 << " steps written to " << argv[1]
                  << "\n";
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
