// leaf_capacity_sweep.cpp -- leaf-capacity / opening-angle sweep of the GPU
// Barnes-Hut calculator, the counterpart of the reference's
// CPU/tests/test3_computeLeafCapacity.cpp, written against the drop-in classes.
//
// For every particle count and experiment it draws the reference's inputs
// (std::mt19937(42 + experiment); mass U(0.5, 2), positions U(-100, 100)^3,
// zero velocities, drawn in the order z, y, x, m of the reference's g++
// build), computes the FP64 direct-sum reference with
// ForceCalculatorNaive, then for every (theta, leaf capacity) times
// `iterations` Barnes-Hut evaluations after 5 warmups and accumulates the mean
// and maximum relative force error (targets with |a_ref| < 1e-12 skipped, the
// sum divided by N, as the reference does).  Results are averaged over the
// experiments and printed sorted by time and by accuracy, as the reference
// prints them, followed by one JSON line per particle count with the fastest
// configuration whose mean relative error is below --max-error.
//
// With --input plummer the inputs are seeded Plummer spheres (a = 1, M = 1,
// G as the calculators use it) instead of the reference's uniform cubes.
// --write-tuning PATH writes the leaf-capacity table that NB_LEAF_AUTO
// calculators read (nb_bh_set_tuning_file): one "n_min leaf_capacity" line
// per particle count, the fastest capacity under --max-error, switching at
// the geometric midpoint between consecutive counts -- the per-N
// recommendation table of the reference (CPU/documentation.md:406-418) in a
// form the library consumes.
//
// usage: leaf_capacity_sweep [--n 100,500,1000,5000,10000] [--caps 1,2,4,8,16,32,64]
//                            [--theta 0.5] [--experiments 5] [--iterations 100]
//                            [--softening 0] [--monopole] [--max-error 1e-3]
//                            [--input reference|plummer] [--write-tuning PATH]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "nbody.hpp"

namespace {

struct Result {
    double theta = 0.0;
    int leafCapacity = 0;
    double averageTime = 0.0;  // ms per force evaluation
    double averageRelativeError = 0.0;
    double maximumRelativeError = 0.0;
};

std::vector<double> parse_list(const char* s) {
    std::vector<double> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) out.push_back(std::atof(tok.c_str()));
    return out;
}

void print_table(const std::vector<Result>& rs, bool by_time) {
    std::cout << std::left << std::setw(8) << "Theta" << std::setw(8) << "Leaf";
    if (by_time)
        std::cout << std::setw(15) << "Time(ms)" << std::setw(18) << "Avg Rel Error"
                  << std::setw(18) << "Max Rel Error" << "\n";
    else
        std::cout << std::setw(18) << "Avg Rel Error" << std::setw(18) << "Max Rel Error"
                  << std::setw(15) << "Time(ms)" << "\n";
    for (const Result& r : rs) {
        std::cout << std::setw(8) << r.theta << std::setw(8) << r.leafCapacity;
        if (by_time)
            std::cout << std::setw(15) << r.averageTime << std::setw(18) << r.averageRelativeError
                      << std::setw(18) << r.maximumRelativeError << "\n";
        else
            std::cout << std::setw(18) << r.averageRelativeError << std::setw(18)
                      << r.maximumRelativeError << std::setw(15) << r.averageTime << "\n";
    }
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<double> counts = {100, 500, 1000, 5000, 10000};
    std::vector<double> caps = {1, 2, 4, 8, 16, 32, 64};
    std::vector<double> thetas = {0.5};
    int experiments = 5, iterations = 100;
    double softening = 0.0, max_error = 1e-3;
    int multipole = NB_QUADRUPOLE;
    bool plummer = false;
    const char* tuning_path = nullptr;
    for (int a = 1; a < argc; ++a) {
        const bool more = a + 1 < argc;
        if (!std::strcmp(argv[a], "--n") && more) counts = parse_list(argv[++a]);
        else if (!std::strcmp(argv[a], "--caps") && more) caps = parse_list(argv[++a]);
        else if (!std::strcmp(argv[a], "--theta") && more) thetas = parse_list(argv[++a]);
        else if (!std::strcmp(argv[a], "--experiments") && more) experiments = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--iterations") && more) iterations = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--softening") && more) softening = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--max-error") && more) max_error = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--monopole")) multipole = NB_MONOPOLE;
        else if (!std::strcmp(argv[a], "--input") && more) {
            const char* v = argv[++a];
            if (!std::strcmp(v, "plummer")) plummer = true;
            else if (std::strcmp(v, "reference")) {
                std::cerr << "--input must be reference or plummer\n";
                return 2;
            }
        } else if (!std::strcmp(argv[a], "--write-tuning") && more) tuning_path = argv[++a];
        else {
            std::cerr << "unknown argument " << argv[a] << "\n";
            return 2;
        }
    }
    if (experiments < 1 || iterations < 1) {
        std::cerr << "experiments and iterations must be positive\n";
        return 2;
    }
    std::sort(counts.begin(), counts.end());
    std::vector<std::pair<int, int>> table;  // (n, recommended leaf capacity)
    try {
        for (double Nd : counts) {
            const int N = (int)Nd;
            std::cout << "\n=============================================================\n";
            std::cout << "Particle Count : " << N << "\n";
            std::cout << "=============================================================\n";
            std::vector<Result> results;
            for (double th : thetas)
                for (double c : caps) {
                    Result r;
                    r.theta = th;
                    r.leafCapacity = (int)c;
                    results.push_back(r);
                }
            for (int experiment = 0; experiment < experiments; experiment++) {
                std::vector<Particle> particles;
                particles.reserve(N);
                if (plummer) {
                    std::vector<double> m(N), x(N), y(N), z(N), vx(N), vy(N), vz(N);
                    nbody::check(nb_ic_plummer(N, 42 + experiment, 1.0, 1.0, constant::G, 0.0,
                                               m.data(), x.data(), y.data(), z.data(), vx.data(),
                                               vy.data(), vz.data()),
                                 "nb_ic_plummer");
                    for (int i = 0; i < N; i++)
                        particles.emplace_back(m[i], Vector3D(x[i], y[i], z[i]), Vector3D(0, 0, 0),
                                               "Particle");
                } else {
                    // The reference's inputs: std::mt19937(42 + experiment), drawn
                    // inside one emplace_back call, which g++ evaluates right to
                    // left: the stream order is z, y, x, m.
                    std::vector<double> m(N), x(N), y(N), z(N);
                    nbody::check(nb_ic_reference_fixture(N, 42 + experiment, 1, 0, m.data(),
                                                         x.data(), y.data(), z.data()),
                                 "nb_ic_reference_fixture");
                    for (int i = 0; i < N; i++)
                        particles.emplace_back(m[i], Vector3D(x[i], y[i], z[i]), Vector3D(0, 0, 0),
                                               "Particle");
                }
                std::vector<Particle> reference = particles;
                ForceCalculatorNaive naive(softening);
                naive.computeAccelerations(reference);

                for (Result& res : results) {
                    std::vector<Particle> current = particles;
                    BarnesHutForceCalculator bh(softening, res.theta, res.leafCapacity, multipole);
                    ParticleArrays d{};
                    allocateParticles(d, N);
                    {
                        nbody::HostSoA h;
                        h.from(current);
                        copyParticlesToDevice(d, h.x.data(), h.y.data(), h.z.data(), h.vx.data(),
                                              h.vy.data(), h.vz.data(), h.m.data());
                    }
                    for (int i = 0; i < 5; i++) bh.computeAccelerations(d);
                    bh.resetTimeStatistics();
                    for (int i = 0; i < iterations; i++) bh.computeAccelerations(d);
                    const double ms = bh.getTotalComputeTime() / iterations;
                    std::vector<double> ax(N), ay(N), az(N);
                    copyParticlesToHost(d, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                        ax.data(), ay.data(), az.data(), nullptr);
                    freeParticles(d);
                    double avgRel = 0.0, maxRel = 0.0;
                    for (int i = 0; i < N; i++) {
                        const double ref = reference[i].acceleration.norm();
                        if (ref < 1e-12) continue;
                        const double rel =
                            (Vector3D(ax[i], ay[i], az[i]) - reference[i].acceleration).norm() / ref;
                        avgRel += rel;
                        maxRel = std::max(maxRel, rel);
                    }
                    avgRel /= N;
                    res.averageTime += ms;
                    res.averageRelativeError += avgRel;
                    res.maximumRelativeError += maxRel;
                }
            }
            for (Result& r : results) {
                r.averageTime /= experiments;
                r.averageRelativeError /= experiments;
                r.maximumRelativeError /= experiments;
            }
            auto byTime = results;
            std::sort(byTime.begin(), byTime.end(),
                      [](const Result& a, const Result& b) { return a.averageTime < b.averageTime; });
            std::cout << "\n================ Fastest =================\n\n";
            print_table(byTime, true);
            auto byAccuracy = results;
            std::sort(byAccuracy.begin(), byAccuracy.end(), [](const Result& a, const Result& b) {
                return a.averageRelativeError < b.averageRelativeError;
            });
            std::cout << "\n============== Most Accurate ==============\n\n";
            print_table(byAccuracy, false);

            const Result* best = nullptr;
            for (const Result& r : byTime)
                if (r.averageRelativeError < max_error) {
                    best = &r;
                    break;
                }
            if (best) table.emplace_back(N, best->leafCapacity);
            std::cout << "\n{\"n\": " << N << ", \"max_error\": " << max_error;
            if (best)
                std::cout << ", \"theta\": " << best->theta << ", \"leaf_capacity\": "
                          << best->leafCapacity << ", \"ms\": " << best->averageTime
                          << ", \"avg_rel_error\": " << best->averageRelativeError;
            else
                std::cout << ", \"theta\": null, \"leaf_capacity\": null";
            std::cout << ", \"results\": [";
            for (size_t k = 0; k < results.size(); ++k)
                std::cout << (k ? ", " : "") << "[" << results[k].theta << ", "
                          << results[k].leafCapacity << ", " << results[k].averageTime << ", "
                          << results[k].averageRelativeError << ", "
                          << results[k].maximumRelativeError << "]";
            std::cout << "]}\n";
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    if (tuning_path) {
        if (table.empty()) {
            std::cerr << "no configuration met --max-error: no table written\n";
            return 1;
        }
        std::ofstream t(tuning_path);
        t << "# leaf capacity per particle count for NB_LEAF_AUTO Barnes-Hut calculators\n"
          << "# command:";
        for (int a = 0; a < argc; ++a) t << " " << argv[a];
        t << "\n# (leaf_capacity_sweep: input " << (plummer ? "plummer" : "reference")
          << ", theta";
        for (double th : thetas) t << " " << th;
        t << ", " << (multipole == NB_QUADRUPOLE ? "quadrupole" : "monopole")
          << ", fastest with mean relative error < " << max_error << ")\n"
          << "# n_min leaf_capacity\n";
        for (size_t k = 0; k < table.size(); ++k) {
            const long long nmin =
                k == 0 ? 0
                       : (long long)std::llround(std::sqrt((double)table[k - 1].first * table[k].first));
            t << nmin << " " << table[k].second << "   # measured at n = " << table[k].first << "\n";
        }
        if (!t) {
            std::cerr << "cannot write " << tuning_path << "\n";
            return 1;
        }
        std::cout << "wrote " << tuning_path << "\n";
    }
    return 0;
}
