// host_checks.cpp -- exercises the host-side C++ of the library (initial
// conditions, the reference fixture stream, particle files, argument
// checks, error messages) and the CPU oracle on small inputs.  Built by
// tests/test_sanitize_host.py with AddressSanitizer and
// UndefinedBehaviorSanitizer (g++, CPU only); any report fails the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nbody_amd.h"

extern "C" {
void orc_direct(int64_t, double, double, const double*, const double*, const double*,
                const double*, double*, double*, double*);
void orc_direct_subset(int64_t, double, double, const double*, const double*, const double*,
                       const double*, int64_t, const int64_t*, double*, double*, double*);
void orc_leapfrog_direct(int64_t, double, double, double, int, int, const double*, double*,
                         double*, double*, double*, double*, double*, double*, double*, double*);
void orc_rk4_direct(int64_t, double, double, double, int, const double*, double*, double*,
                    double*, double*, double*, double*, double*, double*, double*);
void orc_step1_direct(int64_t, int, double, double, double, int, const double*, double*, double*,
                      double*, double*, double*, double*, double*, double*, double*);
void orc_bhopt(int64_t, double, double, double, int, const double*, const double*, const double*,
               const double*, double*, double*, double*, int64_t*);
void orc_bhopt_sample(int64_t, double, double, double, int, const double*, const double*,
                      const double*, const double*, int64_t, double*, double*, double*, double*);
int64_t orc_bhopt_leaves(int64_t, int, const double*, const double*, const double*, const double*,
                         int32_t*, uint64_t*, int64_t*);
void orc_bhrec(int64_t, double, double, double, const double*, const double*, const double*,
               const double*, double*, double*, double*);
void orc_root(int64_t, const double*, const double*, const double*, double*, double*);
void orc_keys63(int64_t, const double*, const double*, const double*, const double*, double,
                uint64_t*);
void orc_diag(int64_t, double, double, const double*, const double*, const double*, const double*,
              const double*, const double*, const double*, double*);
void orc_error_stats(int64_t, const double*, const double*, const double*, const double*,
                     const double*, const double*, double*);
void orc_mt19937_particles(uint32_t, int64_t, int, double*, double*, double*, double*);
}

static int failures = 0;
#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static bool finite(const std::vector<double>& v) {
    for (double a : v)
        if (!std::isfinite(a)) return false;
    return true;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int64_t n = 700;
    std::vector<double> m(n), x(n), y(n), z(n), vx(n), vy(n), vz(n);

    // Initial conditions and their argument checks.
    CHECK(nb_ic_plummer(n, 3, 1.0, 1.0, 1.0, 0.0, m.data(), x.data(), y.data(), z.data(),
                        vx.data(), vy.data(), vz.data()) == NB_OK);
    CHECK(finite(x) && finite(vz));
    CHECK(nb_ic_plummer(n, 3, -1.0, 1.0, 1.0, 0.0, m.data(), x.data(), y.data(), z.data(),
                        vx.data(), vy.data(), vz.data()) == NB_ERR_INVALID);
    CHECK(std::strlen(nb_last_error()) > 0);
    CHECK(nb_ic_uniform_gaussian(n, 4, 0.1, m.data(), x.data(), y.data(), z.data(), vx.data(),
                                 vy.data(), vz.data()) == NB_OK);
    CHECK(nb_ic_plummer_disk(n, 5, 1.0, 1.0, 1.0, 0.1, 1.0, 0.05, m.data(), x.data(), y.data(),
                             z.data(), vx.data(), vy.data(), vz.data()) == NB_OK);
    CHECK(finite(m) && finite(y));
    for (int order = 0; order < 3; ++order) {
        // The fixture stream after 13 dropped particles equals the oracle's
        // stream from particle 13 on.
        const int64_t k = n + 13;
        std::vector<double> a(4 * n), b(4 * k);
        CHECK(nb_ic_reference_fixture(n, 42, order, 13, a.data(), a.data() + n, a.data() + 2 * n,
                                      a.data() + 3 * n) == NB_OK);
        orc_mt19937_particles(42, k, order, b.data(), b.data() + k, b.data() + 2 * k,
                              b.data() + 3 * k);
        for (int c = 0; c < 4; ++c)
            CHECK(std::memcmp(a.data() + c * n, b.data() + c * k + 13, n * sizeof(double)) == 0);
    }
    CHECK(nb_ic_reference_fixture(n, 42, 7, 0, m.data(), x.data(), y.data(), z.data()) ==
          NB_ERR_INVALID);
    CHECK(nb_ic_reference_fixture(0, 42, 0, 0, nullptr, nullptr, nullptr, nullptr) == NB_OK);
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
    nb_philox4x32_10(ctr, key, out);

    // Particle files: round trip, short capacity, missing and malformed files.
    CHECK(nb_ic_plummer(n, 9, 1.0, 1.0, 1.0, 0.0, m.data(), x.data(), y.data(), z.data(),
                        vx.data(), vy.data(), vz.data()) == NB_OK);
    const std::string path = dir + "/round_trip.txt";
    CHECK(nb_file_write(path.c_str(), n, m.data(), x.data(), y.data(), z.data(), vx.data(),
                        vy.data(), vz.data()) == NB_OK);
    int64_t nr = -1;
    CHECK(nb_file_read(path.c_str(), &nr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, 0) == NB_OK && nr == n);
    std::vector<double> r[7];
    for (auto& v : r) v.assign(n, 0.0);
    CHECK(nb_file_read(path.c_str(), &nr, r[0].data(), r[1].data(), r[2].data(), r[3].data(),
                       r[4].data(), r[5].data(), r[6].data(), n) == NB_OK);
    CHECK(r[1] == x && r[6] == vz && r[0] == m);
    CHECK(nb_file_read(path.c_str(), &nr, r[0].data(), r[1].data(), r[2].data(), r[3].data(),
                       r[4].data(), r[5].data(), r[6].data(), n - 1) == NB_ERR_INVALID);
    CHECK(nb_file_read((dir + "/missing.txt").c_str(), &nr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, 0) == NB_ERR_INVALID);
    const std::string bad = dir + "/bad.txt";
    FILE* f = std::fopen(bad.c_str(), "w");
    std::fputs("3\n1 2 3 4 5 6 7\n1 2\n", f);
    std::fclose(f);
    CHECK(nb_file_read(bad.c_str(), &nr, r[0].data(), r[1].data(), r[2].data(), r[3].data(),
                       r[4].data(), r[5].data(), r[6].data(), n) == NB_ERR_INVALID);
    CHECK(nb_file_write(path.c_str(), -1, m.data(), x.data(), y.data(), z.data(), vx.data(),
                        vy.data(), vz.data()) == NB_ERR_INVALID);

    // The oracle on the same small system.
    std::vector<double> ax(n), ay(n), az(n), bx(n), by(n), bz(n);
    orc_direct(n, 1.0, 1e-4, m.data(), x.data(), y.data(), z.data(), ax.data(), ay.data(), az.data());
    CHECK(finite(ax) && finite(az));
    std::vector<int64_t> idx = {0, 5, 699};
    orc_direct_subset(n, 1.0, 1e-4, m.data(), x.data(), y.data(), z.data(), 3, idx.data(),
                      bx.data(), by.data(), bz.data());
    CHECK(bx[2] == ax[699] && bz[1] == az[5]);
    int64_t stats[4];
    orc_bhopt(n, 1.0, 1e-4, 0.5, 8, m.data(), x.data(), y.data(), z.data(), bx.data(), by.data(),
              bz.data(), stats);
    CHECK(finite(bx) && stats[0] > 0);
    double ms[2];
    orc_bhopt_sample(n, 1.0, 1e-4, 0.5, 8, m.data(), x.data(), y.data(), z.data(), 70, bx.data(),
                     by.data(), bz.data(), ms);
    std::vector<int32_t> depth(n);
    std::vector<uint64_t> lpath(n), keys(n);
    int64_t internal = 0;
    CHECK(orc_bhopt_leaves(n, 8, m.data(), x.data(), y.data(), z.data(), depth.data(),
                           lpath.data(), &internal) > 0);
    orc_bhrec(n, 1.0, 1e-4, 0.5, m.data(), x.data(), y.data(), z.data(), bx.data(), by.data(),
              bz.data());
    CHECK(finite(by));
    double center[3], size = 0;
    orc_root(n, x.data(), y.data(), z.data(), center, &size);
    orc_keys63(n, x.data(), y.data(), z.data(), center, size, keys.data());
    double dg[13], es[4];
    orc_diag(n, 1.0, 1e-4, m.data(), x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), dg);
    CHECK(std::isfinite(dg[0]) && std::isfinite(dg[1]));
    orc_error_stats(n, bx.data(), by.data(), bz.data(), ax.data(), ay.data(), az.data(), es);
    CHECK(es[3] >= es[2]);
    std::vector<double> s[9];
    for (auto& v : s) v.assign(n, 0.0);
    auto state = [&]() {
        s[0] = x; s[1] = y; s[2] = z; s[3] = vx; s[4] = vy; s[5] = vz;
    };
    state();
    orc_leapfrog_direct(n, 1.0, 1e-4, 1e-3, 3, 1, m.data(), s[0].data(), s[1].data(), s[2].data(),
                        s[3].data(), s[4].data(), s[5].data(), s[6].data(), s[7].data(), s[8].data());
    CHECK(finite(s[0]) && finite(s[8]));
    state();
    orc_rk4_direct(n, 1.0, 1e-4, 1e-3, 2, m.data(), s[0].data(), s[1].data(), s[2].data(),
                   s[3].data(), s[4].data(), s[5].data(), s[6].data(), s[7].data(), s[8].data());
    for (int kind = 0; kind < 3; ++kind) {
        state();
        orc_step1_direct(n, kind, 1.0, 1e-4, 1e-3, 2, m.data(), s[0].data(), s[1].data(),
                         s[2].data(), s[3].data(), s[4].data(), s[5].data(), s[6].data(),
                         s[7].data(), s[8].data());
        CHECK(finite(s[3]));
    }
    if (failures) return 1;
    std::printf("host checks passed\n");
    return 0;
}
