// nb_host.cpp -- host-side initial conditions (Philox4x32-10 streams) and the
// particle-file reader/writer.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "nb_internal.hpp"

namespace {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with a Weyl key
// schedule.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Stream of uniform doubles for one particle: counter (i_lo, i_hi, stream, block).
struct Stream {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int used = 4;
    Stream(uint64_t seed, uint64_t particle, uint32_t stream) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        ctr[0] = (uint32_t)particle;
        ctr[1] = (uint32_t)(particle >> 32);
        ctr[2] = stream;
        ctr[3] = 0;
    }
    uint32_t next32() {
        if (used == 4) {
            philox(ctr, key, buf);
            ctr[3]++;
            used = 0;
        }
        return buf[used++];
    }
    // Uniform in (0, 1): 53 random bits, offset by half an ulp step.
    double uniform() {
        const uint64_t a = next32() >> 5, b = next32() >> 6;
        return ((double)((a << 26) | b) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double normal() {  // Box-Muller (one of the pair)
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    void direction(double& dx, double& dy, double& dz) {
        const double cz = 2.0 * uniform() - 1.0;
        const double ph = 2.0 * M_PI * uniform();
        const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        dx = sz * std::cos(ph);
        dy = sz * std::sin(ph);
        dz = cz;
    }
};

// Plummer sphere, G = M = a = 1 units scaled afterwards (Aarseth, Henon &
// Wielen 1974): r = (X^{-2/3} - 1)^{-1/2}; speed q * v_esc with q drawn from
// g(q) = q^2 (1-q^2)^{7/2} by rejection; v_esc = sqrt(2) (1 + r^2)^{-1/4}.
void plummer_particle(Stream& s, double rmax, double& x, double& y, double& z, double& vx,
                      double& vy, double& vz) {
    double r;
    for (;;) {
        const double X = s.uniform();
        r = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
        if (rmax <= 0.0 || r <= rmax) break;
    }
    double ux, uy, uz;
    s.direction(ux, uy, uz);
    x = r * ux;
    y = r * uy;
    z = r * uz;
    double q;
    for (;;) {
        const double X4 = s.uniform(), X5 = s.uniform();
        if (0.1 * X5 < X4 * X4 * std::pow(1.0 - X4 * X4, 3.5)) {
            q = X4;
            break;
        }
    }
    const double v = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    s.direction(ux, uy, uz);
    vx = v * ux;
    vy = v * uy;
    vz = v * uz;
}

void to_com_frame(int64_t n, const double* m, double* x, double* y, double* z, double* vx,
                  double* vy, double* vz) {
    long double M = 0, c[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        M += m[i];
        c[0] += (long double)m[i] * x[i];
        c[1] += (long double)m[i] * y[i];
        c[2] += (long double)m[i] * z[i];
        c[3] += (long double)m[i] * vx[i];
        c[4] += (long double)m[i] * vy[i];
        c[5] += (long double)m[i] * vz[i];
    }
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = (double)(c[k] / M);
    for (int64_t i = 0; i < n; ++i) {
        x[i] -= d[0];
        y[i] -= d[1];
        z[i] -= d[2];
        vx[i] -= d[3];
        vy[i] -= d[4];
        vz[i] -= d[5];
    }
}

bool bad(int64_t n, const void* const* ptrs, int k) {
    if (n < 0) return true;
    for (int i = 0; i < k; ++i)
        if (!ptrs[i]) return true;
    return false;
}

}  // namespace

extern "C" {

void nb_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

int nb_ic_plummer(int64_t n, uint64_t seed, double a, double M, double G, double rmax, double* m,
                  double* x, double* y, double* z, double* vx, double* vy, double* vz) {
    const void* p[7] = {m, x, y, z, vx, vy, vz};
    if (bad(n, p, 7) || !(a > 0) || !(M > 0) || !(G > 0)) {
        nb::set_error("nb_ic_plummer: invalid arguments");
        return NB_ERR_INVALID;
    }
    const double vscale = std::sqrt(G * M / a);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Stream s(seed, (uint64_t)i, 0);
        double px, py, pz, ux, uy, uz;
        plummer_particle(s, rmax, px, py, pz, ux, uy, uz);
        m[i] = M / (double)n;
        x[i] = a * px;
        y[i] = a * py;
        z[i] = a * pz;
        vx[i] = vscale * ux;
        vy[i] = vscale * uy;
        vz[i] = vscale * uz;
    }
    to_com_frame(n, m, x, y, z, vx, vy, vz);
    return NB_OK;
}

int nb_ic_uniform_gaussian(int64_t n, uint64_t seed, double sigma, double* m, double* x, double* y,
                           double* z, double* vx, double* vy, double* vz) {
    const void* p[7] = {m, x, y, z, vx, vy, vz};
    if (bad(n, p, 7) || !(sigma >= 0)) {
        nb::set_error("nb_ic_uniform_gaussian: invalid arguments");
        return NB_ERR_INVALID;
    }
    for (int64_t i = 0; i < n; ++i) {
        Stream s(seed, (uint64_t)i, 1);
        m[i] = 1.0 / (double)n;
        x[i] = 2.0 * s.uniform() - 1.0;
        y[i] = 2.0 * s.uniform() - 1.0;
        z[i] = 2.0 * s.uniform() - 1.0;
        vx[i] = sigma * s.normal();
        vy[i] = sigma * s.normal();
        vz[i] = sigma * s.normal();
    }
    return NB_OK;
}

int nb_ic_plummer_disk(int64_t n, uint64_t seed, double a, double M, double G, double disk_frac,
                       double rd, double zd, double* m, double* x, double* y, double* z,
                       double* vx, double* vy, double* vz) {
    const void* p[7] = {m, x, y, z, vx, vy, vz};
    if (bad(n, p, 7) || !(a > 0) || !(M > 0) || !(G > 0) || !(disk_frac >= 0 && disk_frac < 1) ||
        !(rd > 0) || !(zd >= 0)) {
        nb::set_error("nb_ic_plummer_disk: invalid arguments");
        return NB_ERR_INVALID;
    }
    const int64_t nd = (int64_t)std::llround(disk_frac * (double)n);
    const int64_t nb_ = n - nd;
    const double Mb = M * (1.0 - disk_frac), Md = M * disk_frac;
    const double vscale = std::sqrt(G * Mb / a);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Stream s(seed, (uint64_t)i, 2);
        if (i < nb_) {
            double px, py, pz, ux, uy, uz;
            plummer_particle(s, 0.0, px, py, pz, ux, uy, uz);
            m[i] = Mb / (double)nb_;
            x[i] = a * px;
            y[i] = a * py;
            z[i] = a * pz;
            vx[i] = vscale * ux;
            vy[i] = vscale * uy;
            vz[i] = vscale * uz;
        } else {
            // Surface density ~ exp(-R/rd): R ~ Gamma(2, rd); sech^2(z/zd) vertically.
            const double R = -rd * std::log(s.uniform() * s.uniform());
            const double ph = 2.0 * M_PI * s.uniform();
            const double u = s.uniform();
            const double h = zd * std::atanh(2.0 * u - 1.0);
            m[i] = Md / (double)nd;
            x[i] = R * std::cos(ph);
            y[i] = R * std::sin(ph);
            z[i] = h;
            // Circular speed from the mass enclosed in the sphere of radius R:
            // Plummer M(<R) plus the disk mass inside R.
            const double Mp = Mb * R * R * R / std::pow(R * R + a * a, 1.5);
            const double Mdisk = Md * (1.0 - (1.0 + R / rd) * std::exp(-R / rd));
            const double vc = R > 0 ? std::sqrt(G * (Mp + Mdisk) / R) : 0.0;
            vx[i] = -vc * std::sin(ph);
            vy[i] = vc * std::cos(ph);
            vz[i] = 0.0;
        }
    }
    to_com_frame(n, m, x, y, z, vx, vy, vz);
    return NB_OK;
}

// The reference's fixture stream: std::mt19937 with libstdc++'s
// uniform_real_distribution<double> (the same engine and distribution the
// reference's generator and test mains use, so the draws are the same bits).
int nb_ic_reference_fixture(int64_t n, uint32_t seed, int order, int64_t skip, double* m,
                            double* x, double* y, double* z) {
    if (n < 0 || skip < 0 || order < 0 || order > 2 || (n > 0 && (!m || !x || !y || !z))) {
        nb::set_error("nb_ic_reference_fixture: invalid arguments (n=%lld order=%d skip=%lld)",
                      (long long)n, order, (long long)skip);
        return NB_ERR_INVALID;
    }
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> pos(-100.0, 100.0);
    std::uniform_real_distribution<double> mass(0.5, 2.0);
    double dm, dx, dy, dz;
    auto draw = [&](double& pm, double& px, double& py, double& pz) {
        switch (order) {
            case 0:  // m, x, y, z
                pm = mass(rng);
                px = pos(rng);
                py = pos(rng);
                pz = pos(rng);
                break;
            case 1:  // z, y, x, m
                pz = pos(rng);
                py = pos(rng);
                px = pos(rng);
                pm = mass(rng);
                break;
            default:  // x, y, z, m
                px = pos(rng);
                py = pos(rng);
                pz = pos(rng);
                pm = mass(rng);
        }
    };
    for (int64_t i = 0; i < skip; ++i) draw(dm, dx, dy, dz);
    for (int64_t i = 0; i < n; ++i) draw(m[i], x[i], y[i], z[i]);
    return NB_OK;
}

int nb_file_read(const char* path, int64_t* n_out, double* m, double* x, double* y, double* z,
                 double* vx, double* vy, double* vz, int64_t capacity) {
    if (!path || !n_out) {
        nb::set_error("nb_file_read: null argument");
        return NB_ERR_INVALID;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        nb::set_error("nb_file_read: cannot open %s: %s", path, std::strerror(errno));
        return NB_ERR_INVALID;
    }
    long long n = -1;
    if (std::fscanf(f, "%lld", &n) != 1 || n < 0) {
        std::fclose(f);
        nb::set_error("nb_file_read: %s: missing particle count", path);
        return NB_ERR_INVALID;
    }
    *n_out = n;
    if (!m) {
        std::fclose(f);
        return NB_OK;
    }
    if (n > capacity || !x || !y || !z || !vx || !vy || !vz) {
        std::fclose(f);
        nb::set_error("nb_file_read: %s holds %lld particles, capacity %lld", path, n,
                      (long long)capacity);
        return NB_ERR_INVALID;
    }
    for (long long i = 0; i < n; ++i) {
        if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf", &m[i], &x[i], &y[i], &z[i], &vx[i],
                        &vy[i], &vz[i]) != 7) {
            std::fclose(f);
            nb::set_error("nb_file_read: %s: malformed line %lld", path, i + 2);
            return NB_ERR_INVALID;
        }
    }
    std::fclose(f);
    return NB_OK;
}

int nb_file_write(const char* path, int64_t n, const double* m, const double* x, const double* y,
                  const double* z, const double* vx, const double* vy, const double* vz) {
    const void* p[7] = {m, x, y, z, vx, vy, vz};
    if (!path || bad(n, p, n > 0 ? 7 : 0)) {
        nb::set_error("nb_file_write: invalid arguments");
        return NB_ERR_INVALID;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        nb::set_error("nb_file_write: cannot open %s: %s", path, std::strerror(errno));
        return NB_ERR_INVALID;
    }
    std::fprintf(f, "%lld\n", (long long)n);
    for (int64_t i = 0; i < n; ++i)
        std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", m[i], x[i], y[i], z[i],
                     vx[i], vy[i], vz[i]);
    if (std::fclose(f) != 0) {
        nb::set_error("nb_file_write: write error on %s", path);
        return NB_ERR_INVALID;
    }
    return NB_OK;
}

}  // extern "C"
