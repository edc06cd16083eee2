// nbody.hpp -- C++ drop-in adapter: the reference's force-calculator and
// integrator classes, with both of its call signatures, implemented over the
// C ABI of libnbody_amd.so (include/nbody_amd.h).
//
//  * CPU-style API (CPU/include/force_calculator.hpp:13-31, CPU/include/integrator.hpp:5-26):
//      calculator.computeAccelerations(std::vector<Particle>&)
//      integrator.step(std::vector<Particle>&, ForceCalculator&, double dt)
//    The host vector is uploaded, the work runs on the GPU, results come back
//    into the vector (the reference's semantics: the call returns with the
//    vector updated).
//  * CUDA-style API (CUDA/include/force_calculator.hpp:14-30, CUDA/include/integrator.hpp:7-17,
//    CUDA/include/particle_data.hpp:6-73):
//      calculator.computeAccelerations(ParticleArrays&)
//      integrator.step(ForceCalculator&, ParticleArrays&, double dt)
//    State stays resident on the device; allocateParticles / copyParticlesToDevice /
//    copyParticlesToHost / freeParticles move it.
//
// Errors: a non-zero C-ABI status throws nbody::Error (the reference calls
// exit() or assert()).  G defaults to the reference's 4*pi^2
// (CPU/include/common.hpp:28); setG() changes it.
#pragma once

#include <climits>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbody_amd.h"

namespace constant {
constexpr double PI = 3.14159265358979323846;
constexpr double G = 4.0 * PI * PI;
}  // namespace constant

// Same layout and operator rounding as CPU/include/particle.hpp:5-93.
struct Vector3D {
    double x, y, z;
    Vector3D() : x(0), y(0), z(0) {}
    Vector3D(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    Vector3D normalize() const {
        const double n = norm();
        if (n == 0) return Vector3D(0, 0, 0);
        return Vector3D(x / n, y / n, z / n);
    }
    Vector3D operator+(const Vector3D& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector3D operator-(const Vector3D& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vector3D operator*(double s) const { return {x * s, y * s, z * s}; }
    friend Vector3D operator*(double s, const Vector3D& v) { return {v.x * s, v.y * s, v.z * s}; }
    Vector3D operator/(double s) const { return {x / s, y / s, z / s}; }
    Vector3D& operator+=(const Vector3D& o) {
        x += o.x;
        y += o.y;
        z += o.z;
        return *this;
    }
    double dot(const Vector3D& o) const { return x * o.x + y * o.y + z * o.z; }
    Vector3D cross(const Vector3D& o) const {
        return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
    }
};

struct Particle {
    double mass;
    Vector3D position, velocity, acceleration;
    std::string name;
    Particle() : mass(1) {}
    Particle(double m, const Vector3D& p, const Vector3D& v, const std::string& n = "")
        : mass(m), position(p), velocity(v), name(n) {}
    Particle(double m, const Vector3D& p, const Vector3D& v, const Vector3D& a,
             const std::string& n = "")
        : mass(m), position(p), velocity(v), acceleration(a), name(n) {}
    void info() const {
        if (name.size()) std::cout << "Particle: " << name << "\n";
        std::cout << "Mass: " << mass << "\n\n";
        std::cout << "Position:     (" << position.x << ", " << position.y << ", " << position.z
                  << ")\n";
        std::cout << "Velocity:     (" << velocity.x << ", " << velocity.y << ", " << velocity.z
                  << ")\n";
        std::cout << "Acceleration: (" << acceleration.x << ", " << acceleration.y << ", "
                  << acceleration.z << ")\n";
    }
};

// CPU/include/common.hpp:40-51
struct TimingStatistics {
    std::string name;
    double total_time = 0.0;
    size_t calls = 0;
    double average_time() const { return calls > 0 ? total_time / calls : 0.0; }
    void reset() {
        total_time = 0.0;
        calls = 0;
    }
};

// Device state with the reference's field names (ParticleArrays,
// CUDA/include/particle_data.hpp:6-43): FP64 device arrays owned by
// allocateParticles / freeParticles.  The reference's Barnes-Hut scratch
// arrays (Morton codes, sorted FP32 copies) live in the calculator's
// workspace here, and velocity Verlet keeps its previous acceleration inside
// the integrator, so the old_acc_* fields stay null.
struct ParticleArrays {
    int num_particles = 0;
    double* pos_x = nullptr;
    double* pos_y = nullptr;
    double* pos_z = nullptr;
    double* vel_x = nullptr;
    double* vel_y = nullptr;
    double* vel_z = nullptr;
    double* acc_x = nullptr;
    double* acc_y = nullptr;
    double* acc_z = nullptr;
    double* old_acc_x = nullptr;
    double* old_acc_y = nullptr;
    double* old_acc_z = nullptr;
    double* mass = nullptr;
};

namespace nbody {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* where) {
    if (rc != NB_OK) throw Error(rc, std::string(where) + ": " + nb_last_error());
}

// The C ABI's view of a ParticleArrays (same device arrays, no copy).
inline nb_particles view(const ParticleArrays& p) {
    nb_particles v{};
    v.n = p.num_particles;
    v.x = p.pos_x;
    v.y = p.pos_y;
    v.z = p.pos_z;
    v.vx = p.vel_x;
    v.vy = p.vel_y;
    v.vz = p.vel_z;
    v.ax = p.acc_x;
    v.ay = p.acc_y;
    v.az = p.acc_z;
    v.m = p.mass;
    return v;
}

// Host <-> device staging for the CPU-style API.
struct HostSoA {
    std::vector<double> x, y, z, vx, vy, vz, ax, ay, az, m;
    void from(const std::vector<Particle>& p) {
        const size_t n = p.size();
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &ax, &ay, &az, &m}) v->resize(n);
        for (size_t i = 0; i < n; ++i) {
            x[i] = p[i].position.x;
            y[i] = p[i].position.y;
            z[i] = p[i].position.z;
            vx[i] = p[i].velocity.x;
            vy[i] = p[i].velocity.y;
            vz[i] = p[i].velocity.z;
            m[i] = p[i].mass;
        }
    }
};

// Device mirror of a host particle vector, reused across calls.
struct DeviceMirror {
    nb_particles d{};
    HostSoA h;
    ~DeviceMirror() { nb_particles_free(&d); }
    void upload(const std::vector<Particle>& p) {
        if (d.n != (int64_t)p.size()) {
            check(nb_particles_free(&d), "freeParticles");
            check(nb_particles_alloc(&d, (int64_t)p.size()), "allocateParticles");
        }
        h.from(p);
        check(nb_particles_upload(&d, h.x.data(), h.y.data(), h.z.data(), h.vx.data(), h.vy.data(),
                                  h.vz.data(), h.m.data()),
              "copyParticlesToDevice");
    }
    void download(std::vector<Particle>& p, bool state) {
        check(nb_particles_download(&d, state ? h.x.data() : nullptr, state ? h.y.data() : nullptr,
                                    state ? h.z.data() : nullptr, state ? h.vx.data() : nullptr,
                                    state ? h.vy.data() : nullptr, state ? h.vz.data() : nullptr,
                                    h.ax.data(), h.ay.data(), h.az.data(), nullptr),
              "copyParticlesToHost");
        for (size_t i = 0; i < p.size(); ++i) {
            if (state) {
                p[i].position = Vector3D(h.x[i], h.y[i], h.z[i]);
                p[i].velocity = Vector3D(h.vx[i], h.vy[i], h.vz[i]);
            }
            p[i].acceleration = Vector3D(h.ax[i], h.ay[i], h.az[i]);
        }
    }
};

}  // namespace nbody

// CUDA/include/particle_data.hpp:45-73
inline void allocateParticles(ParticleArrays& p, size_t n) {
    if (n > (size_t)INT_MAX)  // num_particles is an int, as in the reference
        throw nbody::Error(NB_ERR_INVALID, "allocateParticles: too many particles");
    nb_particles v{};
    nbody::check(nb_particles_alloc(&v, (int64_t)n), "allocateParticles");
    p = ParticleArrays{};
    p.num_particles = (int)n;
    p.pos_x = v.x;
    p.pos_y = v.y;
    p.pos_z = v.z;
    p.vel_x = v.vx;
    p.vel_y = v.vy;
    p.vel_z = v.vz;
    p.acc_x = v.ax;
    p.acc_y = v.ay;
    p.acc_z = v.az;
    p.mass = v.m;
}
inline void freeParticles(ParticleArrays& p) {
    nb_particles v = nbody::view(p);
    nbody::check(nb_particles_free(&v), "freeParticles");
    p = ParticleArrays{};
}
inline void copyParticlesToDevice(ParticleArrays& p, const double* x, const double* y,
                                  const double* z, const double* vx, const double* vy,
                                  const double* vz, const double* m) {
    nb_particles v = nbody::view(p);
    nbody::check(nb_particles_upload(&v, x, y, z, vx, vy, vz, m), "copyParticlesToDevice");
}
inline void copyParticlesToHost(const ParticleArrays& p, double* x, double* y, double* z,
                                double* vx, double* vy, double* vz, double* ax, double* ay,
                                double* az, double* m) {
    const nb_particles v = nbody::view(p);
    nbody::check(nb_particles_download(&v, x, y, z, vx, vy, vz, ax, ay, az, m),
                 "copyParticlesToHost");
}

// ForceCalculator base: CPU/include/force_calculator.hpp:13-31 and
// CUDA/include/force_calculator.hpp:14-30 in one class.
class ForceCalculator {
   protected:
    double softening_squared;
    double softening_;
    double G_ = constant::G;
    nb_force* h_ = nullptr;
    nbody::DeviceMirror mirror_;
    mutable TimingStatistics stats_;

    void sync_stats() const {
        double t[4];
        if (h_ && nb_force_timing(h_, t) == NB_OK) {
            stats_.total_time = t[0];
            stats_.calls = (size_t)t[3];
        }
    }

   public:
    ForceCalculator() : softening_squared(0.0), softening_(0.0) {}
    explicit ForceCalculator(double softening)
        : softening_squared(softening * softening), softening_(softening) {}
    virtual ~ForceCalculator() { nb_force_destroy(h_); }
    ForceCalculator(const ForceCalculator&) = delete;
    ForceCalculator& operator=(const ForceCalculator&) = delete;

    virtual void computeAccelerations(std::vector<Particle>& particles) {
        mirror_.upload(particles);
        nbody::check(nb_force_compute(h_, &mirror_.d), "computeAccelerations");
        mirror_.download(particles, false);
        checkDevice();
    }
    virtual void computeAccelerations(ParticleArrays& particles) {
        nb_particles v = nbody::view(particles);
        nbody::check(nb_force_compute(h_, &v), "computeAccelerations");
    }
    // Throws if a check that ran on the device failed in the last evaluation
    // (the synchronous std::vector calls run it after downloading).
    virtual void checkDevice() const {}
    virtual std::string name() const { return nb_force_name(h_); }
    virtual const TimingStatistics& getTimeStatistics() const {
        sync_stats();
        return stats_;
    }
    nb_force* handle() const { return h_; }
};

// ForceCalculatorNaive(double softening): CPU/include/force_calculator_naive.hpp:12
// (FP64, bit-identical to the reference CPU) or the FP32 tiled kernel.
class ForceCalculatorNaive : public ForceCalculator {
   public:
    explicit ForceCalculatorNaive(double softening = 0.0, int precision = NB_DIRECT_F64_EXACT,
                                  double G = constant::G)
        : ForceCalculator(softening) {
        G_ = G;
        nbody::check(nb_force_create_direct(&h_, precision, G, softening), "ForceCalculatorNaive");
        nb_force_set_timing(h_, 1);
        stats_.name = "Naive Force Calculation";
    }
};

// BarnesHutForceCalculator(double softening, double theta, int leaf_capacity = 8):
// CUDA/include/force_calculator_barnes_hut.hpp:56.
class BarnesHutForceCalculator : public ForceCalculator {
   public:
    BarnesHutForceCalculator(double softening, double theta, int leaf_capacity = 8,
                             int multipole = NB_QUADRUPOLE, double G = constant::G)
        : ForceCalculator(softening) {
        G_ = G;
        nbody::check(nb_force_create_bh(&h_, G, softening, theta, leaf_capacity, multipole),
                     "BarnesHutForceCalculator");
        nb_force_set_timing(h_, 2);  // build + walk: the getters below need no more
        stats_.name = "Barnes-Hut Force Calculation";
    }
    nb_bh_stats stats() const {
        nb_bh_stats s{};
        nbody::check(nb_force_bh_stats(h_, &s), "BarnesHutForceCalculator::stats");
        return s;
    }
    // Walk frontier, node pool, tree consistency and sort checks of the last
    // evaluation (nb_force_bh_stats fails if any of them did; NB_ERR_STATE
    // only means nothing was evaluated yet, e.g. an empty particle set).
    void checkDevice() const override {
        nb_bh_stats s{};
        const int rc = nb_force_bh_stats(h_, &s);
        if (rc != NB_ERR_STATE) nbody::check(rc, "Barnes-Hut device checks");
    }
    // CUDA/include/force_calculator_barnes_hut.hpp:120-125
    double getTotalComputeTime() const {
        double t[4];
        nbody::check(nb_force_timing(h_, t), "timing");
        return t[0];
    }
    double getBuildTreeTime() const {
        double t[4];
        nbody::check(nb_force_timing(h_, t), "timing");
        return t[1];
    }
    double getTraversalTime() const {
        double t[4];
        nbody::check(nb_force_timing(h_, t), "timing");
        return t[2];
    }
    void resetTimeStatistics() { nb_force_reset_timing(h_); }
};

// ForceCalculatorBarnesHutOptimized(double softening, double theta, int leaf_capacity):
// CPU/include/force_calculator_barnes_hut_optimized.hpp:40.  Same tree as the
// reference CPU; forces from the GPU group walk.
class ForceCalculatorBarnesHutOptimized : public BarnesHutForceCalculator {
   public:
    ForceCalculatorBarnesHutOptimized(double softening, double theta, int leaf_capacity)
        : BarnesHutForceCalculator(softening, theta, leaf_capacity) {}
};

// ForceCalculatorBarnesHut() / (double softening, double theta):
// CPU/include/force_calculator_barnes_hut.hpp:33-34, the recursive build
// (CPU/src/force_calculator_barnes_hut.cpp:97-173): a leaf splits when a second
// particle arrives unless its half size is below 1e-9, i.e. the same tree
// rule with leaf capacity 1.  The three statistics getters of :51-59 report
// the GPU tree build, the walk, and no destruction (the device tree is reused).
class ForceCalculatorBarnesHut : public BarnesHutForceCalculator {
    mutable TimingStatistics building_{"Barnes-Hut Tree Building"}, traversal_{"Barnes-Hut Tree Traversal"},
        destroying_{"Barnes-Hut Tree Destruction"};
    void sync_parts() const {
        double t[4];
        if (nb_force_timing(h_, t) == NB_OK) {
            building_.total_time = t[1];
            traversal_.total_time = t[2];
            building_.calls = traversal_.calls = destroying_.calls = (size_t)t[3];
        }
    }

   public:
    ForceCalculatorBarnesHut() : ForceCalculatorBarnesHut(0.0, 0.5) {}
    ForceCalculatorBarnesHut(double softening, double theta)
        : BarnesHutForceCalculator(softening, theta, 1) {}
    TimingStatistics& getBuildingTreeStatistics() const {
        sync_parts();
        return building_;
    }
    TimingStatistics& getTraversalStatistics() const {
        sync_parts();
        return traversal_;
    }
    TimingStatistics& getDestroyingTreeStatistics() const {
        sync_parts();
        return destroying_;
    }
};

// TimeIntegrator base with both step() signatures
// (CPU/include/integrator.hpp:16-20, CUDA/include/integrator.hpp:10-14).
class TimeIntegrator {
   protected:
    nb_integrator* h_ = nullptr;
    nbody::DeviceMirror mirror_;

   public:
    explicit TimeIntegrator(int kind) {
        nbody::check(nb_integrator_create(&h_, kind), "TimeIntegrator");
    }
    virtual ~TimeIntegrator() { nb_integrator_destroy(h_); }
    TimeIntegrator(const TimeIntegrator&) = delete;
    TimeIntegrator& operator=(const TimeIntegrator&) = delete;

    virtual void step(std::vector<Particle>& particles, ForceCalculator& calculator, double dt) {
        mirror_.upload(particles);
        nb_integrator_reset(h_);  // host state may have changed since the last call
        nbody::check(nb_integrator_step(h_, calculator.handle(), &mirror_.d, dt), "step");
        mirror_.download(particles, true);
        calculator.checkDevice();
    }
    virtual void step(ForceCalculator& calculator, ParticleArrays& particles, double dt) {
        nb_particles v = nbody::view(particles);
        nbody::check(nb_integrator_step(h_, calculator.handle(), &v, dt), "step");
    }
    // Several steps without host round trips (device state).
    void run(ForceCalculator& calculator, ParticleArrays& particles, double dt, int steps) {
        nb_particles v = nbody::view(particles);
        nbody::check(nb_integrator_run(h_, calculator.handle(), &v, dt, steps), "run");
    }
    virtual std::string name() const { return nb_integrator_name(h_); }
};

// CPU/include/integrator_leapfrog.hpp, CUDA/include/integrator_leapfrog.hpp
class LeapfrogIntegrator : public TimeIntegrator {
   public:
    LeapfrogIntegrator() : TimeIntegrator(NB_LEAPFROG) {}
};

// CPU/include/integrator_rk4.hpp
class RK4Integrator : public TimeIntegrator {
   public:
    RK4Integrator() : TimeIntegrator(NB_RK4) {}
};

// CPU/include/integrator_euler.hpp, CUDA/include/integrator_euler.hpp
class EulerIntegrator : public TimeIntegrator {
   public:
    EulerIntegrator() : TimeIntegrator(NB_EULER) {}
};

// CPU/include/integrator_euler-cromer.hpp, CUDA/include/integrator_euler-cromer.hpp
class EulerCromerIntegrator : public TimeIntegrator {
   public:
    EulerCromerIntegrator() : TimeIntegrator(NB_EULER_CROMER) {}
};

// CPU/include/integrator_velocity_verlet.hpp, CUDA/include/integrator_velocity_verlet.hpp
class VelocityVerletIntegrator : public TimeIntegrator {
   public:
    VelocityVerletIntegrator() : TimeIntegrator(NB_VELOCITY_VERLET) {}
};

// ------------------------------------------------------------------ test helpers
// runSimulation and the diagnostics of CPU/include/tests_common.hpp:19-101,
// with the same names and meaning.  The sums run on the GPU in FP64 with
// compensated (Neumaier) accumulation; the potential is the reference's
// unsoftened pair sum with G = constant::G.
template <typename Integrator, typename Calculator>
std::vector<Particle> runSimulation(std::vector<Particle> particles, Integrator& integrator,
                                    Calculator& calculator, double dt, int steps) {
    for (int i = 0; i < steps; i++) integrator.step(particles, calculator, dt);
    return particles;
}

namespace nbody {
inline nb_diag diagnostics(const std::vector<Particle>& particles, bool with_potential,
                           double G = constant::G) {
    DeviceMirror mirror;
    mirror.upload(particles);
    nb_diag d{};
    check(nb_diagnostics(&mirror.d, G, 0.0, with_potential ? 1 : 0, &d), "diagnostics");
    return d;
}
}  // namespace nbody

inline double kineticEnergy(const std::vector<Particle>& particles) {
    return nbody::diagnostics(particles, false).ke;
}
inline double potentialEnergy(const std::vector<Particle>& particles) {
    return nbody::diagnostics(particles, true).pe;
}
inline double totalEnergy(const std::vector<Particle>& particles) {
    const nb_diag d = nbody::diagnostics(particles, true);
    return d.ke + d.pe;
}
inline Vector3D angularMomentum(const std::vector<Particle>& particles) {
    const nb_diag d = nbody::diagnostics(particles, false);
    return Vector3D(d.l[0], d.l[1], d.l[2]);
}
inline Vector3D centerOfMass(const std::vector<Particle>& particles) {
    const nb_diag d = nbody::diagnostics(particles, false);
    return Vector3D(d.com[0], d.com[1], d.com[2]);
}
inline double positionError(const std::vector<Particle>& a, const std::vector<Particle>& b) {
    double err = 0.0;
    for (size_t i = 0; i < a.size(); i++) err += (a[i].position - b[i].position).norm();
    return err / a.size();
}
inline double velocityError(const std::vector<Particle>& a, const std::vector<Particle>& b) {
    double err = 0.0;
    for (size_t i = 0; i < a.size(); i++) err += (a[i].velocity - b[i].velocity).norm();
    return err / a.size();
}

namespace nbody {
// Trajectory rows in the column layout of the reference's orbit.csv
// (CPU/src/n-body.cpp:6,19-58): the positions of two bodies (`body`, then
// `center`), kinetic, potential and total energy, angular momentum and centre
// of mass, printed fixed with 15 decimals.  The reference prints its first row
// with the stream's default 6 significant digits; every row here uses 15
// decimals.  Diagnostics are reduced on the GPU from device-resident state.
class OrbitCsvWriter {
   public:
    OrbitCsvWriter(std::ostream& out, int64_t body, int64_t center, double G = constant::G)
        : out_(out), body_(body), center_(center), G_(G) {
        out_ << "earth_x,earth_y,earth_z,sun_x,sun_y,sun_z,kinetic_energy,potential_energy,"
                "total_energy,angular_momentum_x,angular_momentum_y,angular_momentum_z,"
                "center_of_mass_x,center_of_mass_y,center_of_mass_z\n";
    }
    void row(const ParticleArrays& pa) {
        const nb_particles p = view(pa);
        std::vector<double> x(p.n), y(p.n), z(p.n);
        check(nb_particles_download(&p, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr),
              "OrbitCsvWriter");
        nb_diag d{};
        check(nb_diagnostics(&p, G_, 0.0, 1, &d), "OrbitCsvWriter");
        const double v[15] = {x[body_], y[body_], z[body_], x[center_], y[center_], z[center_],
                              d.ke, d.pe, d.ke + d.pe, d.l[0], d.l[1], d.l[2],
                              d.com[0], d.com[1], d.com[2]};
        out_ << std::fixed << std::setprecision(15);
        for (int k = 0; k < 15; ++k) out_ << v[k] << (k < 14 ? ',' : '\n');
    }

   private:
    std::ostream& out_;
    int64_t body_, center_;
    double G_;
};
}  // namespace nbody
