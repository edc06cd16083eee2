// nbody_oracle.cpp -- CPU restatement of the reference N-body force and
// integration path, used ONLY as the checker by tests/, by smoke() and by the
// cpu_baseline leg of bench.py.  Nothing in the product links or calls it.
//
// Every function states the reference file:line range whose arithmetic it
// follows.  FP64 results are meant to be bit-identical to the reference CPU
// build (g++ -O3, x86-64 baseline: SSE2 scalar double, no FMA contraction);
// this file is therefore compiled with -ffp-contract=off.
//
// What pins each part (checked in tests/test_oracle_*.py):
//   direct / leapfrog / rk4 /    bit-exact hashes from the compiled reference
//   euler / euler-cromer / vv
//                                 (tests/golden/ref_golden.json) and the 200
//                                 error lines of CPU/outputs/output2_4.txt
//   bhopt (topology + monopole)   compiled-reference hashes + output2_4.txt
//   bhrec (recursive, leaf cap 1) compiled-reference hashes + output2_4.txt
//   keys63 / root                 octant paths of the compiled reference's
//                                 insertion tree (leaf path == key prefix)
//   diag                          long-double sums vs closed forms / brute force
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

namespace {

// Three-vector with the reference's operator rounding order
// (CPU/include/particle.hpp:5-62): every operator is component-wise and
// dot() is ((x*x + y*y) + z*z).
struct V3 {
    double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 divs(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct SoA {
    int64_t n;
    const double *m, *x, *y, *z;
    V3 pos(int64_t i) const { return V3{x[i], y[i], z[i]}; }
};

// ---------------------------------------------------------------------------
// Direct sum.  CPU/src/force_calculator_naive.cpp:20-30:
//   r = p_j - p_i; d2 = dot(r,r) + eps2; inv = 1/sqrt(d2);
//   inv3 = (inv*inv)*inv; a_i += r * ((G*m_j)*inv3), j ascending, j != i.
// ---------------------------------------------------------------------------
V3 direct_row(const SoA& s, int64_t i, double G, double soft2) {
    V3 pi = s.pos(i);
    V3 acc{0.0, 0.0, 0.0};
    for (int64_t j = 0; j < s.n; ++j) {
        if (j == i) continue;
        V3 r = sub(s.pos(j), pi);
        double d2 = dot(r, r) + soft2;
        double inv = 1.0 / std::sqrt(d2);
        double inv3 = inv * inv * inv;
        acc = add(acc, scale(r, G * s.m[j] * inv3));
    }
    return acc;
}

void direct_into(const SoA& s, double G, double soft2, double* ax, double* ay, double* az) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < s.n; ++i) {
        const V3 acc = direct_row(s, i, G, soft2);
        ax[i] = acc.x;
        ay[i] = acc.y;
        az[i] = acc.z;
    }
}

// ---------------------------------------------------------------------------
// Barnes-Hut with a node pool, contiguous 8-child blocks and leaf capacity k.
// CPU/src/force_calculator_barnes_hut_optimized.cpp:48-231.  The walk stack is
// a growable vector instead of the reference's int[64]
// (force_calculator_barnes_hut_optimized.cpp:52), which overflows on deep
// (Plummer, N >= ~2^20) trees; traversal order is unchanged (LIFO, children
// scanned 0..7, accepted cells added immediately).
// ---------------------------------------------------------------------------
struct PoolNode {
    V3 center{0, 0, 0};
    double size = 0.0, size_sqr = 0.0, mass = 0.0;
    V3 com{0, 0, 0};
    int64_t first_child = -1;
    std::vector<int64_t> members;
    int depth = 0;
};

struct PoolTree {
    const SoA* s = nullptr;
    int cap = 8;
    double split_eps = 1e-9;  // force_calculator_barnes_hut_optimized.hpp:35
    std::vector<PoolNode> pool;
    int64_t root = -1;

    static int octant(V3 p, V3 c) {
        int k = 0;
        if (p.x >= c.x) k |= 1;
        if (p.y >= c.y) k |= 2;
        if (p.z >= c.z) k |= 4;
        return k;
    }

    int64_t alloc() {
        pool.emplace_back();
        return (int64_t)pool.size() - 1;
    }

    // force_calculator_barnes_hut_optimized.cpp:206-231
    bool split(int64_t ni) {
        double half = pool[ni].size / 2.0;
        if (half < split_eps) return false;
        int64_t first = (int64_t)pool.size();
        V3 c = pool[ni].center;
        int d = pool[ni].depth + 1;
        for (int k = 0; k < 8; ++k) {
            int64_t ci = alloc();
            PoolNode& ch = pool[ci];
            ch.size = half;
            ch.size_sqr = half * half;
            double ox = (k & 1) ? half / 2.0 : -half / 2.0;
            double oy = (k & 2) ? half / 2.0 : -half / 2.0;
            double oz = (k & 4) ? half / 2.0 : -half / 2.0;
            ch.center = V3{c.x + ox, c.y + oy, c.z + oz};
            ch.depth = d;
        }
        pool[ni].first_child = first;
        return true;
    }

    // force_calculator_barnes_hut_optimized.cpp:131-185.  Mass and COM of
    // every visited node are updated before descending, in arrival order.
    void insert(int64_t ni, int64_t pi) {
        const double mp = s->m[pi];
        const V3 pp = s->pos(pi);
        for (;;) {
            {
                PoolNode& nd = pool[ni];
                nd.mass += mp;
                nd.com = divs(add(scale(nd.com, nd.mass - mp), scale(pp, mp)), nd.mass);
            }
            if (pool[ni].first_child == -1) {
                if ((int64_t)pool[ni].members.size() < cap) {
                    pool[ni].members.push_back(pi);
                    return;
                }
                if (!split(ni)) {
                    pool[ni].members.push_back(pi);
                    return;
                }
                std::vector<int64_t> old;
                old.swap(pool[ni].members);
                V3 c = pool[ni].center;
                int64_t fc = pool[ni].first_child;
                for (int64_t q : old) insert(fc + octant(s->pos(q), c), q);
                ni = fc + octant(pp, c);
                // The new particle is inserted into its child with a fresh
                // mass/COM update, exactly like the reference's recursive call.
                continue;
            }
            ni = pool[ni].first_child + octant(pp, pool[ni].center);
        }
    }

    // force_calculator_barnes_hut_optimized.cpp:99-129
    void build(const SoA& soa, int leaf_cap) {
        s = &soa;
        cap = leaf_cap;
        pool.clear();
        root = -1;
        if (soa.n == 0) return;
        double mnx, mxx, mny, mxy, mnz, mxz;
        mnx = mny = mnz = std::numeric_limits<double>::max();
        mxx = mxy = mxz = std::numeric_limits<double>::lowest();
        for (int64_t i = 0; i < soa.n; ++i) {
            mnx = std::min(mnx, soa.x[i]);
            mxx = std::max(mxx, soa.x[i]);
            mny = std::min(mny, soa.y[i]);
            mxy = std::max(mxy, soa.y[i]);
            mnz = std::min(mnz, soa.z[i]);
            mxz = std::max(mxz, soa.z[i]);
        }
        double size = std::max(mxx - mnx, std::max(mxy - mny, mxz - mnz));
        size += split_eps;
        root = alloc();
        pool[root].center = V3{(mnx + mxx) / 2, (mny + mxy) / 2, (mnz + mxz) / 2};
        pool[root].size = size;
        pool[root].size_sqr = size * size;
        for (int64_t i = 0; i < soa.n; ++i) insert(root, i);
    }

    // force_calculator_barnes_hut_optimized.cpp:48-98
    V3 walk(int64_t pi, double G, double soft2, double theta_sqr, int* max_stack) const {
        V3 acc{0.0, 0.0, 0.0};
        if (root < 0) return acc;
        const V3 p = s->pos(pi);
        std::vector<int64_t> stack;
        stack.reserve(256);
        stack.push_back(root);
        int hi = 1;
        while (!stack.empty()) {
            int64_t ni = stack.back();
            stack.pop_back();
            const PoolNode& nd = pool[ni];
            if (nd.first_child == -1) {
                for (int64_t q : nd.members) {
                    if (q == pi) continue;
                    V3 r = sub(s->pos(q), p);
                    double d2 = dot(r, r) + soft2;
                    double inv = 1.0 / std::sqrt(d2);
                    double inv3 = inv * inv * inv;
                    acc = add(acc, scale(scale(r, G), s->m[q] * inv3));
                }
                continue;
            }
            int own = octant(p, nd.center);
            for (int k = 0; k < 8; ++k) {
                int64_t ci = nd.first_child + k;
                if (k == own) {
                    stack.push_back(ci);
                    continue;
                }
                const PoolNode& ch = pool[ci];
                V3 r = sub(ch.com, p);
                double d2 = dot(r, r);
                if (ch.size_sqr < theta_sqr * d2) {
                    double inv = 1.0 / std::sqrt(d2 + soft2);
                    double inv3 = inv * inv * inv;
                    acc = add(acc, scale(scale(r, G), ch.mass * inv3));
                } else {
                    stack.push_back(ci);
                }
            }
            hi = std::max(hi, (int)stack.size());
        }
        if (max_stack) *max_stack = hi;
        return acc;
    }
};

// ---------------------------------------------------------------------------
// Recursive pointer-tree Barnes-Hut with one particle per leaf.
// CPU/src/force_calculator_barnes_hut.cpp:47-209.  Nodes live in a vector and
// are addressed by index; arithmetic and visiting order follow the reference.
// ---------------------------------------------------------------------------
struct RecNode {
    V3 center{0, 0, 0};
    double size = 0.0, size_sqr = 0.0, mass = 0.0;
    V3 com{0, 0, 0};
    int64_t child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    std::vector<int64_t> members;
    bool leaf = true;
};

struct RecTree {
    const SoA* s = nullptr;
    double split_eps = 1e-9;
    std::vector<RecNode> nodes;
    int64_t root = -1;

    static int octant(V3 p, V3 c) { return PoolTree::octant(p, c); }

    bool split(int64_t ni) {
        double half = nodes[ni].size / 2.0;
        if (half < split_eps) return false;
        V3 c = nodes[ni].center;
        for (int k = 0; k < 8; ++k) {
            RecNode ch;
            ch.size = half;
            ch.size_sqr = half * half;
            ch.leaf = true;
            double ox = (k & 1) ? half / 2.0 : -half / 2.0;
            double oy = (k & 2) ? half / 2.0 : -half / 2.0;
            double oz = (k & 4) ? half / 2.0 : -half / 2.0;
            ch.center = V3{c.x + ox, c.y + oy, c.z + oz};
            nodes.push_back(ch);
            nodes[ni].child[k] = (int64_t)nodes.size() - 1;
        }
        return true;
    }

    // force_calculator_barnes_hut.cpp:133-174
    void insert(int64_t ni, int64_t pi) {
        const double mp = s->m[pi];
        const V3 pp = s->pos(pi);
        if (nodes[ni].leaf && nodes[ni].members.empty()) {
            nodes[ni].members.push_back(pi);
            nodes[ni].mass = mp;
            nodes[ni].com = pp;
            return;
        }
        if (nodes[ni].leaf) {
            if (!split(ni)) {
                RecNode& nd = nodes[ni];
                nd.mass += mp;
                nd.com = divs(add(scale(nd.com, nd.mass - mp), scale(pp, mp)), nd.mass);
                nd.members.push_back(pi);
                return;
            }
            int64_t q = nodes[ni].members.front();
            nodes[ni].leaf = false;
            nodes[ni].members.clear();
            V3 c = nodes[ni].center;
            insert(nodes[ni].child[octant(s->pos(q), c)], q);
            insert(nodes[ni].child[octant(pp, c)], pi);
            RecNode& nd = nodes[ni];
            nd.mass = s->m[q] + mp;
            nd.com = divs(add(scale(s->pos(q), s->m[q]), scale(pp, mp)), nd.mass);
            return;
        }
        insert(nodes[ni].child[octant(pp, nodes[ni].center)], pi);
        RecNode& nd = nodes[ni];
        nd.mass += mp;
        nd.com = divs(add(scale(nd.com, nd.mass - mp), scale(pp, mp)), nd.mass);
    }

    // force_calculator_barnes_hut.cpp:94-131
    void build(const SoA& soa) {
        s = &soa;
        nodes.clear();
        root = -1;
        if (soa.n == 0) return;
        double mnx, mxx, mny, mxy, mnz, mxz;
        mnx = mny = mnz = std::numeric_limits<double>::max();
        mxx = mxy = mxz = std::numeric_limits<double>::lowest();
        for (int64_t i = 0; i < soa.n; ++i) {
            mnx = std::min(mnx, soa.x[i]);
            mxx = std::max(mxx, soa.x[i]);
            mny = std::min(mny, soa.y[i]);
            mxy = std::max(mxy, soa.y[i]);
            mnz = std::min(mnz, soa.z[i]);
            mxz = std::max(mxz, soa.z[i]);
        }
        double size = std::max(mxx - mnx, std::max(mxy - mny, mxz - mnz));
        size += split_eps;
        RecNode r;
        r.center = V3{(mnx + mxx) / 2, (mny + mxy) / 2, (mnz + mxz) / 2};
        r.size = size;
        r.size_sqr = size * size;
        r.leaf = true;
        nodes.push_back(r);
        root = 0;
        for (int64_t i = 0; i < soa.n; ++i) insert(root, i);
    }

    // force_calculator_barnes_hut.cpp:47-92
    void walk(int64_t ni, int64_t pi, V3 p, double G, double soft2, double theta_sqr, V3& acc) const {
        const RecNode& nd = nodes[ni];
        if (nd.leaf) {
            for (int64_t q : nd.members) {
                if (q == pi) continue;
                V3 r = sub(s->pos(q), p);
                double d2 = dot(r, r) + soft2;
                double inv = 1.0 / std::sqrt(d2);
                double inv3 = inv * inv * inv;
                acc = add(acc, scale(scale(r, G), s->m[q] * inv3));
            }
            return;
        }
        int own = octant(p, nd.center);
        for (int k = 0; k < 8; ++k) {
            int64_t ci = nd.child[k];
            if (ci < 0) continue;
            if (k == own) {
                walk(ci, pi, p, G, soft2, theta_sqr, acc);
                continue;
            }
            const RecNode& ch = nodes[ci];
            V3 r = sub(ch.com, p);
            double d2 = dot(r, r);
            if (ch.size_sqr < theta_sqr * d2) {
                double inv = 1.0 / std::sqrt(d2 + soft2);
                double inv3 = inv * inv * inv;
                acc = add(acc, scale(scale(r, G), ch.mass * inv3));
            } else {
                walk(ci, pi, p, G, soft2, theta_sqr, acc);
            }
        }
    }
};

// Kick/drift helpers with the reference's element-wise update order.
inline void kick(int64_t n, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                 const double* az, double h) {
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = vx[i] + ax[i] * h;
        vy[i] = vy[i] + ay[i] * h;
        vz[i] = vz[i] + az[i] * h;
    }
}
inline void drift(int64_t n, double* x, double* y, double* z, const double* vx, const double* vy,
                  const double* vz, double dt) {
    for (int64_t i = 0; i < n; ++i) {
        x[i] = x[i] + vx[i] * dt;
        y[i] = y[i] + vy[i] * dt;
        z[i] = z[i] + vz[i] * dt;
    }
}

// Long-double Neumaier-free accumulation is enough for the diagnostics checks
// (80-bit extended on x86-64: 64-bit mantissa).
typedef long double LD;

}  // namespace

extern "C" {

// CPU/src/force_calculator_naive.cpp:3-35
void orc_direct(int64_t n, double G, double soft2, const double* m, const double* x, const double* y,
                const double* z, double* ax, double* ay, double* az) {
    SoA s{n, m, x, y, z};
    direct_into(s, G, soft2, ax, ay, az);
}

// The direct sum for the targets idx[0..k) only (same arithmetic per target).
void orc_direct_subset(int64_t n, double G, double soft2, const double* m, const double* x,
                       const double* y, const double* z, int64_t k, const int64_t* idx,
                       double* ax, double* ay, double* az) {
    SoA s{n, m, x, y, z};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < k; ++t) {
        const V3 acc = direct_row(s, idx[t], G, soft2);
        ax[t] = acc.x;
        ay[t] = acc.y;
        az[t] = acc.z;
    }
}

// Kick-drift-kick leapfrog with the direct force.
// two_evals=1 follows CPU/src/integrator_leapfrog.cpp:3-26 (force evaluated at
// the start and at the end of every step); two_evals=0 follows the cached form
// of CUDA/src/integrator_leafrog.cu:47-103 (first force evaluated once, then
// one evaluation per step).  Both produce the same bits because the force is a
// pure function of the positions.  On return a* holds the last force.
void orc_leapfrog_direct(int64_t n, double G, double soft2, double dt, int steps, int two_evals,
                         const double* m, double* x, double* y, double* z, double* vx, double* vy,
                         double* vz, double* ax, double* ay, double* az) {
    SoA s{n, m, x, y, z};
    const double h = dt / 2.0;
    if (!two_evals && steps > 0) direct_into(s, G, soft2, ax, ay, az);
    for (int k = 0; k < steps; ++k) {
        if (two_evals) direct_into(s, G, soft2, ax, ay, az);
        kick(n, vx, vy, vz, ax, ay, az, h);
        drift(n, x, y, z, vx, vy, vz, dt);
        direct_into(s, G, soft2, ax, ay, az);
        kick(n, vx, vy, vz, ax, ay, az, h);
    }
}

// Classic RK4 with the direct force.  CPU/src/integrator_rk4.cpp:3-47:
//   a1=f(x); x2=x+v*(dt/2), v2=v+a1*(dt/2); a2=f(x2);
//   x3=x+v2*(dt/2), v3=v+a2*(dt/2); a3=f(x3);
//   x4=x+v3*dt, v4=v+a3*dt; a4=f(x4);
//   x += (((v+2*v2)+2*v3)+v4)*(dt/6); v += (((a1+2*a2)+2*a3)+a4)*(dt/6).
// On return a* holds a1 of the last step (the reference leaves it there too).
void orc_rk4_direct(int64_t n, double G, double soft2, double dt, int steps, const double* m,
                    double* x, double* y, double* z, double* vx, double* vy, double* vz, double* ax,
                    double* ay, double* az) {
    std::vector<double> x2(n), y2(n), z2(n), vx2(n), vy2(n), vz2(n), ax2(n), ay2(n), az2(n);
    std::vector<double> x3(n), y3(n), z3(n), vx3(n), vy3(n), vz3(n), ax3(n), ay3(n), az3(n);
    std::vector<double> x4(n), y4(n), z4(n), vx4(n), vy4(n), vz4(n), ax4(n), ay4(n), az4(n);
    const double h = dt / 2.0, sixth = dt / 6.0;
    for (int k = 0; k < steps; ++k) {
        SoA s1{n, m, x, y, z};
        direct_into(s1, G, soft2, ax, ay, az);
        for (int64_t i = 0; i < n; ++i) {
            x2[i] = x[i] + vx[i] * h;
            y2[i] = y[i] + vy[i] * h;
            z2[i] = z[i] + vz[i] * h;
            vx2[i] = vx[i] + ax[i] * h;
            vy2[i] = vy[i] + ay[i] * h;
            vz2[i] = vz[i] + az[i] * h;
        }
        SoA s2{n, m, x2.data(), y2.data(), z2.data()};
        direct_into(s2, G, soft2, ax2.data(), ay2.data(), az2.data());
        for (int64_t i = 0; i < n; ++i) {
            x3[i] = x[i] + vx2[i] * h;
            y3[i] = y[i] + vy2[i] * h;
            z3[i] = z[i] + vz2[i] * h;
            vx3[i] = vx[i] + ax2[i] * h;
            vy3[i] = vy[i] + ay2[i] * h;
            vz3[i] = vz[i] + az2[i] * h;
        }
        SoA s3{n, m, x3.data(), y3.data(), z3.data()};
        direct_into(s3, G, soft2, ax3.data(), ay3.data(), az3.data());
        for (int64_t i = 0; i < n; ++i) {
            x4[i] = x[i] + vx3[i] * dt;
            y4[i] = y[i] + vy3[i] * dt;
            z4[i] = z[i] + vz3[i] * dt;
            vx4[i] = vx[i] + ax3[i] * dt;
            vy4[i] = vy[i] + ay3[i] * dt;
            vz4[i] = vz[i] + az3[i] * dt;
        }
        SoA s4{n, m, x4.data(), y4.data(), z4.data()};
        direct_into(s4, G, soft2, ax4.data(), ay4.data(), az4.data());
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] + (((vx[i] + 2.0 * vx2[i]) + 2.0 * vx3[i]) + vx4[i]) * sixth;
            y[i] = y[i] + (((vy[i] + 2.0 * vy2[i]) + 2.0 * vy3[i]) + vy4[i]) * sixth;
            z[i] = z[i] + (((vz[i] + 2.0 * vz2[i]) + 2.0 * vz3[i]) + vz4[i]) * sixth;
            vx[i] = vx[i] + (((ax[i] + 2.0 * ax2[i]) + 2.0 * ax3[i]) + ax4[i]) * sixth;
            vy[i] = vy[i] + (((ay[i] + 2.0 * ay2[i]) + 2.0 * ay3[i]) + ay4[i]) * sixth;
            vz[i] = vz[i] + (((az[i] + 2.0 * az2[i]) + 2.0 * az3[i]) + az4[i]) * sixth;
        }
    }
}

// First-order and velocity-Verlet integrators with the direct force, in the
// CPU form: the force is evaluated at the start of every step.
// kind 0 = Euler (CPU/src/integrator_euler.cpp:3-15):
//   a = f(x); x = x + v*dt; v = v + a*dt.
// kind 1 = Euler-Cromer (CPU/src/integrator_euler-cromer.cpp:3-15):
//   a = f(x); v = v + a*dt; x = x + v*dt.
// kind 2 = velocity Verlet (CPU/src/integrator_velocity_verlet.cpp:3-23):
//   a = f(x); x = x + ((v*dt) + (((a*0.5)*dt)*dt)); a_old = a; a = f(x);
//   v = v + ((a_old + a)*(0.5*dt)).
// On return a* holds the last force evaluated (the reference leaves it there).
void orc_step1_direct(int64_t n, int kind, double G, double soft2, double dt, int steps,
                      const double* m, double* x, double* y, double* z, double* vx, double* vy,
                      double* vz, double* ax, double* ay, double* az) {
    std::vector<double> ox(n), oy(n), oz(n);
    const double hdt = 0.5 * dt;
    for (int k = 0; k < steps; ++k) {
        SoA s{n, m, x, y, z};
        direct_into(s, G, soft2, ax, ay, az);
        if (kind == 0) {
            drift(n, x, y, z, vx, vy, vz, dt);
            kick(n, vx, vy, vz, ax, ay, az, dt);
        } else if (kind == 1) {
            kick(n, vx, vy, vz, ax, ay, az, dt);
            drift(n, x, y, z, vx, vy, vz, dt);
        } else {
            for (int64_t i = 0; i < n; ++i) {
                x[i] = x[i] + ((vx[i] * dt) + (((ax[i] * 0.5) * dt) * dt));
                y[i] = y[i] + ((vy[i] * dt) + (((ay[i] * 0.5) * dt) * dt));
                z[i] = z[i] + ((vz[i] * dt) + (((az[i] * 0.5) * dt) * dt));
                ox[i] = ax[i];
                oy[i] = ay[i];
                oz[i] = az[i];
            }
            direct_into(s, G, soft2, ax, ay, az);
            for (int64_t i = 0; i < n; ++i) {
                vx[i] = vx[i] + ((ox[i] + ax[i]) * hdt);
                vy[i] = vy[i] + ((oy[i] + ay[i]) * hdt);
                vz[i] = vz[i] + ((oz[i] + az[i]) * hdt);
            }
        }
    }
}

// Optimized Barnes-Hut force (monopole, reference opening criterion).
// stats (may be NULL): [0]=pool nodes incl. empty children, [1]=internal
// nodes, [2]=max node depth, [3]=max walk stack length.
void orc_bhopt(int64_t n, double G, double soft2, double theta, int leaf_cap, const double* m,
               const double* x, const double* y, const double* z, double* ax, double* ay,
               double* az, int64_t* stats) {
    SoA s{n, m, x, y, z};
    PoolTree t;
    t.build(s, leaf_cap);
    const double theta_sqr = theta * theta;
    int max_stack = 0;
#pragma omp parallel for schedule(static) reduction(max : max_stack)
    for (int64_t i = 0; i < n; ++i) {
        int hi = 0;
        V3 a = t.walk(i, G, soft2, theta_sqr, &hi);
        ax[i] = a.x;
        ay[i] = a.y;
        az[i] = a.z;
        max_stack = std::max(max_stack, hi);
    }
    if (stats) {
        int64_t internal = 0;
        int maxd = 0;
        for (const PoolNode& nd : t.pool) {
            if (nd.first_child != -1) ++internal;
            maxd = std::max(maxd, nd.depth);
        }
        stats[0] = (int64_t)t.pool.size();
        stats[1] = internal;
        stats[2] = maxd;
        stats[3] = max_stack;
    }
}

// Times orc_bhopt on a bounded sample (CPU baseline, not a checker): the full
// tree build, then the walk of n_sample targets spaced n / n_sample apart
// (OpenMP over targets, as the reference walks them).  ms = {build, walk}.
void orc_bhopt_sample(int64_t n, double G, double soft2, double theta, int leaf_cap,
                      const double* m, const double* x, const double* y, const double* z,
                      int64_t n_sample, double* ax, double* ay, double* az, double* ms) {
    SoA s{n, m, x, y, z};
    PoolTree t;
    auto t0 = std::chrono::steady_clock::now();
    t.build(s, leaf_cap);
    auto t1 = std::chrono::steady_clock::now();
    const double theta_sqr = theta * theta;
    const int64_t stride = n_sample > 0 && n / n_sample > 0 ? n / n_sample : 1;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n_sample; ++k) {
        int hi = 0;
        const V3 a = t.walk((k * stride) % n, G, soft2, theta_sqr, &hi);
        ax[k] = a.x;
        ay[k] = a.y;
        az[k] = a.z;
    }
    auto t2 = std::chrono::steady_clock::now();
    ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
}

// Canonical topology of the optimized Barnes-Hut tree: for every particle the
// depth of the leaf holding it and the octant path to that leaf (3 bits per
// level, root first, digit = x | y<<1 | z<<2).  n_internal receives the number
// of internal (split) nodes.  Returns the number of pool nodes including the
// empty children the reference allocates.
int64_t orc_bhopt_leaves(int64_t n, int leaf_cap, const double* m, const double* x,
                         const double* y, const double* z, int32_t* leaf_depth,
                         uint64_t* leaf_path, int64_t* n_internal) {
    SoA s{n, m, x, y, z};
    PoolTree t;
    t.build(s, leaf_cap);
    if (n == 0) return 0;
    // DFS from the root recording paths.
    struct Item {
        int64_t node;
        uint64_t path;
        int depth;
    };
    std::vector<Item> st{{t.root, 0, 0}};
    int64_t internal = 0;
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        const PoolNode& nd = t.pool[it.node];
        if (nd.first_child == -1) {
            for (int64_t q : nd.members) {
                leaf_depth[q] = it.depth;
                leaf_path[q] = it.path;
            }
            continue;
        }
        ++internal;
        for (int k = 0; k < 8; ++k)
            st.push_back(Item{nd.first_child + k, (it.path << 3) | (uint64_t)k, it.depth + 1});
    }
    if (n_internal) *n_internal = internal;
    return (int64_t)t.pool.size();
}

// Recursive Barnes-Hut force (leaf capacity 1).
void orc_bhrec(int64_t n, double G, double soft2, double theta, const double* m, const double* x,
               const double* y, const double* z, double* ax, double* ay, double* az) {
    SoA s{n, m, x, y, z};
    RecTree t;
    t.build(s);
    const double theta_sqr = theta * theta;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        V3 a{0.0, 0.0, 0.0};
        if (t.root >= 0) t.walk(t.root, i, s.pos(i), G, soft2, theta_sqr, a);
        ax[i] = a.x;
        ay[i] = a.y;
        az[i] = a.z;
    }
}

// Root cell of the reference CPU trees (force_calculator_barnes_hut_optimized.cpp:105-124):
// centre = (min+max)/2 per axis, size = max extent + 1e-9.
void orc_root(int64_t n, const double* x, const double* y, const double* z, double* center,
              double* size) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = std::numeric_limits<double>::max();
        mx[a] = std::numeric_limits<double>::lowest();
    }
    const double* p[3] = {x, y, z};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], p[a][i]);
            mx[a] = std::max(mx[a], p[a][i]);
        }
    double s = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2]));
    *size = s + 1e-9;
    for (int a = 0; a < 3; ++a) center[a] = (mn[a] + mx[a]) / 2;
}

// 63-bit octant-path keys: 21 levels of the reference's FP64 subdivision
// replayed per particle (child centre = centre +/- (size/2)/2, octant bit =
// coord >= centre; force_calculator_barnes_hut_optimized.cpp:159-161,219-227).
// Level 0 occupies bits 62..60 (digit = x | y<<1 | z<<2), so sorting keys
// orders particles by the reference's child order at every level.
void orc_keys63(int64_t n, const double* x, const double* y, const double* z, const double* center,
                double size, uint64_t* keys) {
    for (int64_t i = 0; i < n; ++i) {
        double c[3] = {center[0], center[1], center[2]};
        const double p[3] = {x[i], y[i], z[i]};
        double s = size;
        uint64_t k = 0;
        for (int l = 0; l < 21; ++l) {
            double half = s / 2.0;
            double off = half / 2.0;
            uint64_t d = 0;
            for (int a = 0; a < 3; ++a) {
                if (p[a] >= c[a]) {
                    d |= (uint64_t)1 << a;
                    c[a] = c[a] + off;
                } else {
                    c[a] = c[a] + (-off);
                }
            }
            k = (k << 3) | d;
            s = half;
        }
        keys[i] = k;
    }
}

// Diagnostics in 80-bit long double.  out: [0]=kinetic energy,
// [1]=potential energy with the reference's unsoftened pair sum
// (CPU/include/tests_common.hpp:42-53), [2]=softened potential
// -G m_i m_j / sqrt(r^2+eps2), [3..5]=momentum, [6..8]=angular momentum
// (CPU/include/tests_common.hpp:59-65), [9..11]=centre of mass, [12]=total mass.
void orc_diag(int64_t n, double G, double soft2, const double* m, const double* x, const double* y,
              const double* z, const double* vx, const double* vy, const double* vz, double* out) {
    LD ke = 0, pe = 0, pes = 0, P[3] = {0, 0, 0}, L[3] = {0, 0, 0}, C[3] = {0, 0, 0}, M = 0;
    for (int64_t i = 0; i < n; ++i) {
        LD mi = m[i];
        ke += 0.5L * mi * ((LD)vx[i] * vx[i] + (LD)vy[i] * vy[i] + (LD)vz[i] * vz[i]);
        P[0] += mi * vx[i];
        P[1] += mi * vy[i];
        P[2] += mi * vz[i];
        L[0] += mi * ((LD)y[i] * vz[i] - (LD)z[i] * vy[i]);
        L[1] += mi * ((LD)z[i] * vx[i] - (LD)x[i] * vz[i]);
        L[2] += mi * ((LD)x[i] * vy[i] - (LD)y[i] * vx[i]);
        C[0] += mi * x[i];
        C[1] += mi * y[i];
        C[2] += mi * z[i];
        M += mi;
    }
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pe, pes)
    for (int64_t i = 0; i < n; ++i) {
        LD a = 0, b = 0;
        for (int64_t j = i + 1; j < n; ++j) {
            LD dx = (LD)x[j] - x[i], dy = (LD)y[j] - y[i], dz = (LD)z[j] - z[i];
            LD r2 = dx * dx + dy * dy + dz * dz;
            LD mm = (LD)m[i] * m[j];
            a -= mm / sqrtl(r2);
            b -= mm / sqrtl(r2 + soft2);
        }
        pe += a;
        pes += b;
    }
    out[0] = (double)ke;
    out[1] = (double)(G * pe);
    out[2] = (double)(G * pes);
    for (int a = 0; a < 3; ++a) {
        out[3 + a] = (double)P[a];
        out[6 + a] = (double)L[a];
        out[9 + a] = (double)(C[a] / M);
    }
    out[12] = (double)M;
}

// Error statistics of CPU/tests/test2_ForceBenchmark.cpp:84-150: mean and max
// of |a-a_ref|, and of |a-a_ref|/|a_ref| over particles with |a_ref| > 1e-12;
// both means divide by N.  out = {mean_abs, max_abs, mean_rel, max_rel}.
void orc_error_stats(int64_t n, const double* ax, const double* ay, const double* az,
                     const double* rx, const double* ry, const double* rz, double* out) {
    double sa = 0, ma = 0, sr = 0, mr = 0;
    for (int64_t i = 0; i < n; ++i) {
        V3 d = sub(V3{ax[i], ay[i], az[i]}, V3{rx[i], ry[i], rz[i]});
        double e = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        sa += e;
        ma = std::max(ma, e);
        double ref = std::sqrt(rx[i] * rx[i] + ry[i] * ry[i] + rz[i] * rz[i]);
        if (ref > 1e-12) {
            double r = e / ref;
            sr += r;
            mr = std::max(mr, r);
        }
    }
    out[0] = sa / (double)n;
    out[1] = ma;
    out[2] = sr / (double)n;
    out[3] = mr;
}

// Particles drawn from one std::mt19937 stream with
// std::uniform_real_distribution<double>: mass in [0.5,2), coordinates in
// [-100,100).  order=0: m,x,y,z per particle (the particles_*.txt fixtures);
// order=1: z,y,x,m (g++'s right-to-left argument evaluation in
// CPU/tests/test2_ForceBenchmark.cpp:48-57); order=2: x,y,z,m
// (CUDA/tests/test2_ForceBenchmark.cu:63-66).
void orc_mt19937_particles(uint32_t seed, int64_t n, int order, double* m, double* x, double* y,
                           double* z) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> pos(-100.0, 100.0);
    std::uniform_real_distribution<double> mass(0.5, 2.0);
    for (int64_t i = 0; i < n; ++i) {
        if (order == 0) {
            m[i] = mass(rng);
            x[i] = pos(rng);
            y[i] = pos(rng);
            z[i] = pos(rng);
        } else if (order == 1) {
            z[i] = pos(rng);
            y[i] = pos(rng);
            x[i] = pos(rng);
            m[i] = mass(rng);
        } else {
            x[i] = pos(rng);
            y[i] = pos(rng);
            z[i] = pos(rng);
            m[i] = mass(rng);
        }
    }
}

}  // extern "C"
