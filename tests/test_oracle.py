"""CPU tests of the oracle (the checker), pinned against:

* golden hashes produced from the COMPILED REFERENCE (tests/golden/ref_golden.json,
  written by tests/golden/make_golden.py from oracle/_ref/libnbref.so);
* the reference's own logged output CPU/outputs/output2_4.txt
  (tests/golden/output2_4_errors.json: 200 error statistics of test2);
* closed forms and invariants (two-body, Kepler period, momentum, reversibility).
"""
import math

import numpy as np
import pytest

from conftest import FIXTURES, load_fixture, rel_err, sha256

G_REF = 4.0 * math.pi * math.pi
ACC = ("ax", "ay", "az")


@pytest.mark.parametrize("fx", ["p100", "p1000", "c1"])
@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_direct_matches_compiled_reference(O, golden, fixtures, fx, eps):
    d = fixtures[fx]
    a = O.direct(d["m"], d["x"], d["y"], d["z"], G_REF, eps * eps)
    assert sha256(*a) == golden[f"direct/{fx}/eps{eps}"]["sha256"]


@pytest.mark.parametrize("fx,steps", [("p1000", 10), ("c1", 100)])
@pytest.mark.parametrize("two_evals", [True, False])
def test_leapfrog_matches_compiled_reference(O, golden, fixtures, fx, steps, two_evals):
    d = fixtures[fx]
    st = O.leapfrog_direct(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], 1e-3,
                           steps, G_REF, 1e-4, two_evals)
    assert sha256(*st) == golden[f"leapfrog/{fx}/eps0.01/dt1e-3/steps{steps}"]["sha256"]


@pytest.mark.parametrize("fx,steps", [("p1000", 5), ("c1", 10)])
def test_rk4_matches_compiled_reference(O, golden, fixtures, fx, steps):
    d = fixtures[fx]
    st = O.rk4_direct(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], 1e-3, steps,
                      G_REF, 1e-4)
    assert sha256(*st) == golden[f"rk4/{fx}/eps0.01/dt1e-3/steps{steps}"]["sha256"]


@pytest.mark.parametrize("fx", ["p100", "p1000", "c1"])
@pytest.mark.parametrize("eps", [0.0, 0.005])
@pytest.mark.parametrize("cap", [1, 8, 16])
def test_bhopt_matches_compiled_reference(O, golden, fixtures, fx, eps, cap):
    d = fixtures[fx]
    a = O.bhopt(d["m"], d["x"], d["y"], d["z"], 0.5, cap, G_REF, eps * eps)
    assert sha256(*a) == golden[f"bhopt/{fx}/eps{eps}/theta0.5/cap{cap}"]["sha256"]


@pytest.mark.parametrize("fx", ["p100", "p1000", "c1"])
def test_bhrec_matches_compiled_reference(O, golden, fixtures, fx):
    d = fixtures[fx]
    a = O.bhrec(d["m"], d["x"], d["y"], d["z"], 0.5, G_REF, 0.0)
    assert sha256(*a) == golden[f"bhrec/{fx}/eps0.0/theta0.5"]["sha256"]


@pytest.mark.parametrize("fx", ["p100", "p1000", "c1"])
@pytest.mark.parametrize("cap", [1, 4, 8, 16, 33, 64])
def test_topology_matches_compiled_reference(O, golden, fixtures, fx, cap):
    d = fixtures[fx]
    dep, path, nint, pool = O.bhopt_leaves(d["m"], d["x"], d["y"], d["z"], cap)
    g = golden[f"topology/{fx}/cap{cap}"]
    assert sha256(dep, path) == g["sha256"]
    assert nint == g["internal"] and pool == g["pool_nodes"]
    assert pool == 1 + 8 * nint  # the reference allocates all eight children per split


@pytest.mark.parametrize("key", ["topology/plummer65536_seed3/cap8",
                                 "topology/plummer65536_seed3/cap16",
                                 "topology/plummer200000_seed11/cap8"])
def test_topology_plummer_matches_compiled_reference(O, nb, golden, key):
    g = golden[key]
    n = int(key.split("plummer")[1].split("_")[0])
    seed = int(key.split("seed")[1].split("/")[0])
    cap = int(key.split("cap")[1])
    pl = nb.ic_plummer(n, seed)
    assert sha256(pl["m"], pl["x"], pl["y"], pl["z"]) == g["input_sha256"], \
        "Plummer generator output changed; regenerate the goldens"
    dep, path, nint, pool = O.bhopt_leaves(pl["m"], pl["x"], pl["y"], pl["z"], cap)
    assert sha256(dep, path) == g["sha256"] and nint == g["internal"]


def test_reference_test2_error_lines(O):
    """Reproduce all 200 error statistics of CPU/outputs/output2_4.txt
    (CPU/tests/test2_ForceBenchmark.cpp: mt19937(42), g++ draw order z,y,x,m,
    eps=0, theta=0.5, leaf capacity 8) to the printed 6 significant digits."""
    import json
    import os

    from conftest import GOLDEN

    rows = json.load(open(os.path.join(GOLDEN, "output2_4_errors.json")))
    sizes = [100, 500, 1000, 5000, 10000]
    total = 5 * sum(sizes)
    m, x, y, z = O.mt19937_particles(total, order=1, seed=42)
    off = 0
    checked = 0
    for row in rows:
        n = row["n"]
        sl = slice(off, off + n)
        off += n
        a_ref = O.direct(m[sl], x[sl], y[sl], z[sl])
        for name, acc in (("bh_recursive", O.bhrec(m[sl], x[sl], y[sl], z[sl])),
                          ("bh_optimized", O.bhopt(m[sl], x[sl], y[sl], z[sl], 0.5, 8))):
            st = O.error_stats(acc, a_ref)
            for k, v in row[name].items():
                assert f"{st[k]:.6g}" == v, (row["run"], n, name, k, st[k], v)
                checked += 1
    assert checked == 200 and off == total


def test_keys_are_the_reference_octant_paths(O, fixtures, nb):
    """Each key's top 3*depth bits equal the octant path of the particle's leaf
    in the reference insertion tree (depth from the tree)."""
    for d in (fixtures["p1000"], nb.ic_plummer(20000, 5)):
        dep, path, _, _ = O.bhopt_leaves(d["m"], d["x"], d["y"], d["z"], 8)
        k = O.keys63(d["x"], d["y"], d["z"])
        pref = np.where(dep == 0, 0, k >> (63 - 3 * dep.astype(np.uint64)))
        assert np.array_equal(pref, path)


def test_keys_known_answers(O):
    # A cube-shaped cloud: the root cube is centred on the bbox midpoint with
    # side = extent + 1e-9 (force_calculator_barnes_hut_optimized.cpp:117-124).
    x = np.array([-1.0, 1.0, 0.0, 0.3])
    y = np.array([-1.0, 1.0, 0.0, -0.7])
    z = np.array([-1.0, 1.0, 0.0, 0.1])
    c, s = O.root_cell(x, y, z)
    assert np.array_equal(c, [0.0, 0.0, 0.0]) and s == 2.0 + 1e-9
    k = O.keys63(x, y, z)
    assert k[0] == 0                      # minimum corner: every digit 0
    assert k[1] == (1 << 63) - 1          # maximum corner: every digit 7
    assert k[2] == 7 << 60                # exactly on the centre: >= picks the upper child, then 0s
    # (0.3, -0.7, 0.1): level 0 digit x|z = 0b101; level 1 centre (0.5, -0.5, 0.5): digit 0b000;
    # level 2 centre (0.25, -0.75, 0.25): x>=, y>=, z< -> 0b011.
    assert (int(k[3]) >> 54) == (0b101 << 6) | (0b000 << 3) | 0b011


def test_error_stats_semantics(O):
    a = (np.array([1.0, 0.0, 2.0]), np.zeros(3), np.zeros(3))
    r = (np.array([2.0, 0.0, 1.0]), np.zeros(3), np.zeros(3))
    st = O.error_stats(a, r)
    # |a-r| = 1, 0, 1; |r| = 2, 0 (skipped), 1; means divide by N = 3.
    assert st["mean_abs"] == pytest.approx(2 / 3) and st["max_abs"] == 1.0
    assert st["mean_rel"] == pytest.approx((0.5 + 1.0) / 3) and st["max_rel"] == 1.0


def test_two_body_closed_form(O):
    m = np.array([3.0, 5.0])
    x = np.array([0.0, 2.0]); y = np.zeros(2); z = np.zeros(2)
    ax, ay, az = O.direct(m, x, y, z, 1.0, 0.0)
    assert ax[0] == pytest.approx(5.0 / 4.0, rel=1e-15) and ax[1] == pytest.approx(-3.0 / 4.0, rel=1e-15)
    ax, _, _ = O.direct(m, x, y, z, 1.0, 0.25)  # Plummer softening: r / (r^2+eps^2)^{3/2}
    assert ax[0] == pytest.approx(5.0 * 2.0 / (4.25 ** 1.5), rel=1e-15)


def test_direct_momentum_conservation(O, fixtures):
    d = fixtures["p1000"]
    a = O.direct(d["m"], d["x"], d["y"], d["z"], G_REF, 0.0)
    f = np.stack(a, 1) * d["m"][:, None]
    assert np.abs(f.sum(0)).max() < 1e-12 * np.abs(f).sum()


@pytest.mark.parametrize("integ", ["euler", "euler_cromer", "velocity_verlet"])
@pytest.mark.parametrize("fx,steps", [("p1000", 10), ("c1", 20)])
def test_step1_integrators_match_compiled_reference(O, golden, fixtures, integ, fx, steps):
    d = fixtures[fx]
    st = O.step1_direct(integ, d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], 1e-3,
                        steps, G_REF, 1e-4)
    assert sha256(*st) == golden[f"{integ}/{fx}/eps0.01/dt1e-3/steps{steps}"]["sha256"]


def test_direct_subset_equals_full_rows(O, golden, fixtures):
    """The target-subset direct sum reproduces the rows of the full sum, which
    is bit-identical to the compiled reference."""
    d = fixtures["p1000"]
    full = O.direct(d["m"], d["x"], d["y"], d["z"], G_REF, 1e-4)
    assert sha256(*full) == golden["direct/p1000/eps0.01"]["sha256"]
    idx = np.array([999, 0, 500, 7, 7, 123])
    sub = O.direct_subset(d["m"], d["x"], d["y"], d["z"], idx, G_REF, 1e-4)
    for a, b in zip(sub, full):
        assert np.array_equal(a, b[idx])


def test_velocity_verlet_time_reversible(O, fixtures):
    d = fixtures["c1"]
    st = O.step1_direct("velocity_verlet", d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"],
                        d["vz"], 1e-3, 20, 1.0, 1e-4)
    back = O.step1_direct("velocity_verlet", d["m"], st[0], st[1], st[2], -st[3], -st[4], -st[5],
                          1e-3, 20, 1.0, 1e-4)
    for a, b in zip(back[:3], (d["x"], d["y"], d["z"])):
        assert np.abs(a - b).max() < 1e-12


def test_leapfrog_time_reversible(O, fixtures):
    d = fixtures["c1"]
    st = O.leapfrog_direct(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], 1e-3, 20,
                           1.0, 1e-4)
    back = O.leapfrog_direct(d["m"], st[0], st[1], st[2], -st[3], -st[4], -st[5], 1e-3, 20, 1.0,
                             1e-4)
    for a, b in zip(back[:3], (d["x"], d["y"], d["z"])):
        assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("integ,order", [("leapfrog", 2), ("rk4", 4), ("euler", 1),
                                         ("euler_cromer", 1), ("velocity_verlet", 2)])
def test_kepler_convergence_order(O, integ, order):
    """Earth on a circular orbit around the Sun (G=4pi^2, 1 yr period,
    CPU/tests/test1_EarthSun.cpp): the position error after a quarter period
    falls as dt^order.  (A quarter, not a whole period: over a whole circular
    period the first-order error of Euler-Cromer cancels.)"""
    me = 3.003e-6
    m = np.array([1.0, me])
    v = 2 * math.pi * math.sqrt(1 + me)
    x0 = np.array([0.0, 1.0]); y0 = np.zeros(2); z0 = np.zeros(2)
    vx0 = np.zeros(2); vy0 = np.array([-me * v / (1 + me), v / (1 + me)]); vz0 = np.zeros(2)
    period = 1.0 / math.sqrt(1.0 + me)  # 2*pi / sqrt(G (M_sun + M_earth)) at r = 1
    errs = []
    for steps in (200, 400):
        dt = period / steps
        n = steps // 4
        if integ in O.STEP1_KINDS:
            st = O.step1_direct(integ, m, x0, y0, z0, vx0, vy0, vz0, dt, n, G_REF, 0.0)
        else:
            fn = O.leapfrog_direct if integ == "leapfrog" else O.rk4_direct
            st = fn(m, x0, y0, z0, vx0, vy0, vz0, dt, n, G_REF, 0.0)
        rx = st[0][1] - st[0][0]; ry = st[1][1] - st[1][0]
        errs.append(math.hypot(rx, ry - 1.0))  # a quarter turn from (1, 0) is (0, 1)
    ratio = errs[0] / errs[1]
    assert 2 ** order * 0.7 < ratio < 2 ** order * 1.3, (errs, ratio)


def test_velocity_verlet_reproduces_reference_orbit_csv(O, orbit_rows):
    """The Earth-Sun run of CPU/src/n-body.cpp (velocity Verlet, dt = 1/365,
    G = 4 pi^2): positions and diagnostics equal the rows of the reference's
    committed CPU/orbit.csv to its printed precision (15 decimals; the first
    row is printed with 6 significant digits)."""
    m = np.array([1.0, 3.003e-6])
    st = [np.array([0.0, 1.0]), np.zeros(2), np.zeros(2), np.zeros(2),
          np.array([0.0, 2 * math.pi]), np.zeros(2)]
    dt = 1.0 / 365.0
    done = 0
    for r in sorted(int(k) for k in orbit_rows["rows"]):
        if r > done:
            st = O.step1_direct("velocity_verlet", m, *st, dt, r - done, G_REF, 0.0)[:6]
            done = r
        ref = np.array([float(v) for v in orbit_rows["rows"][str(r)].split(",")])
        d = O.diag(m, *st, G=G_REF, soft2=0.0)
        mine = np.array([st[0][1], st[1][1], st[2][1], st[0][0], st[1][0], st[2][0], d["ke"],
                         d["pe"], d["ke"] + d["pe"], *d["l"], *d["com"]])
        if r == 0:
            assert np.allclose(mine, ref, rtol=5e-6, atol=0), r
        else:
            assert np.abs(mine - ref).max() <= 6e-16, r


def test_diag_two_body(O):
    m = np.array([2.0, 3.0])
    x = np.array([0.0, 4.0]); y = np.array([0.0, 3.0]); z = np.zeros(2)
    vx = np.array([1.0, 0.0]); vy = np.array([0.0, -1.0]); vz = np.zeros(2)
    d = O.diag(m, x, y, z, vx, vy, vz, G=1.0, soft2=1.0)
    assert d["ke"] == pytest.approx(0.5 * 2 * 1 + 0.5 * 3 * 1)
    assert d["pe"] == pytest.approx(-6.0 / 5.0)
    assert d["pe_soft"] == pytest.approx(-6.0 / math.sqrt(26.0))
    assert np.allclose(d["p"], [2.0, -3.0, 0.0])
    assert np.allclose(d["l"], [0.0, 0.0, 3.0 * (4.0 * -1.0 - 3.0 * 0.0)])
    assert np.allclose(d["com"], [12.0 / 5.0, 9.0 / 5.0, 0.0])


def test_oracle_matches_compiled_reference_directly(O, fixtures):
    """When the compiled reference is present (build container), compare the
    oracle against it on fresh random inputs too."""
    if not O.have_ref():
        pytest.skip("compiled reference not built here")
    rng = np.random.default_rng(7)
    n = 777
    m = rng.uniform(0.5, 2, n); x, y, z = rng.normal(size=(3, n))
    for eps in (0.0, 0.02):
        assert all(np.array_equal(a, b) for a, b in
                   zip(O.direct(m, x, y, z, G_REF, eps * eps), O.ref_direct(m, x, y, z, eps)))
        assert all(np.array_equal(a, b) for a, b in
                   zip(O.bhopt(m, x, y, z, 0.5, 8, G_REF, eps * eps),
                       O.ref_bhopt(m, x, y, z, 0.5, 8, eps)))
    r = O.ref_bhopt_leaves(m, x, y, z, 8)
    o = O.bhopt_leaves(m, x, y, z, 8)
    assert np.array_equal(r[0], o[0]) and np.array_equal(r[1], o[1]) and r[2] == o[2]


def test_bh_oracle_accuracy_vs_direct(O, fixtures):
    d = fixtures["p1000"]
    e = rel_err(O.bhopt(d["m"], d["x"], d["y"], d["z"], 0.5, 8),
                O.direct(d["m"], d["x"], d["y"], d["z"]))
    assert np.median(e) < 1e-2 and np.percentile(e, 99) < 0.05
