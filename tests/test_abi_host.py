"""CPU tests of the C-ABI library: it loads, exports every symbol declared in
include/nbody_amd.h, validates arguments before touching a device, and its
host-side parts (Philox streams, initial conditions, particle files) are
correct.  No GPU compute is called here."""
import math
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def header_symbols():
    inc = os.path.join(ROOT, "include")
    txt = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                  if f.endswith(".h"))
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(nb):
    import ctypes

    syms = header_symbols()
    assert len(syms) > 30
    lib = ctypes.CDLL(nb.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version(nb):
    assert b"gfx950" in nb.lib.nb_version()


def test_invalid_arguments_are_rejected_before_device_use(nb):
    import ctypes as C

    h = C.c_void_p()
    assert nb.lib.nb_force_create_direct(C.byref(h), 7, 1.0, 0.0) == 1
    assert b"invalid" in nb.lib.nb_last_error()
    assert nb.lib.nb_force_create_bh(C.byref(h), 1.0, 0.0, -1.0, 8, 1) == 1
    assert nb.lib.nb_force_create_bh(C.byref(h), 1.0, 0.0, 0.5, -1, 1) == 1
    assert nb.lib.nb_force_create_bh(C.byref(h), 1.0, 0.0, 0.5, 65, 1) == 1
    assert nb.lib.nb_force_create_bh(C.byref(h), 1.0, 0.0, 0.5, 8, 3) == 1
    assert nb.lib.nb_integrator_create(C.byref(h), 9) == 1
    p = nb._Particles()
    assert nb.lib.nb_particles_alloc(C.byref(p), -5) == 1
    cnt = C.c_int(0)
    assert nb.lib.nb_server_poll(None, None, 0, C.byref(cnt)) == 1
    assert nb.lib.nb_server_poll_admitted(None, None, 0, C.byref(cnt)) == 1
    assert nb.lib.nb_server_set_direct_steps(None, 4) == 1
    with pytest.raises(nb.NBError):
        nb.ic_plummer(10, 1, a=-1.0)


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(nb, ctr, key, expect):
    """Philox4x32-10 known-answer vectors published with Random123 (Salmon et al. 2011)."""
    assert list(nb.philox4x32_10(ctr, key)) == expect


@pytest.mark.parametrize("name", ["particles_100.txt", "particles_1000.txt", "config1_n1000.txt"])
def test_particle_file_round_trip_is_byte_exact(nb, tmp_path, name):
    src = os.path.join(GOLDEN, name)
    d = nb.read_particles(src)
    out = tmp_path / name
    nb.write_particles(out, **d)
    assert open(out, "rb").read() == open(src, "rb").read()


def test_particle_file_errors(nb, tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("3\n1 2 3 4 5 6 7\n1 2 3\n")
    with pytest.raises(nb.NBError):
        nb.read_particles(bad)
    with pytest.raises(nb.NBError):
        nb.read_particles(tmp_path / "missing.txt")
    empty = tmp_path / "empty.txt"
    nb.write_particles(empty, np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
    assert nb.read_particles(empty)["m"].size == 0


def test_reference_fixture_generator(O):
    """particles_*.txt regenerate from one std::mt19937(42) stream drawing
    m, x, y, z per particle, files in the order 100, 500, 1000, 5000, 10000."""
    m, x, y, z = O.mt19937_particles(100 + 500 + 1000, order=0, seed=42)
    d = np.loadtxt(os.path.join(GOLDEN, "particles_1000.txt"), skiprows=1)
    assert np.array_equal(d[:, 0], m[600:1600]) and np.array_equal(d[:, 1], x[600:1600])
    d = np.loadtxt(os.path.join(GOLDEN, "particles_100.txt"), skiprows=1)
    assert np.array_equal(d[:, 3], z[:100])


def test_plummer_closed_forms(nb, O):
    n = 20000
    p = nb.ic_plummer(n, 9)
    r = np.sqrt(p["x"] ** 2 + p["y"] ** 2 + p["z"] ** 2)
    # Cumulative mass M(<r) = r^3 / (r^2 + 1)^{3/2}: Kolmogorov-Smirnov distance.
    rs = np.sort(r)
    cdf = rs ** 3 / (rs ** 2 + 1) ** 1.5
    ks = np.max(np.abs(cdf - np.arange(1, n + 1) / n))
    assert ks < 1.63 / math.sqrt(n)  # 1% level
    assert np.isclose(p["m"].sum(), 1.0) and np.allclose(p["m"], 1.0 / n)
    # Centre-of-mass frame.
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert abs(np.sum(p["m"] * p[k])) < 1e-12
    # Virial equilibrium 2T/|W| = 1 (sampling noise ~1/sqrt(n)).
    q = nb.ic_plummer(4000, 10)
    d = O.diag(q["m"], q["x"], q["y"], q["z"], q["vx"], q["vy"], q["vz"], G=1.0, soft2=0.0)
    assert abs(2 * d["ke"] / abs(d["pe"]) - 1.0) < 0.06
    # Velocities below the local escape speed sqrt(2) (1+r^2)^{-1/4}.
    v = np.sqrt(p["vx"] ** 2 + p["vy"] ** 2 + p["vz"] ** 2)
    assert np.all(v <= math.sqrt(2) * (1 + r ** 2) ** -0.25 * 1.01 + 1e-3)


def test_plummer_is_deterministic_and_seeded(nb):
    a = nb.ic_plummer(1000, 5)
    b = nb.ic_plummer(1000, 5)
    c = nb.ic_plummer(1000, 6)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert not np.array_equal(a["x"], c["x"])


def test_uniform_gaussian(nb):
    d = nb.ic_uniform_gaussian(50000, 3, 0.1)
    for k in ("x", "y", "z"):
        assert d[k].min() >= -1 and d[k].max() <= 1 and abs(d[k].mean()) < 0.02
    for k in ("vx", "vy", "vz"):
        assert abs(d[k].std() - 0.1) < 0.002 and abs(d[k].mean()) < 0.002
    assert np.allclose(d["m"], 1 / 50000)


def test_plummer_disk(nb):
    n = 20000
    d = nb.ic_plummer_disk(n, 4, disk_frac=0.1, rd=1.0, zd=0.05)
    mass = d["m"]
    disk = np.arange(n) >= n - round(0.1 * n)  # the disk particles come last
    assert abs(mass.sum() - 1.0) < 1e-12
    assert abs(mass[disk].sum() - 0.1) < 1e-9
    R = np.hypot(d["x"][disk], d["y"][disk])
    # Exponential disk: R ~ Gamma(2, rd), mean 2 rd (COM shift is tiny).
    assert abs(R.mean() - 2.0) < 0.1
    # sech^2 vertical profile with zd = 0.05: mean |z| = zd ln 2.
    assert abs(np.abs(d["z"][disk]).mean() - 0.05 * math.log(2)) < 0.005
    # Cold circular orbits: tangential, prograde; the radial component is only
    # the centre-of-mass velocity shift (~1/sqrt(n)).
    vr = (d["x"][disk] * d["vx"][disk] + d["y"][disk] * d["vy"][disk]) / R
    vt = (d["x"][disk] * d["vy"][disk] - d["y"][disk] * d["vx"][disk]) / R
    assert np.median(np.abs(vr)) < 0.02 * np.median(vt) and np.all(vt > 0)


@pytest.mark.parametrize("name,skip", [("particles_100.txt", 0), ("particles_1000.txt", 600)])
def test_library_regenerates_reference_fixtures_byte_for_byte(nb, tmp_path, name, skip):
    """nb_ic_reference_fixture (order 0, seed 42) plus nb_file_write rebuild the
    reference's particles_*.txt byte for byte: the files come from one stream,
    100 and 500 particles drawn before particles_1000.txt."""
    n = int(name.split("_")[1].split(".")[0])
    d = nb.ic_reference_fixture(n, seed=42, order=0, skip=skip)
    out = tmp_path / name
    nb.write_particles(out, **d)
    assert open(out, "rb").read() == open(os.path.join(GOLDEN, name), "rb").read()


@pytest.mark.parametrize("order", [0, 1, 2])
def test_reference_fixture_orders_match_the_oracle(nb, O, order):
    """The library's stream equals the oracle's std::mt19937 restatement in
    every draw order, and `skip` drops whole particles of the stream."""
    m, x, y, z = O.mt19937_particles(300, order=order, seed=7)
    d = nb.ic_reference_fixture(200, seed=7, order=order, skip=100)
    for k, ref in zip("mxyz", (m, x, y, z)):
        assert np.array_equal(d[k], ref[100:])
    with pytest.raises(nb.NBError):
        nb.ic_reference_fixture(10, order=3)


def test_node_capacity_host_bound(nb, monkeypatch):
    """The Barnes-Hut node pool: n + 21 ceil-ish(n/(cap+1)) + 1 nodes, never
    above 2^28 (the walk's 32-bit node-record offsets); NB_BH_NODE_CAP can
    only lower it."""
    monkeypatch.delenv("NB_BH_NODE_CAP", raising=False)
    for n, cap in [(1, 1), (1000, 8), (1 << 20, 8), (12345, 33), (1 << 20, 1)]:
        assert nb.bh_node_capacity(n, cap) == n + 21 * ((n + cap) // (cap + 1)) + 1
    assert nb.bh_node_capacity(0, 8) == 0
    lim = 1 << 28
    assert nb.bh_node_capacity(1 << 28, 1) == lim  # 4 nodes per particle at capacity 1
    assert nb.bh_node_capacity(70_000_000, 1) == lim
    assert nb.bh_node_capacity(lim // 30, 1) < lim
    assert nb.bh_node_capacity(200_000_000, 8) == lim
    monkeypatch.setenv("NB_BH_NODE_CAP", "5000")
    assert nb.bh_node_capacity(1 << 20, 8) == 5000
    assert nb.bh_node_capacity(100, 8) == 100 + 21 * 12 + 1  # never raised
    monkeypatch.setenv("NB_BH_NODE_CAP", str(1 << 30))
    assert nb.bh_node_capacity(1 << 28, 1) == lim


def test_leaf_capacity_table(nb, tmp_path):
    """NB_LEAF_AUTO table: the capacity of the last row with n_min <= n;
    malformed tables are rejected; None returns to the default table."""
    t = tmp_path / "caps.txt"
    t.write_text("# comment\n0 16   # small\n5000 8\n\n200000 4\n")
    nb.set_bh_tuning_file(t)
    try:
        assert [nb.auto_leaf_capacity(n) for n in (1, 4999, 5000, 199999, 200000, 10**9)] == \
            [16, 16, 8, 8, 4, 4]
        for bad in ("0 8\n0 4\n", "0 65\n", "0\n", "x 8\n", "", "0 8 9\n"):
            t.write_text(bad)
            with pytest.raises(nb.NBError):
                nb.set_bh_tuning_file(t)
        assert nb.auto_leaf_capacity(5000) == 8  # a rejected table leaves the last one
        with pytest.raises(nb.NBError):
            nb.set_bh_tuning_file(tmp_path / "missing.txt")
    finally:
        nb.set_bh_tuning_file(None)
    default = os.path.join(ROOT, "n-body-gravity-simulation_amd", "tuning", "bh_leaf_capacity.txt")
    rows = [tuple(map(int, ln.split("#")[0].split())) for ln in open(default)
            if ln.split("#")[0].strip()]
    for k, (nmin, cap) in enumerate(rows):
        assert nb.auto_leaf_capacity(nmin) == cap
        if k + 1 < len(rows):
            assert nb.auto_leaf_capacity(rows[k + 1][0] - 1) == cap
