"""ctypes binding of the CPU oracle (liboracle.so) and, when built, of the
compiled reference (_ref/libnbref.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the timed
CPU baseline, never as a product path.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnbref.so")

# The reference's gravitational constant, evaluated exactly as
# CPU/include/common.hpp:28 does: (4.0 * PI) * PI in double precision.
G_REF = 4.0 * math.pi * math.pi

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_I64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_d = C.c_double
_i = C.c_int

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.orc_direct.argtypes = [_i64, _d, _d, _D, _D, _D, _D, _D, _D, _D]
        L.orc_direct_subset.argtypes = [_i64, _d, _d, _D, _D, _D, _D, _i64, _I64, _D, _D, _D]
        L.orc_leapfrog_direct.argtypes = [_i64, _d, _d, _d, _i, _i, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_rk4_direct.argtypes = [_i64, _d, _d, _d, _i, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_step1_direct.argtypes = [_i64, _i, _d, _d, _d, _i] + [_D] * 10
        L.orc_bhopt.argtypes = [_i64, _d, _d, _d, _i, _D, _D, _D, _D, _D, _D, _D, C.c_void_p]
        L.orc_bhopt_sample.argtypes = [_i64, _d, _d, _d, _i, _D, _D, _D, _D, _i64, _D, _D, _D, _D]
        L.orc_bhopt_leaves.argtypes = [_i64, _i, _D, _D, _D, _D, _I32, _U64, C.POINTER(C.c_int64)]
        L.orc_bhopt_leaves.restype = C.c_int64
        L.orc_bhrec.argtypes = [_i64, _d, _d, _d, _D, _D, _D, _D, _D, _D, _D]
        L.orc_root.argtypes = [_i64, _D, _D, _D, _D, C.POINTER(C.c_double)]
        L.orc_keys63.argtypes = [_i64, _D, _D, _D, _D, _d, _U64]
        L.orc_diag.argtypes = [_i64, _d, _d, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_error_stats.argtypes = [_i64, _D, _D, _D, _D, _D, _D, _D]
        L.orc_mt19937_particles.argtypes = [C.c_uint32, _i64, _i, _D, _D, _D, _D]
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError(f"compiled reference not built: {REF_SO}")
        R = C.CDLL(REF_SO)
        R.ref_direct.argtypes = [_i64, _d, _D, _D, _D, _D, _D, _D, _D]
        R.ref_integrate_direct.argtypes = [_i64, _i, _d, _d, _i, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
        R.ref_bhopt.argtypes = [_i64, _d, _d, _i, _D, _D, _D, _D, _D, _D, _D]
        R.ref_bhopt_leaves.argtypes = [_i64, _i, _D, _D, _D, _D, _I32, _U64,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        R.ref_bhrec.argtypes = [_i64, _d, _d, _D, _D, _D, _D, _D, _D, _D]
        R.ref_bhrec_sample.argtypes = [_i64, _d, _d, _D, _D, _D, _D, _i64, _D, _D, _D, _D]
        _ref = R
    return _ref


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------- oracle
def direct(m, x, y, z, G=G_REF, soft2=0.0):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    lib().orc_direct(n, G, soft2, m, x, y, z, ax, ay, az)
    return ax, ay, az


def direct_subset(m, x, y, z, idx, G=G_REF, soft2=0.0):
    """Direct-sum accelerations of the targets idx only."""
    m, x, y, z = map(_c, (m, x, y, z))
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    k = len(idx)
    ax, ay, az = (np.zeros(k) for _ in range(3))
    lib().orc_direct_subset(len(m), G, soft2, m, x, y, z, k, idx, ax, ay, az)
    return ax, ay, az


def leapfrog_direct(m, x, y, z, vx, vy, vz, dt, steps, G=G_REF, soft2=0.0, two_evals=True):
    m = _c(m)
    st = [_c(a).copy() for a in (x, y, z, vx, vy, vz)]
    n = len(m)
    acc = [np.zeros(n) for _ in range(3)]
    lib().orc_leapfrog_direct(n, G, soft2, dt, steps, int(two_evals), m, *st, *acc)
    return st + acc


def rk4_direct(m, x, y, z, vx, vy, vz, dt, steps, G=G_REF, soft2=0.0):
    m = _c(m)
    st = [_c(a).copy() for a in (x, y, z, vx, vy, vz)]
    n = len(m)
    acc = [np.zeros(n) for _ in range(3)]
    lib().orc_rk4_direct(n, G, soft2, dt, steps, m, *st, *acc)
    return st + acc


STEP1_KINDS = {"euler": 0, "euler_cromer": 1, "velocity_verlet": 2}


def step1_direct(kind, m, x, y, z, vx, vy, vz, dt, steps, G=G_REF, soft2=0.0):
    """Euler / Euler-Cromer / velocity Verlet with the direct force (CPU form)."""
    m = _c(m)
    st = [_c(a).copy() for a in (x, y, z, vx, vy, vz)]
    n = len(m)
    acc = [np.zeros(n) for _ in range(3)]
    lib().orc_step1_direct(n, STEP1_KINDS[kind], G, soft2, dt, steps, m, *st, *acc)
    return st + acc


def bhopt(m, x, y, z, theta=0.5, leaf_cap=8, G=G_REF, soft2=0.0, with_stats=False):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    stats = np.zeros(4, dtype=np.int64)
    lib().orc_bhopt(n, G, soft2, theta, leaf_cap, m, x, y, z, ax, ay, az,
                    stats.ctypes.data_as(C.c_void_p))
    if with_stats:
        return (ax, ay, az), dict(pool_nodes=int(stats[0]), internal=int(stats[1]),
                                  max_depth=int(stats[2]), max_stack=int(stats[3]))
    return ax, ay, az


def bhopt_sample(m, x, y, z, n_sample, theta=0.5, leaf_cap=8, G=G_REF, soft2=0.0):
    """Timing of bhopt on a sample (CPU baseline): full build, walk of
    n_sample evenly spaced targets.  Returns (accelerations, {build_ms, walk_ms})."""
    m, x, y, z = map(_c, (m, x, y, z))
    ax, ay, az = (np.zeros(n_sample) for _ in range(3))
    ms = np.zeros(2)
    lib().orc_bhopt_sample(len(m), G, soft2, theta, leaf_cap, m, x, y, z, n_sample, ax, ay, az, ms)
    return (ax, ay, az), {"build_ms": ms[0], "walk_ms": ms[1]}


def bhopt_leaves(m, x, y, z, leaf_cap=8):
    """Per-particle (leaf depth, octant path) of the insertion-built tree."""
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    depth = np.zeros(n, dtype=np.int32)
    path = np.zeros(n, dtype=np.uint64)
    internal = C.c_int64(0)
    pool = lib().orc_bhopt_leaves(n, leaf_cap, m, x, y, z, depth, path, C.byref(internal))
    return depth, path, int(internal.value), int(pool)


def bhrec(m, x, y, z, theta=0.5, G=G_REF, soft2=0.0):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    lib().orc_bhrec(n, G, soft2, theta, m, x, y, z, ax, ay, az)
    return ax, ay, az


def root_cell(x, y, z):
    x, y, z = map(_c, (x, y, z))
    c = np.zeros(3)
    s = C.c_double(0)
    lib().orc_root(len(x), x, y, z, c, C.byref(s))
    return c, s.value


def keys63(x, y, z, center=None, size=None):
    x, y, z = map(_c, (x, y, z))
    if center is None:
        center, size = root_cell(x, y, z)
    k = np.zeros(len(x), dtype=np.uint64)
    lib().orc_keys63(len(x), x, y, z, _c(center), float(size), k)
    return k


def diag(m, x, y, z, vx, vy, vz, G=1.0, soft2=0.0):
    out = np.zeros(13)
    lib().orc_diag(len(m), G, soft2, *map(_c, (m, x, y, z, vx, vy, vz)), out)
    return dict(ke=out[0], pe=out[1], pe_soft=out[2], p=out[3:6].copy(), l=out[6:9].copy(),
                com=out[9:12].copy(), mass=out[12])


def error_stats(a, ref):
    out = np.zeros(4)
    lib().orc_error_stats(len(a[0]), *map(_c, a), *map(_c, ref), out)
    return dict(mean_abs=out[0], max_abs=out[1], mean_rel=out[2], max_rel=out[3])


def mt19937_particles(n, order, seed=42):
    m, x, y, z = (np.zeros(n) for _ in range(4))
    lib().orc_mt19937_particles(seed, n, order, m, x, y, z)
    return m, x, y, z


# --------------------------------------------------------------------------- compiled reference
def ref_direct(m, x, y, z, softening=0.0):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    ref().ref_direct(n, softening, m, x, y, z, ax, ay, az)
    return ax, ay, az


def ref_integrate_direct(integrator, m, x, y, z, vx, vy, vz, dt, steps, softening=0.0):
    m = _c(m)
    st = [_c(a).copy() for a in (x, y, z, vx, vy, vz)]
    n = len(m)
    acc = [np.zeros(n) for _ in range(3)]
    code = {"leapfrog": 0, "rk4": 1, "euler": 2, "euler_cromer": 3, "velocity_verlet": 4}[integrator]
    ref().ref_integrate_direct(n, code, softening, dt, steps, m, *st, *acc)
    return st + acc


def ref_bhopt(m, x, y, z, theta=0.5, leaf_cap=8, softening=0.0):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    ref().ref_bhopt(n, softening, theta, leaf_cap, m, x, y, z, ax, ay, az)
    return ax, ay, az


def ref_bhopt_leaves(m, x, y, z, leaf_cap=8):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    depth = np.zeros(n, dtype=np.int32)
    path = np.zeros(n, dtype=np.uint64)
    internal, nodes = C.c_int64(0), C.c_int64(0)
    ref().ref_bhopt_leaves(n, leaf_cap, m, x, y, z, depth, path, C.byref(internal), C.byref(nodes))
    return depth, path, int(internal.value), int(nodes.value)


def ref_bhrec(m, x, y, z, theta=0.5, softening=0.0):
    m, x, y, z = map(_c, (m, x, y, z))
    n = len(m)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    ref().ref_bhrec(n, softening, theta, m, x, y, z, ax, ay, az)
    return ax, ay, az


def ref_bhrec_sample(m, x, y, z, n_sample, theta=0.5, softening=0.0):
    m, x, y, z = map(_c, (m, x, y, z))
    ax, ay, az = (np.zeros(n_sample) for _ in range(3))
    ms = np.zeros(3)
    ref().ref_bhrec_sample(len(m), softening, theta, m, x, y, z, n_sample, ax, ay, az, ms)
    return (ax, ay, az), dict(build_ms=ms[0], walk_ms=ms[1], destroy_ms=ms[2])


# --------------------------------------------------------------------------- helpers
def fnv1a64(*arrays) -> str:
    """FNV-1a 64-bit hash over the raw little-endian bytes of the arrays."""
    h = 0xCBF29CE484222325
    for a in arrays:
        b = np.ascontiguousarray(a).view(np.uint8)
        # vectorised FNV is not possible; fold 8-byte words for speed instead
        w = np.frombuffer(b.tobytes() + b"\0" * ((-len(b)) % 8), dtype=np.uint64)
        for v in w.tolist():
            h ^= v
            h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"
