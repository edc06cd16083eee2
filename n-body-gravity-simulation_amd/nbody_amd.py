"""Python host binding of libnbody_amd.so (the gfx950 N-body engine's C ABI).

The classes mirror the reference's C++ interface so code written against it
reads the same:

* ``Particles``                 -- ParticleArrays + allocate/free/copy
                                   (CUDA/include/particle_data.hpp:6-73)
* ``ForceCalculatorNaive``      -- ForceCalculatorNaive(softening)
                                   (CPU/include/force_calculator_naive.hpp:12)
* ``BarnesHutForceCalculator``  -- BarnesHutForceCalculator(softening, theta, leaf_capacity)
                                   (CUDA/include/force_calculator_barnes_hut.hpp:56)
* ``LeapfrogIntegrator``, ``RK4Integrator`` -- TimeIntegrator::step(calculator, particles, dt)
                                   (CUDA/include/integrator.hpp:10-14)

Every call goes through the C ABI declared in include/nbody_amd.h; there is no
CPU fallback.  If the shared library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnbody_amd.so")

# The reference's G = 4*pi^2 (CPU/include/common.hpp:28), same rounding.
G_REF = 4.0 * math.pi * math.pi

NB_DIRECT_F64_EXACT, NB_DIRECT_F64, NB_DIRECT_F32 = 0, 1, 2
NB_MONOPOLE, NB_QUADRUPOLE = 0, 1
NB_LEAPFROG, NB_RK4, NB_EULER, NB_EULER_CROMER, NB_VELOCITY_VERLET = 0, 1, 2, 3, 4
STATUS = {0: "NB_OK", 1: "NB_ERR_INVALID", 2: "NB_ERR_HIP", 3: "NB_ERR_NOMEM",
          4: "NB_ERR_CAPACITY", 5: "NB_ERR_BUDGET", 6: "NB_ERR_STATE"}
PRECISIONS = {"f64_exact": NB_DIRECT_F64_EXACT, "f64": NB_DIRECT_F64, "f32": NB_DIRECT_F32}


class NBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class _Particles(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in
                                     ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m")]


class _BHStats(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_leaves", C.c_int64), ("max_level", C.c_int32),
                ("level_limit", C.c_int32), ("oversize_cells", C.c_int64),
                ("pp_interactions", C.c_int64), ("pc_interactions", C.c_int64),
                ("max_frontier", C.c_int32), ("frontier_overflow", C.c_int32),
                ("root_center", C.c_double * 3), ("root_size", C.c_double),
                ("n_groups", C.c_int64), ("flags_since_read", C.c_int32),
                ("leaf_capacity", C.c_int32), ("far_cells", C.c_int64),
                ("far_target_terms", C.c_int64)]


class _Diag(C.Structure):
    _fields_ = [("ke", C.c_double), ("pe", C.c_double), ("pe_soft", C.c_double),
                ("p", C.c_double * 3), ("l", C.c_double * 3), ("com", C.c_double * 3),
                ("mass", C.c_double)]


class _Request(C.Structure):
    _fields_ = [("n", C.c_int64), ("steps", C.c_int32), ("G", C.c_double),
                ("softening", C.c_double), ("dt", C.c_double), ("theta", C.c_double)] + \
               [(k, C.c_void_p) for k in ("m", "x", "y", "z", "vx", "vy", "vz")]


class _ServerStatus(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("pending", "active", "finished_uncollected",
                                         "bytes_in_use", "ticks", "particle_steps")]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_Dn = C.c_void_p  # nullable double*
_vp = C.c_void_p


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    sig = {
        "nb_last_error": ([], C.c_char_p),
        "nb_version": ([], C.c_char_p),
        "nb_probe_fp32_fma": ([C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
        "nb_probe_fp32_rsq_cost": ([C.POINTER(C.c_double)], C.c_int),
        "nb_device_count": ([C.POINTER(C.c_int)], C.c_int),
        "nb_set_device": ([C.c_int], C.c_int),
        "nb_synchronize": ([], C.c_int),
        "nb_particles_alloc": ([C.POINTER(_Particles), C.c_int64], C.c_int),
        "nb_particles_free": ([C.POINTER(_Particles)], C.c_int),
        "nb_particles_upload": ([C.POINTER(_Particles)] + [_Dn] * 7, C.c_int),
        "nb_particles_download": ([C.POINTER(_Particles)] + [_Dn] * 10, C.c_int),
        "nb_force_create_direct": ([C.POINTER(_vp), C.c_int, C.c_double, C.c_double], C.c_int),
        "nb_force_create_bh": ([C.POINTER(_vp), C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_int], C.c_int),
        "nb_force_destroy": ([_vp], C.c_int),
        "nb_force_compute": ([_vp, C.POINTER(_Particles)], C.c_int),
        "nb_force_name": ([_vp], C.c_char_p),
        "nb_force_direct_f32_noalloc": ([C.POINTER(_Particles), C.c_double, C.c_double, _vp],
                                        C.c_int),
        "nb_force_set_timing": ([_vp, C.c_int], C.c_int),
        "nb_force_timing": ([_vp, _D], C.c_int),
        "nb_force_stage_timing": ([_vp, _D, C.c_int], C.c_int),
        "nb_force_reset_timing": ([_vp], C.c_int),
        "nb_force_bh_stats": ([_vp, C.POINTER(_BHStats)], C.c_int),
        "nb_force_bh_topology": ([_vp, C.c_int64, _vp, _vp, _vp, _vp], C.c_int),
        "nb_force_bh_nodes": ([_vp, C.c_int64, C.POINTER(C.c_int64)] + [_vp] * 7, C.c_int),
        "nb_bh_set_tuning_file": ([C.c_char_p], C.c_int),
        "nb_bh_auto_leaf_capacity": ([C.c_int64, C.POINTER(C.c_int)], C.c_int),
        "nb_bh_node_capacity": ([C.c_int64, C.c_int], C.c_int64),
        "nb_integrator_create": ([C.POINTER(_vp), C.c_int], C.c_int),
        "nb_integrator_destroy": ([_vp], C.c_int),
        "nb_integrator_step": ([_vp, _vp, C.POINTER(_Particles), C.c_double], C.c_int),
        "nb_integrator_run": ([_vp, _vp, C.POINTER(_Particles), C.c_double, C.c_int], C.c_int),
        "nb_integrator_reset": ([_vp], C.c_int),
        "nb_integrator_name": ([_vp], C.c_char_p),
        "nb_diagnostics": ([C.POINTER(_Particles), C.c_double, C.c_double, C.c_int,
                            C.POINTER(_Diag)], C.c_int),
        "nb_philox4x32_10": ([_vp, _vp, _vp], None),
        "nb_ic_plummer": ([C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_double,
                           C.c_double] + [_D] * 7, C.c_int),
        "nb_ic_uniform_gaussian": ([C.c_int64, C.c_uint64, C.c_double] + [_D] * 7, C.c_int),
        "nb_ic_plummer_disk": ([C.c_int64, C.c_uint64] + [C.c_double] * 6 + [_D] * 7, C.c_int),
        "nb_ic_reference_fixture": ([C.c_int64, C.c_uint32, C.c_int, C.c_int64] + [_D] * 4,
                                    C.c_int),
        "nb_file_read": ([C.c_char_p, C.POINTER(C.c_int64)] + [_Dn] * 7 + [C.c_int64], C.c_int),
        "nb_file_write": ([C.c_char_p, C.c_int64] + [_D] * 7, C.c_int),
        "nb_ensemble_create": ([C.POINTER(_vp), C.c_int, _I64, _D, _D, _D], C.c_int),
        "nb_ensemble_destroy": ([_vp], C.c_int),
        "nb_ensemble_upload": ([_vp] + [_D] * 7, C.c_int),
        "nb_ensemble_download": ([_vp] + [_Dn] * 9, C.c_int),
        "nb_ensemble_run": ([_vp, C.c_int, C.c_int], C.c_int),
        "nb_ensemble_run_unbatched": ([_vp, C.c_int], C.c_int),
        "nb_ensemble_autotune": ([_vp, C.c_char_p], C.c_int),
        "nb_ensemble_load_tuning": ([_vp, C.c_char_p], C.c_int),
        "nb_ensemble_tuning": ([_vp, _vp, C.c_int], C.c_int),
        "nb_ensemble_time_force": ([_vp, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "nb_server_create": ([C.POINTER(_vp), C.c_double, C.c_int64, C.c_int, C.c_int], C.c_int),
        "nb_server_destroy": ([_vp], C.c_int),
        "nb_server_submit": ([_vp, C.POINTER(_Request), C.POINTER(C.c_int64)], C.c_int),
        "nb_server_tick": ([_vp, C.POINTER(C.c_int)], C.c_int),
        "nb_server_poll": ([_vp, _vp, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "nb_server_poll_admitted": ([_vp, _vp, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "nb_server_result": ([_vp, C.c_int64] + [_Dn] * 6, C.c_int),
        "nb_server_get_status": ([_vp, C.POINTER(_ServerStatus)], C.c_int),
        "nb_server_set_direct_steps": ([_vp, C.c_int], C.c_int),
        "nb_server_set_bh_steps": ([_vp, C.c_int], C.c_int),
        "nb_server_set_tick_budget": ([_vp, C.c_double], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


lib = _load()


def _check(rc):
    if rc != 0:
        raise NBError(rc, lib.nb_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    c = C.c_int(0)
    _check(lib.nb_device_count(C.byref(c)))
    return c.value


def probe_fp32_fma() -> dict:
    """FP32 FMA issue-rate probe (VGPR operands, no packed FP32): the
    arithmetic ceiling of this library's force kernels on the current GPU."""
    tf, ms = C.c_double(0.0), C.c_double(0.0)
    _check(lib.nb_probe_fp32_fma(C.byref(tf), C.byref(ms)))
    return {"tflops": tf.value, "ms": ms.value}


def probe_fp32_rsq_cost() -> float:
    """Issue cost of one v_rsq_f32 in FMA slots, measured in a softened pair
    term's mix (one rsqrt per 12 FMAs) against FMAs alone."""
    c = C.c_double(0.0)
    _check(lib.nb_probe_fp32_rsq_cost(C.byref(c)))
    return c.value


def synchronize():
    _check(lib.nb_synchronize())


# --------------------------------------------------------------------------- state
class Particles:
    """Device SoA particle state (ParticleArrays)."""

    def __init__(self, n: int):
        self._p = _Particles()
        _check(lib.nb_particles_alloc(C.byref(self._p), int(n)))
        self.n = int(n)

    @classmethod
    def from_arrays(cls, m, x, y, z, vx=None, vy=None, vz=None):
        p = cls(len(m))
        p.upload(m, x, y, z, vx, vy, vz)
        return p

    def upload(self, m, x, y, z, vx=None, vy=None, vz=None):
        arrs = [_f64(a) for a in (x, y, z)]
        vel = [None if a is None else _f64(a) for a in (vx, vy, vz)]
        mm = _f64(m)
        for a in arrs + [mm] + [v for v in vel if v is not None]:
            if len(a) != self.n:
                raise NBError(1, f"array length {len(a)} != n={self.n}")
        _check(lib.nb_particles_upload(C.byref(self._p), *[_ptr(a) for a in arrs],
                                       *[_ptr(v) for v in vel], _ptr(mm)))

    def download(self, fields=("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m")):
        names = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m")
        out = {k: np.empty(self.n) for k in fields}
        _check(lib.nb_particles_download(C.byref(self._p),
                                         *[_ptr(out.get(k)) for k in names]))
        return out

    def acc(self):
        d = self.download(("ax", "ay", "az"))
        return d["ax"], d["ay"], d["az"]

    def free(self):
        if getattr(self, "_p", None) is not None and self._p.n:
            _check(lib.nb_particles_free(C.byref(self._p)))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self._p)


# --------------------------------------------------------------------------- forces
class ForceCalculator:
    _h = None

    def computeAccelerations(self, particles: Particles):
        _check(lib.nb_force_compute(self._h, particles.ref))

    compute = computeAccelerations

    def name(self) -> str:
        return lib.nb_force_name(self._h).decode()

    def set_timing(self, on=True, stages=True):
        """on=False: off; stages=True: every Barnes-Hut stage; stages=False:
        tree build and walk only (fewer events on the stream)."""
        _check(lib.nb_force_set_timing(self._h, 0 if not on else 1 if stages else 2))

    def timing(self):
        out = np.zeros(4)
        _check(lib.nb_force_timing(self._h, out))
        return dict(total_ms=out[0], build_ms=out[1], walk_ms=out[2], calls=int(out[3]))

    STAGES = ("bbox", "keys", "sort", "gather", "build", "upsweep", "walk")

    def stage_timing(self):
        """Accumulated ms per stage (bbox, keys, sort, gather, build, upsweep,
        walk) and the number of timed calls."""
        out = np.zeros(8)
        _check(lib.nb_force_stage_timing(self._h, out, 8))
        d = {k: float(v) for k, v in zip(self.STAGES, out[:7])}
        d["calls"] = int(out[7])
        return d

    def reset_timing(self):
        _check(lib.nb_force_reset_timing(self._h))

    def close(self):
        if self._h:
            lib.nb_force_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ForceCalculatorNaive(ForceCalculator):
    def __init__(self, softening: float = 0.0, precision: str = "f64_exact", G: float = G_REF):
        h = C.c_void_p()
        _check(lib.nb_force_create_direct(C.byref(h), PRECISIONS[precision], float(G),
                                          float(softening)))
        self._h = h
        self.softening, self.G, self.precision = softening, G, precision


class BarnesHutForceCalculator(ForceCalculator):
    def __init__(self, softening: float = 0.0, theta: float = 0.5, leaf_capacity: int = 8,
                 multipole: str = "quadrupole", G: float = G_REF):
        h = C.c_void_p()
        mp = {"monopole": NB_MONOPOLE, "quadrupole": NB_QUADRUPOLE}[multipole]
        _check(lib.nb_force_create_bh(C.byref(h), float(G), float(softening), float(theta),
                                      int(leaf_capacity), mp))
        self._h = h
        self.softening, self.theta, self.leaf_capacity, self.G = softening, theta, leaf_capacity, G

    def stats(self, check=True):
        s = _BHStats()
        rc = lib.nb_force_bh_stats(self._h, C.byref(s))
        if check:
            _check(rc)
        return dict(n_nodes=s.n_nodes, n_leaves=s.n_leaves, max_level=s.max_level,
                    level_limit=s.level_limit, oversize_cells=s.oversize_cells,
                    pp_interactions=s.pp_interactions, pc_interactions=s.pc_interactions,
                    max_frontier=s.max_frontier, frontier_overflow=s.frontier_overflow,
                    root_center=np.array(s.root_center[:]), root_size=s.root_size,
                    n_groups=s.n_groups, flags_since_read=s.flags_since_read,
                    leaf_capacity=s.leaf_capacity, far_cells=s.far_cells,
                    far_target_terms=s.far_target_terms, status=rc)

    def topology(self, n):
        depth = np.zeros(n, np.int32)
        path = np.zeros(n, np.uint64)
        keys = np.zeros(n, np.uint64)
        perm = np.zeros(n, np.uint32)
        _check(lib.nb_force_bh_topology(self._h, n, depth.ctypes.data, path.ctypes.data,
                                        keys.ctypes.data, perm.ctypes.data))
        return dict(depth=depth, path=path, keys=keys, perm=perm)

    def nodes(self):
        cnt = C.c_int64(0)
        _check(lib.nb_force_bh_nodes(self._h, 0, C.byref(cnt), *([None] * 7)))
        k = cnt.value
        level = np.zeros(k, np.int32)
        path = np.zeros(k, np.uint64)
        first = np.zeros(k, np.int32)
        last = np.zeros(k, np.int32)
        mass = np.zeros(k)
        com = np.zeros((k, 3))
        quad = np.zeros((k, 6))
        _check(lib.nb_force_bh_nodes(self._h, k, C.byref(cnt), level.ctypes.data,
                                     path.ctypes.data, first.ctypes.data, last.ctypes.data,
                                     mass.ctypes.data, com.ctypes.data, quad.ctypes.data))
        return dict(level=level, path=path, first=first, last=last, mass=mass, com=com, quad=quad)


NB_LEAF_AUTO = 0
NB_NOALLOC_MAX_N = 1 << 20


def direct_f32_noalloc(particles: Particles, softening: float = 0.0, G: float = G_REF,
                       stream=None):
    """The NB_DIRECT_F32 force without any device allocation (the packed
    sources live in the library's static scratch array; n <= NB_NOALLOC_MAX_N).
    Asynchronous on `stream` (a hipStream_t handle as an int, None = null
    stream)."""
    _check(lib.nb_force_direct_f32_noalloc(particles.ref, float(G), float(softening), stream))


def set_bh_tuning_file(path=None):
    """Leaf-capacity table of NB_LEAF_AUTO calculators ("n_min cap" lines);
    None returns to the default table."""
    _check(lib.nb_bh_set_tuning_file(None if path is None else os.fsencode(path)))


def auto_leaf_capacity(n) -> int:
    c = C.c_int(0)
    _check(lib.nb_bh_auto_leaf_capacity(int(n), C.byref(c)))
    return c.value


def bh_node_capacity(n, leaf_capacity) -> int:
    """Node pool a Barnes-Hut calculator reserves (at most 2^28 nodes)."""
    return int(lib.nb_bh_node_capacity(int(n), int(leaf_capacity)))


class ForceCalculatorBarnesHutOptimized(BarnesHutForceCalculator):
    """ForceCalculatorBarnesHutOptimized(softening, theta, leaf_capacity)
    (CPU/include/force_calculator_barnes_hut_optimized.hpp:40)."""

    def __init__(self, softening: float, theta: float, leaf_capacity: int, G: float = G_REF):
        super().__init__(softening, theta, leaf_capacity, "quadrupole", G)


class ForceCalculatorBarnesHut(BarnesHutForceCalculator):
    """ForceCalculatorBarnesHut(softening, theta) (CPU/include/force_calculator_barnes_hut.hpp:33-34):
    the recursive build splits a leaf when a second particle arrives (unless its
    half size is below 1e-9), i.e. the same tree rule with leaf capacity 1."""

    def __init__(self, softening: float = 0.0, theta: float = 0.5, G: float = G_REF):
        super().__init__(softening, theta, 1, "quadrupole", G)


# --------------------------------------------------------------------------- integrators
class TimeIntegrator:
    _kind = NB_LEAPFROG

    def __init__(self):
        h = C.c_void_p()
        _check(lib.nb_integrator_create(C.byref(h), self._kind))
        self._h = h

    def step(self, calculator: ForceCalculator, particles: Particles, dt: float):
        _check(lib.nb_integrator_step(self._h, calculator._h, particles.ref, float(dt)))

    def run(self, calculator: ForceCalculator, particles: Particles, dt: float, steps: int):
        _check(lib.nb_integrator_run(self._h, calculator._h, particles.ref, float(dt),
                                     int(steps)))

    def reset(self):
        _check(lib.nb_integrator_reset(self._h))

    def name(self) -> str:
        return lib.nb_integrator_name(self._h).decode()

    def __del__(self):
        try:
            if self._h:
                lib.nb_integrator_destroy(self._h)
                self._h = None
        except Exception:
            pass


class LeapfrogIntegrator(TimeIntegrator):
    _kind = NB_LEAPFROG


class RK4Integrator(TimeIntegrator):
    _kind = NB_RK4


class EulerIntegrator(TimeIntegrator):
    _kind = NB_EULER


class EulerCromerIntegrator(TimeIntegrator):
    _kind = NB_EULER_CROMER


class VelocityVerletIntegrator(TimeIntegrator):
    _kind = NB_VELOCITY_VERLET


# --------------------------------------------------------------------------- diagnostics
def diagnostics(particles: Particles, G: float = G_REF, softening: float = 0.0,
                with_potential: bool = True) -> dict:
    d = _Diag()
    _check(lib.nb_diagnostics(particles.ref, float(G), float(softening), int(with_potential),
                              C.byref(d)))
    return dict(ke=d.ke, pe=d.pe, pe_soft=d.pe_soft, p=np.array(d.p[:]), l=np.array(d.l[:]),
                com=np.array(d.com[:]), mass=d.mass)


ORBIT_CSV_HEADER = ("earth_x,earth_y,earth_z,sun_x,sun_y,sun_z,kinetic_energy,potential_energy,"
                    "total_energy,angular_momentum_x,angular_momentum_y,angular_momentum_z,"
                    "center_of_mass_x,center_of_mass_y,center_of_mass_z")


class OrbitCsvWriter:
    """Trajectory rows in the column layout of the reference's orbit.csv
    (CPU/src/n-body.cpp:6,19-58): positions of `body` then `center`, kinetic,
    potential (unsoftened) and total energy, angular momentum and centre of
    mass, fixed with 15 decimals.  Diagnostics are reduced on the GPU."""

    def __init__(self, path, body=1, center=0, G=G_REF):
        self._f = open(path, "w")
        self._f.write(ORBIT_CSV_HEADER + "\n")
        self.body, self.center, self.G = body, center, G

    def row(self, particles: Particles):
        g = particles.download(("x", "y", "z"))
        d = diagnostics(particles, self.G, 0.0, True)
        b, c = self.body, self.center
        vals = [g["x"][b], g["y"][b], g["z"][b], g["x"][c], g["y"][c], g["z"][c],
                d["ke"], d["pe"], d["ke"] + d["pe"], *d["l"], *d["com"]]
        self._f.write(",".join("%.15f" % v for v in vals) + "\n")

    def close(self):
        self._f.close()


# --------------------------------------------------------------------------- initial conditions
def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib.nb_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def _alloc7(n):
    return [np.zeros(n) for _ in range(7)]


def ic_plummer(n, seed, a=1.0, M=1.0, G=1.0, rmax=0.0):
    arr = _alloc7(n)
    _check(lib.nb_ic_plummer(int(n), int(seed), a, M, G, rmax, *arr))
    return dict(zip(("m", "x", "y", "z", "vx", "vy", "vz"), arr))


def ic_uniform_gaussian(n, seed, sigma=0.1):
    arr = _alloc7(n)
    _check(lib.nb_ic_uniform_gaussian(int(n), int(seed), sigma, *arr))
    return dict(zip(("m", "x", "y", "z", "vx", "vy", "vz"), arr))


def ic_plummer_disk(n, seed, a=1.0, M=1.0, G=1.0, disk_frac=0.1, rd=1.0, zd=0.05):
    arr = _alloc7(n)
    _check(lib.nb_ic_plummer_disk(int(n), int(seed), a, M, G, disk_frac, rd, zd, *arr))
    return dict(zip(("m", "x", "y", "z", "vx", "vy", "vz"), arr))


def ic_reference_fixture(n, seed=42, order=0, skip=0):
    """The reference's fixture stream (std::mt19937(seed), mass U[0.5,2),
    positions U[-100,100)): order 0 = m,x,y,z (particles_*.txt), 1 = z,y,x,m,
    2 = x,y,z,m; the first `skip` particles are drawn and dropped.  Zero
    velocities."""
    m, x, y, z = (np.zeros(n) for _ in range(4))
    _check(lib.nb_ic_reference_fixture(int(n), int(seed), int(order), int(skip), m, x, y, z))
    return dict(m=m, x=x, y=y, z=z, vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n))


# --------------------------------------------------------------------------- particle files
def read_particles(path) -> dict:
    n = C.c_int64(0)
    _check(lib.nb_file_read(os.fsencode(path), C.byref(n), *([None] * 7), 0))
    arr = _alloc7(n.value)
    _check(lib.nb_file_read(os.fsencode(path), C.byref(n), *[a.ctypes.data for a in arr],
                            n.value))
    return dict(zip(("m", "x", "y", "z", "vx", "vy", "vz"), arr))


def write_particles(path, m, x, y, z, vx=None, vy=None, vz=None):
    n = len(m)
    zero = np.zeros(n)
    arrs = [_f64(a) for a in (m, x, y, z)] + [zero if v is None else _f64(v) for v in (vx, vy, vz)]
    _check(lib.nb_file_write(os.fsencode(path), n, *arrs))


# --------------------------------------------------------------------------- ensembles
class Ensemble:
    """Many independent systems advanced together (one persistent force launch
    per step over all systems, optional HIP-graph replay)."""

    def __init__(self, sizes, G=1.0, softening=0.01, dt=1e-3):
        sizes = np.asarray(sizes, dtype=np.int64)
        self.S = len(sizes)
        self.offsets = np.zeros(self.S + 1, np.int64)
        self.offsets[1:] = np.cumsum(sizes)
        self.N = int(self.offsets[-1])
        per = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (self.S,)))
        h = C.c_void_p()
        _check(lib.nb_ensemble_create(C.byref(h), self.S, self.offsets, per(G), per(softening),
                                      per(dt)))
        self._h = h

    def upload(self, m, x, y, z, vx, vy, vz):
        _check(lib.nb_ensemble_upload(self._h, *[_f64(a) for a in (m, x, y, z, vx, vy, vz)]))

    def download(self):
        names = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
        out = {k: np.empty(self.N) for k in names}
        _check(lib.nb_ensemble_download(self._h, *[out[k].ctypes.data for k in names]))
        return out

    def run(self, steps, use_graph=True):
        _check(lib.nb_ensemble_run(self._h, int(steps), int(use_graph)))

    def run_unbatched(self, steps):
        _check(lib.nb_ensemble_run_unbatched(self._h, int(steps)))

    def autotune(self, path=None):
        _check(lib.nb_ensemble_autotune(self._h, None if path is None else os.fsencode(path)))

    def load_tuning(self, path):
        _check(lib.nb_ensemble_load_tuning(self._h, os.fsencode(path)))

    def time_force(self, reps=5):
        """Mean ms of one batched force launch (HIP events on the ensemble's
        stream); the state keeps its bits."""
        ms = C.c_double(0.0)
        _check(lib.nb_ensemble_time_force(self._h, int(reps), C.byref(ms)))
        return ms.value

    def tuning(self):
        out = np.zeros(8, np.int32)
        _check(lib.nb_ensemble_tuning(self._h, out.ctypes.data, 8))
        return out

    def __del__(self):
        try:
            if self._h:
                lib.nb_ensemble_destroy(self._h)
                self._h = None
        except Exception:
            pass


# --------------------------------------------------------------------------- serving
class Server:
    """Continuous-batching server (see nb_server_* in include/nbody_amd.h)."""

    def __init__(self, hbm_budget_bytes=200e9, direct_threshold=32768, bh_lanes=8,
                 leaf_capacity=8, direct_steps_per_tick=None, bh_steps_per_tick=None,
                 tick_budget=None):
        h = C.c_void_p()
        _check(lib.nb_server_create(C.byref(h), float(hbm_budget_bytes), int(direct_threshold),
                                    int(bh_lanes), int(leaf_capacity)))
        self._h = h
        self._n = {}
        if direct_steps_per_tick is not None:
            _check(lib.nb_server_set_direct_steps(h, int(direct_steps_per_tick)))
        if bh_steps_per_tick is not None:
            _check(lib.nb_server_set_bh_steps(h, int(bh_steps_per_tick)))
        if tick_budget is not None:
            _check(lib.nb_server_set_tick_budget(h, float(tick_budget)))

    def submit(self, system, steps, G=1.0, softening=0.005, dt=1e-4, theta=0.5):
        arr = [_f64(system[k]) for k in ("m", "x", "y", "z", "vx", "vy", "vz")]
        r = _Request(len(arr[0]), int(steps), G, softening, dt, theta,
                     *[a.ctypes.data for a in arr])
        rid = C.c_int64(0)
        _check(lib.nb_server_submit(self._h, C.byref(r), C.byref(rid)))
        self._n[rid.value] = len(arr[0])
        return rid.value

    def tick(self):
        c = C.c_int(0)
        _check(lib.nb_server_tick(self._h, C.byref(c)))
        return c.value

    def poll(self, capacity=4096):
        ids = np.zeros(capacity, np.int64)
        c = C.c_int(0)
        _check(lib.nb_server_poll(self._h, ids.ctypes.data, capacity, C.byref(c)))
        return ids[:c.value].tolist()

    def poll_admitted(self, capacity=4096):
        """Ids admitted (state resident on the device) since the last call."""
        ids = np.zeros(capacity, np.int64)
        c = C.c_int(0)
        _check(lib.nb_server_poll_admitted(self._h, ids.ctypes.data, capacity, C.byref(c)))
        return ids[:c.value].tolist()

    def result(self, rid):
        """Final state of a finished request (collected once).  An id that is
        not finished, unknown or already collected raises NBError; so does a
        request that failed, with the state it reached as the error's `state`."""
        n = self._n.get(rid)
        if n is None:  # unknown here or already collected: the library says which
            _check(lib.nb_server_result(self._h, int(rid), *([None] * 6)))
            raise NBError(6, f"request {rid} was not submitted to this server")
        out = {k: np.empty(n) for k in ("x", "y", "z", "vx", "vy", "vz")}
        rc = lib.nb_server_result(self._h, rid, *[out[k].ctypes.data for k in
                                                  ("x", "y", "z", "vx", "vy", "vz")])
        if rc != 6:  # anything but "not finished" collects the request
            del self._n[rid]
        if rc != 0:
            err = NBError(rc, lib.nb_last_error().decode())
            err.state = out
            raise err
        return out

    def status(self):
        st = _ServerStatus()
        _check(lib.nb_server_get_status(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStatus._fields_}

    def __del__(self):
        try:
            if self._h:
                lib.nb_server_destroy(self._h)
                self._h = None
        except Exception:
            pass
