import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "n-body-gravity-simulation_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950)")


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def nb():
    import nbody_amd

    return nbody_amd


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "ref_golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orbit_rows():
    """Header and sampled rows of the reference's CPU/orbit.csv."""
    with open(os.path.join(GOLDEN, "orbit_csv_rows.json")) as f:
        return json.load(f)


def load_fixture(name):
    """Particle file from tests/golden (reference layout: N, then m x y z vx vy vz)."""
    with open(os.path.join(GOLDEN, name)) as f:
        n = int(f.readline())
        a = np.loadtxt(f).reshape(n, 7)
    return {k: a[:, i].copy() for i, k in enumerate(("m", "x", "y", "z", "vx", "vy", "vz"))}


FIXTURES = {"p100": "particles_100.txt", "p1000": "particles_1000.txt", "c1": "config1_n1000.txt"}


@pytest.fixture(scope="session")
def fixtures():
    return {k: load_fixture(v) for k, v in FIXTURES.items()}


def sha256(*arrays):
    import hashlib

    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def rel_err(a, r):
    a = np.stack(a, 1)
    r = np.stack(r, 1)
    return np.linalg.norm(a - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-300)
