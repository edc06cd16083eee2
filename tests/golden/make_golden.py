"""Writes the golden fixtures under tests/golden/ from the COMPILED REFERENCE
(oracle/_ref/libnbref.so, built by `make -C oracle ref` from the reference's
own CPU sources) and from the reference's committed outputs.

Run in a container that has /root/reference:  python tests/golden/make_golden.py

Outputs (data only -- inputs, hashes and expected values):
  ref_golden.json        sha256 of the reference's FP64 outputs per case, plus
                         a few %.17g sample values per case
  output2_4_errors.json  the 200 error statistics printed by the reference's
                         CPU/tests/test2_ForceBenchmark.cpp in CPU/outputs/output2_4.txt
  orbit_csv_rows.json    header and a sample of rows of the reference's CPU/orbit.csv
                         (Earth-Sun, velocity Verlet, dt = 1/365, printed by CPU/src/n-body.cpp)
  config1_n1000.txt      the config-1-shaped input (N=1000, m=1/N, x~U[-1,1]^3,
                         v~N(0,0.1^2) per axis) in the particle-file layout
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "n-body-gravity-simulation_amd"))

import oracle as O  # noqa: E402

REF = "/root/reference"


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def sample(arrs, names, idx=(0, 1, 7, 99, 500, 999)):
    out = {}
    for nm, a in zip(names, arrs):
        out[nm] = {str(i): repr(float(a[i])) for i in idx if i < len(a)}
    return out


def load(path):
    with open(path) as f:
        n = int(f.readline())
        a = np.loadtxt(f).reshape(n, 7)
    return {k: a[:, i].copy() for i, k in enumerate(("m", "x", "y", "z", "vx", "vy", "vz"))}


def parse_output2_4():
    txt = open(os.path.join(REF, "CPU/outputs/output2_4.txt")).read()
    runs = re.split(r"Run : (\d+)", txt)[1:]
    res = []
    for k in range(0, len(runs), 2):
        run = int(runs[k])
        body = runs[k + 1]
        for blk in re.split(r"Number of Particles : ", body)[1:]:
            n = int(blk.split()[0])
            parts = re.split(r"\nBarnes-Hut Optimized\n", blk)
            stats = []
            for p in parts:
                vals = {}
                for key, tag in (("mean_abs", "Average Absolute Error"),
                                 ("max_abs", "Maximum Absolute Error"),
                                 ("mean_rel", "Average Relative Error"),
                                 ("max_rel", "Maximum Relative Error")):
                    mm = re.search(tag + r" : (\S+)", p)
                    vals[key] = mm.group(1)
                stats.append(vals)
            res.append({"run": run, "n": n, "bh_recursive": stats[0], "bh_optimized": stats[1]})
    return res


ORBIT_ROWS = (0, 1, 2, 3, 10, 50, 91, 100, 182, 183, 250, 300, 364, 365, 366)


def orbit_rows():
    with open(os.path.join(REF, "CPU/orbit.csv")) as f:
        lines = f.read().splitlines()
    return {"header": lines[0], "n_rows": len(lines) - 1,
            "rows": {str(i): lines[1 + i] for i in ORBIT_ROWS}}


def main():
    assert O.have_ref(), "build the compiled reference first: make -C oracle ref"
    with open(os.path.join(HERE, "orbit_csv_rows.json"), "w") as f:
        json.dump(orbit_rows(), f, indent=1)
    gold = {}
    fixtures = {"p100": load(os.path.join(HERE, "particles_100.txt")),
                "p1000": load(os.path.join(HERE, "particles_1000.txt"))}

    import nbody_amd as nb  # host-only functions (IC generator, file writer)
    c1 = nb.ic_uniform_gaussian(1000, 1, 0.1)
    c1path = os.path.join(HERE, "config1_n1000.txt")
    nb.write_particles(c1path, c1["m"], c1["x"], c1["y"], c1["z"], c1["vx"], c1["vy"], c1["vz"])
    fixtures["c1"] = load(c1path)

    acc_names = ("ax", "ay", "az")
    st_names = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
    for fx in ("p100", "p1000", "c1"):
        d = fixtures[fx]
        for soft in (0.0, 0.01):
            a = O.ref_direct(d["m"], d["x"], d["y"], d["z"], soft)
            gold[f"direct/{fx}/eps{soft}"] = {"sha256": digest(*a), "sample": sample(a, acc_names)}
    for fx, steps in (("p1000", 10), ("c1", 100)):
        d = fixtures[fx]
        st = O.ref_integrate_direct("leapfrog", d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"],
                                    d["vz"], 1e-3, steps, 0.01)
        gold[f"leapfrog/{fx}/eps0.01/dt1e-3/steps{steps}"] = {"sha256": digest(*st),
                                                               "sample": sample(st, st_names)}
    for fx, steps in (("p1000", 5), ("c1", 10)):
        d = fixtures[fx]
        st = O.ref_integrate_direct("rk4", d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"],
                                    d["vz"], 1e-3, steps, 0.01)
        gold[f"rk4/{fx}/eps0.01/dt1e-3/steps{steps}"] = {"sha256": digest(*st),
                                                          "sample": sample(st, st_names)}
    for integ in ("euler", "euler_cromer", "velocity_verlet"):
        for fx, steps in (("p1000", 10), ("c1", 20)):
            d = fixtures[fx]
            st = O.ref_integrate_direct(integ, d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"],
                                        d["vz"], 1e-3, steps, 0.01)
            gold[f"{integ}/{fx}/eps0.01/dt1e-3/steps{steps}"] = {"sha256": digest(*st),
                                                                  "sample": sample(st, st_names)}
    for fx in ("p100", "p1000", "c1"):
        d = fixtures[fx]
        for soft in (0.0, 0.005):
            for cap in (1, 8, 16):
                a = O.ref_bhopt(d["m"], d["x"], d["y"], d["z"], 0.5, cap, soft)
                gold[f"bhopt/{fx}/eps{soft}/theta0.5/cap{cap}"] = {"sha256": digest(*a),
                                                                   "sample": sample(a, acc_names)}
        a = O.ref_bhrec(d["m"], d["x"], d["y"], d["z"], 0.5, 0.0)
        gold[f"bhrec/{fx}/eps0.0/theta0.5"] = {"sha256": digest(*a), "sample": sample(a, acc_names)}
        for cap in (1, 4, 8, 16, 33, 64):
            dep, path, nint, nodes = O.ref_bhopt_leaves(d["m"], d["x"], d["y"], d["z"], cap)
            gold[f"topology/{fx}/cap{cap}"] = {"sha256": digest(dep, path), "internal": nint,
                                               "pool_nodes": nodes,
                                               "max_depth": int(dep.max())}
    # Larger topology pins on Plummer spheres from the engine's own IC generator.
    for n, seed, cap in ((65536, 3, 8), (65536, 3, 16), (200000, 11, 8)):
        pl = nb.ic_plummer(n, seed)
        dep, path, nint, nodes = O.ref_bhopt_leaves(pl["m"], pl["x"], pl["y"], pl["z"], cap)
        gold[f"topology/plummer{n}_seed{seed}/cap{cap}"] = {
            "input_sha256": digest(pl["m"], pl["x"], pl["y"], pl["z"]),
            "sha256": digest(dep, path), "internal": nint, "pool_nodes": nodes,
            "max_depth": int(dep.max())}
    with open(os.path.join(HERE, "ref_golden.json"), "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "output2_4_errors.json"), "w") as f:
        json.dump(parse_output2_4(), f, indent=1)
    print(f"wrote {len(gold)} golden cases")


if __name__ == "__main__":
    main()
