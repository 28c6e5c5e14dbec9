"""Writes the golden fixtures under tests/golden/ by running the reference
program (oracle/_ref/gramsci, built from the reference sources by
`make -C oracle ref`) on small seeded catalogues.

Each case directory holds
    input.npz      xyz_d, w_d, xyz_r, w_r (float64; w_* absent = unit weights)
    <run>.txt      the reference's output file for that run, byte for byte
and manifest.json lists the parameters and the runs:
    DR  -gal D -ran R        (the normal invocation)
    D   -gal D               (N2 = DD / N_d^2 -> exact ordered DD counts)
    R   -gal R               (N2 = RR / N_r^2)
    T   -gal D+R as one file (N2 = (DD + DR + RD + RR) / N^2)
The D/R/T runs recover exact integer counts per bin from the printed N2.

Usage:  python tests/golden/make_golden.py [case ...]   (needs the reference binary;
        with case names, only those cases are rewritten)
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from gramsci_amd import catalog as C  # noqa: E402


def case_defs():
    cases = []
    # 2PCF, linear bins, nmu = 1 (the shape of the plain CPU configuration, scaled down)
    d = C.uniform(1200, 100.0, 11)
    r = C.uniform(2400, 100.0, 12)
    cases.append(dict(name="c1_lin", mode="-2pcf", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=2.0, rmax=10.0, nbins=10, nmu=1, log_bins=False), runs=["DR", "D", "R", "T"]))
    # 2PCF (r, mu), log bins, clustered data, as in the GPU (r, mu) configuration
    d = C.clustered(1500, 150.0, 21, members=10, sigma=2.0)
    r = C.uniform(3000, 150.0, 22)
    cases.append(dict(name="rmu_log", mode="-2pcf", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=1.0, rmax=30.0, nbins=8, nmu=6, log_bins=True), runs=["DR", "D", "R", "T"]))
    cases.append(dict(name="rmu_log_w", mode="-2pcf", d=d, r=r,
                      wd=C.lognormal_weights(1500, 0.2, 23), wr=C.lognormal_weights(3000, 0.2, 24),
                      p=dict(rmin=1.0, rmax=30.0, nbins=8, nmu=6, log_bins=True), runs=["DR"]))
    # the parameters of the reference's own regression test (tests/run_correlation_tests.py:20)
    d = C.clustered(1500, 120.0, 31, members=10, sigma=2.0)
    r = C.uniform(3000, 120.0, 32)
    cases.append(dict(name="rmu_lin", mode="-2pcf", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=1.0, rmax=30.0, nbins=10, nmu=10, log_bins=False), runs=["DR", "D", "R", "T"]))
    # 3PCF
    d = C.clustered(1200, 80.0, 41, members=10, sigma=2.0)
    r = C.uniform(2400, 80.0, 42)
    cases.append(dict(name="p3", mode="-3pcf", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=4.0, rmax=12.0, nbins=5, nmu=1, log_bins=False), runs=["DR", "D", "R", "T"]))
    cases.append(dict(name="p3_w", mode="-3pcf", d=d, r=r,
                      wd=C.lognormal_weights(1200, 0.2, 43), wr=C.lognormal_weights(2400, 0.2, 44),
                      p=dict(rmin=4.0, rmax=12.0, nbins=5, nmu=1, log_bins=False), runs=["DR"]))
    # 4PCF
    d = C.clustered(1000, 60.0, 51, members=10, sigma=2.0)
    r = C.uniform(1000, 60.0, 52)
    cases.append(dict(name="p4", mode="-4pcf", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=3.0, rmax=8.0, nbins=3, nmu=1, log_bins=False), runs=["DR", "D", "R", "T"]))
    cases.append(dict(name="p4_w", mode="-4pcf", d=d, r=r,
                      wd=C.lognormal_weights(1000, 0.2, 53), wr=C.lognormal_weights(1000, 0.2, 54),
                      p=dict(rmin=3.0, rmax=8.0, nbins=3, nmu=1, log_bins=False), runs=["DR"]))
    # exact-edge cases: a cubic lattice puts many pairs exactly on bin edges
    lat = C.lattice(12, 1.0, 0.0)
    cases.append(dict(name="lattice_lin", mode="-2pcf", d=lat, r=None, wd=None, wr=None,
                      p=dict(rmin=1.0, rmax=4.0, nbins=3, nmu=1, log_bins=False), runs=["D"]))
    cases.append(dict(name="lattice_log_mu", mode="-2pcf", d=lat + 1.0, r=None, wd=None, wr=None,
                      p=dict(rmin=1.0, rmax=8.0, nbins=3, nmu=4, log_bins=True), runs=["D"]))
    lat10 = C.lattice(10, 1.0, 0.0)
    cases.append(dict(name="lattice_p3", mode="-3pcf", d=lat10, r=None, wd=None, wr=None,
                      p=dict(rmin=1.0, rmax=3.0, nbins=2, nmu=1, log_bins=False), runs=["D"]))
    # coincident points and rmin = 0
    d = C.uniform(1100, 50.0, 61)
    d = np.concatenate([d, d[:60]], axis=0)
    cases.append(dict(name="dups_rmin0", mode="-2pcf", d=d, r=None, wd=None, wr=None,
                      p=dict(rmin=0.0, rmax=5.0, nbins=5, nmu=1, log_bins=False), runs=["D"]))
    # RSD 3PCF (-3pcf with nmu >= 2: find_normal mu bins)
    d = C.clustered(1200, 80.0, 41, members=10, sigma=2.0)
    r = C.uniform(2400, 80.0, 42)
    cases.append(dict(name="p3_rsd", mode="-3pcf", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=4.0, rmax=12.0, nbins=3, nmu=4, log_bins=False), runs=["DR", "D", "R", "T"]))
    cases.append(dict(name="p3_rsd_w", mode="-3pcf", d=d, r=r,
                      wd=C.lognormal_weights(1200, 0.2, 43), wr=C.lognormal_weights(2400, 0.2, 44),
                      p=dict(rmin=4.0, rmax=12.0, nbins=3, nmu=5, log_bins=True), runs=["DR"]))
    # equilateral 3PCF (-equi), isotropic and RSD
    cases.append(dict(name="equi", mode="-equi", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=4.0, rmax=12.0, nbins=4, nmu=1, log_bins=False), runs=["DR", "D", "R", "T"]))
    cases.append(dict(name="equi_w", mode="-equi", d=d, r=r,
                      wd=C.lognormal_weights(1200, 0.2, 43), wr=C.lognormal_weights(2400, 0.2, 44),
                      p=dict(rmin=4.0, rmax=12.0, nbins=4, nmu=1, log_bins=True), runs=["DR"]))
    cases.append(dict(name="equi_rsd", mode="-equi", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=4.0, rmax=12.0, nbins=3, nmu=6, log_bins=False), runs=["DR", "D"]))
    # regular-tetrahedron 4PCF (-tetra)
    d = C.clustered(1000, 60.0, 51, members=10, sigma=2.0)
    r = C.uniform(1000, 60.0, 52)
    cases.append(dict(name="tetra", mode="-tetra", d=d, r=r, wd=None, wr=None,
                      p=dict(rmin=3.0, rmax=8.0, nbins=3, nmu=1, log_bins=False), runs=["DR", "D", "R", "T"]))
    cases.append(dict(name="tetra_w", mode="-tetra", d=d, r=r,
                      wd=C.lognormal_weights(1000, 0.2, 53), wr=C.lognormal_weights(1000, 0.2, 54),
                      p=dict(rmin=3.0, rmax=8.0, nbins=2, nmu=1, log_bins=False), runs=["DR"]))
    # lattice: sides 1, sqrt(2), sqrt(3) on and around the bin edges
    lat10b = C.lattice(10, 1.0, 0.0)   # the reference needs N >= 999 points (create_graph(1,999))
    cases.append(dict(name="lattice_equi", mode="-equi", d=lat10b, r=None, wd=None, wr=None,
                      p=dict(rmin=0.5, rmax=2.0, nbins=3, nmu=1, log_bins=False), runs=["D"]))
    cases.append(dict(name="lattice_tetra", mode="-tetra", d=lat10b, r=None, wd=None, wr=None,
                      p=dict(rmin=1.0, rmax=2.0, nbins=1, nmu=1, log_bins=False), runs=["D"]))
    return cases


def main():
    if not os.path.exists(O.REF_BIN):
        sys.exit("reference binary missing: run `make -C oracle ref` first")
    only = set(sys.argv[1:])
    path = os.path.join(HERE, "manifest.json")
    old = {}
    if only and os.path.exists(path):
        with open(path) as f:
            old = {m["name"]: m for m in json.load(f)}
    manifest = []
    for cs in case_defs():
        if only and cs["name"] not in only:
            if cs["name"] in old:
                manifest.append(old[cs["name"]])
            continue
        cdir = os.path.join(HERE, cs["name"])
        os.makedirs(cdir, exist_ok=True)
        arrs = dict(xyz_d=cs["d"])
        if cs["r"] is not None:
            arrs["xyz_r"] = cs["r"]
        if cs["wd"] is not None:
            arrs["w_d"] = cs["wd"]
        if cs["wr"] is not None:
            arrs["w_r"] = cs["wr"]
        np.savez_compressed(os.path.join(cdir, "input.npz"), **arrs)
        p = O.Params(**cs["p"])
        for run in cs["runs"]:
            if run == "DR":
                txt = O.reference_table(cs["d"], cs["wd"], cs["r"], cs["wr"], p, cs["mode"], threads=1)
            elif run == "D":
                txt = O.reference_table(cs["d"], cs["wd"], None, None, p, cs["mode"], threads=1)
            elif run == "R":
                txt = O.reference_table(cs["r"], cs["wr"], None, None, p, cs["mode"], threads=1)
            elif run == "T":
                txt = O.reference_table(np.concatenate([cs["d"], cs["r"]]), None, None, None, p, cs["mode"],
                                        threads=1)
            with open(os.path.join(cdir, f"{run}.txt"), "w") as f:
                f.write(txt)
        manifest.append(dict(name=cs["name"], mode=cs["mode"], params=cs["p"], runs=cs["runs"],
                             n_data=int(len(cs["d"])), n_rand=int(0 if cs["r"] is None else len(cs["r"])),
                             weighted=cs["wd"] is not None))
        print("wrote", cs["name"])
    with open(path, "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
