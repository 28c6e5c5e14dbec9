"""Benchmark of the counting hot path on one MI355X.

Workload (the (r, mu) 2PCF GPU configuration): 1,000,000 clustered galaxies
in a 1000^3 box (500,000 Poisson points + 25,000 Neyman-Scott clusters of 20
members, Gaussian sigma = 2) with lognormal(0, 0.2) weights, plus 3,000,000
uniform randoms (unit weights); rmin = 1, rmax = 150, 30 log-spaced bins,
nmu = 20.

A step is one full pass of the hot path with the catalogue already resident
in HBM: device-to-device catalogue hand-over, bounding box, GPU cell grid
(counting sort), and the persistent pair kernel, ending with the histograms
on the host.  value = binned ordered pairs (the reference's N2 increments)
per second of the FP32 + FP64-guard path; the plain FP64 path is timed on the
same input and reported beside it.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "binned pair/triplet evals/s on 1 MI355X: batched FP32+guard vs plain FP64 vs CPU"
# HBM bytes per dispatch of the FP32 + guard pair kernel on the (r, mu) workload, measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch) and scaled by the same counters'
# readings for a copy of known size (x2.0 for FETCH_SIZE, x1.0 for WRITE_SIZE)
PAIR_RMU_TRAFFIC_BYTES = (34_301_089.1 * 2.0 + 103_883.1 * 1.0) * 1024
# VALU wave-instructions per dispatch of the same kernel on the same workload (rocprofv3 --pmc SQ_INSTS_VALU);
# divided by the kernel time it is the issue rate, read against the rate of a pure FP32 FMA stream
PAIR_RMU_VALU_WAVE_INSTR = 183_368_978_373
FLOPS_PER_PAIR_RMU = 22      # 3 sub, 3 mul, 2 add for s; 3 mid adds, m.s (5), |m|^2 (5), 1 product
PEAK_FP32 = 157.3e12         # MI355X FP32 vector, dense
PEAK_FP64 = 78.6e12
HBM_PEAK = 8.0e12


def make_workload():
    import numpy as np
    from gramsci_amd import catalog as C
    d = C.neyman_scott(500_000, 25_000, 20, 2.0, 1000.0, 201)
    r = C.uniform(3_000_000, 1000.0, 202)
    wd = C.lognormal_weights(len(d), 0.2, 203)
    xyz = np.concatenate([d, r])
    w = np.concatenate([wd, np.ones(len(r))])
    return xyz, w, len(d)


PARAMS = dict(rmin=1.0, rmax=150.0, nbins=30, nmu=20, log_bins=True)


def time_gpu(eng, dxyz, dw, nd, precision, steps, warmup):
    import torch
    for _ in range(warmup):
        eng.set_catalogue(dxyz, dw, nd)
        eng.pcf2(**PARAMS, precision=precision)
    torch.cuda.synchronize()
    kern_ms, grid_ms = [], []
    out = None
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.set_catalogue(dxyz, dw, nd)
        out = eng.pcf2(**PARAMS, precision=precision)
        kern_ms.append(out["stats"]["ms_count"])
        grid_ms.append(out["stats"]["ms_grid"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return out, dt / steps, sum(kern_ms) / len(kern_ms), sum(grid_ms) / len(grid_ms)


def cpu_baseline(xyz, w, nd):
    """The reference program (oracle/_ref/gramsci_fast: the reference sources
    built with its own -O3 -ffast-math optimisation) on a density-matched
    150^3 sub-box of the same catalogue, same bins, all host threads used.
    Throughput = binned ordered pairs / (graph + query time printed by the
    reference).  The reference's cost per pair grows with the neighbour count
    k (per-point selection sort), and k is several times smaller in the
    sub-box than in the full box (rmax = 150 is the sub-box side), so this
    over-states its throughput on the full workload."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.REF_BIN_FAST if os.path.exists(O.REF_BIN_FAST) else O.REF_BIN
    if not os.path.exists(ref):
        return None
    side = 150.0
    lo = 425.0
    inb = np.all((xyz >= lo) & (xyz < lo + side), axis=1)
    idx = np.nonzero(inb)[0]
    d_idx = idx[idx < nd]
    r_idx = idx[idx >= nd]
    p = O.Params(**PARAMS)
    cores = min(os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory() as td:
        O.write_catalogue(os.path.join(td, "d.txt"), xyz[d_idx], w[d_idx])
        O.write_catalogue(os.path.join(td, "r.txt"), xyz[r_idx], w[r_idx])
        args = ["-gal", "d.txt", "-ran", "r.txt", "-rmin", "1.0", "-rmax", "150.0", "-nbins", "30", "-nmu", "20",
                "-log", "-out", "o.txt", "-2pcf"]
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
        t0 = time.perf_counter()
        res = subprocess.run([ref] + args, cwd=td, env=env, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if res.returncode != 0:
            return None
        tg = tq = None
        for line in res.stdout.splitlines():
            if "Creating the graph took" in line:
                tg = float(line.split("took")[1].split()[0])
            if "Querying graph took" in line:
                tq = float(line.split("took")[1].split()[0])
    sub = np.concatenate([xyz[d_idx], xyz[r_idx]])
    pairs = int(O.pcf2(sub, None, len(d_idx), p)["cnt"].sum())
    t = (tg or 0.0) + (tq or 0.0)
    if t <= 0:
        t = wall
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": pairs / t, "unit": "pair evals/s", "cores": cores, "kind": "reference",
            "sample": f"reference gramsci (-O3 -ffast-math build) on a density-matched {side:.0f}^3 sub-box "
                      f"({len(d_idx)} galaxies + {len(r_idx)} randoms) of the same catalogue, same bins; "
                      f"{pairs} binned ordered pairs in {t:.2f} s graph+query ({wall:.2f} s wall incl. ASCII I/O); "
                      f"{cpu_model}; OMP_NUM_THREADS={cores}"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def _ref_phase_times(stdout):
    tg = tq = None
    for line in stdout.splitlines():
        if "Creating the graph took" in line:
            tg = float(line.split("took")[1].split()[0])
        if "Querying graph took" in line:
            tq = float(line.split("took")[1].split()[0])
    return tg, tq


def run_reference(d, wd, r, wr, args, timeout=900):
    """Runs the reference program (-O3 -ffast-math build) on all host threads
    (at most 16); returns (graph+query seconds, wall seconds, threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.REF_BIN_FAST if os.path.exists(O.REF_BIN_FAST) else O.REF_BIN
    if not os.path.exists(ref):
        return None
    cores = min(os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory() as td:
        O.write_catalogue(os.path.join(td, "d.txt"), d, wd)
        a = ["-gal", "d.txt"]
        if r is not None:
            O.write_catalogue(os.path.join(td, "r.txt"), r, wr)
            a += ["-ran", "r.txt"]
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
        t0 = time.perf_counter()
        res = subprocess.run([ref] + a + args, cwd=td, env=env, capture_output=True, text=True, timeout=timeout)
        wall = time.perf_counter() - t0
    if res.returncode != 0:
        return None
    tg, tq = _ref_phase_times(res.stdout)
    t = (tg or 0.0) + (tq or 0.0)
    return (t if t > 0 else wall), wall, cores


def line_base(metric, unit, value, steps, warmup, t_step, dtype, workload, **cfg):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (seeded Neyman-Scott + uniform, numpy PCG64)",
            "config": dict(workload=workload, **cfg)}


def bench_rmu(args):
    import numpy as np
    import torch
    from gramsci_amd.engine import Engine

    xyz, w, nd = make_workload()
    dxyz = torch.from_numpy(xyz).cuda()
    dw = torch.from_numpy(w).cuda()
    eng = Engine(0)
    res = {}
    precisions = ["fp64"] if args.fp64_only else (["fp32"] + ([] if args.no_fp64 else ["fp64"]))
    for prec in precisions:
        out, t_step, t_kern, t_grid = time_gpu(eng, dxyz, dw, nd, prec, args.steps, args.warmup)
        res[prec] = dict(out=out, t_step=t_step, t_kern=t_kern, t_grid=t_grid)
        print(f"[bench] {prec}: {t_step * 1e3:.1f} ms/step, pair kernel {t_kern:.1f} ms, grid {t_grid:.2f} ms, "
              f"binned {out['stats']['n_binned']:.4g}, rechecks {out['stats']['n_recheck']}", file=sys.stderr,
              flush=True)
    if "fp32" in res and "fp64" in res:
        assert np.array_equal(res["fp32"]["out"]["cnt"], res["fp64"]["out"]["cnt"]), "fp32/fp64 counts differ"
    main_prec = precisions[0]
    R = res[main_prec]
    U = R["out"]["stats"]["n_binned"]
    value = U / R["t_step"]
    peak = PEAK_FP32 if main_prec == "fp32" else PEAK_FP64
    achieved = U * FLOPS_PER_PAIR_RMU / (R["t_kern"] * 1e-3)
    line = line_base(METRIC, "pair evals/s", value, args.steps, args.warmup, R["t_step"],
                     "f32" if main_prec == "fp32" else "f64",
                     "2PCF (r,mu): 1M NS-clustered galaxies (lognormal w) + 3M uniform randoms, 1000^3 box, "
                     "1<r<150, 30 log bins, nmu=20",
                     n_points=int(len(xyz)), binned_pairs_per_step=int(U),
                     precision="FP32 + FP64 guard-band recheck" if main_prec == "fp32" else "plain FP64")
    line["roofline"] = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / peak, "traffic": PAIR_RMU_TRAFFIC_BYTES if main_prec == "fp32" else None,
                        "kernel": f"k_pair2<{'true' if main_prec == 'fp64' else 'false'}, true, true, true>",
                        "kernel_ms": R["t_kern"], "flops_per_pair": FLOPS_PER_PAIR_RMU}
    if main_prec == "fp32":
        # instruction-issue reading of the same launch: its VALU wave-instructions (counter) per second
        # against those of a pure FP32 FMA stream timed now on the same GPU
        c = valu_ceiling(3, 1)
        rate = PAIR_RMU_VALU_WAVE_INSTR / (R["t_kern"] * 1e-3)
        line["roofline"]["issue"] = {"valu_wave_instr": PAIR_RMU_VALU_WAVE_INSTR, "rate": rate,
                                     "ceiling": c["wave_instr_per_s"], "frac": rate / c["wave_instr_per_s"],
                                     "unit": "VALU wave-instructions/s",
                                     "ceiling_kernel": "k_valu_fma32 (FP32 FMA stream, %.1f TFLOP/s)" % c["tflops"]}
    if "fp64" in res and main_prec != "fp64":
        F = res["fp64"]
        line["fp64"] = {"value": F["out"]["stats"]["n_binned"] / F["t_step"], "ms_per_step": F["t_step"] * 1e3,
                        "kernel_ms": F["t_kern"], "speedup_fp32_vs_fp64": F["t_step"] / R["t_step"]}
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(xyz, w, nd)
    if line["cpu_baseline"]:
        line["gpu_vs_cpu"] = value / line["cpu_baseline"]["value"]
    eng.close()
    return line


def bench_c1(args):
    """The plain CPU configuration (20k + 40k uniform in 500^3, 10<r<50, 10
    bins, nmu = 1): the reference runs the WHOLE workload here."""
    import numpy as np
    import torch
    from gramsci_amd import catalog as C
    from gramsci_amd.engine import Engine
    d = C.uniform(20000, 500.0, 101)
    r = C.uniform(40000, 500.0, 102)
    xyz = np.concatenate([d, r])
    eng = Engine(0)
    dxyz = torch.from_numpy(xyz).cuda()
    p = dict(rmin=10.0, rmax=50.0, nbins=10, nmu=1, log_bins=False)
    for _ in range(args.warmup):
        eng.set_catalogue(dxyz, None, 20000)
        eng.pcf2(**p)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.set_catalogue(dxyz, None, 20000)
        out = eng.pcf2(**p)
    dt = (time.perf_counter() - t0) / args.steps
    U = out["stats"]["n_binned"]
    line = line_base("binned pair evals/s, plain 2PCF (config 1)", "pair evals/s", U / dt, args.steps, args.warmup,
                     dt, "f32", "2PCF: 20k + 40k uniform, [0,500)^3, 10<r<50, 10 bins, nmu=1",
                     binned_pairs_per_step=int(U))
    ref = None if args.no_cpu else run_reference(d, None, r, None, ["-rmin", "10", "-rmax", "50", "-nbins", "10",
                                                                    "-out", "o.txt", "-2pcf"])
    if ref:
        t, wall, cores = ref
        line["cpu_baseline"] = {"value": U / t, "unit": "pair evals/s", "cores": cores, "kind": "reference",
                                "sample": f"whole workload: {U} binned ordered pairs in {t:.3f} s graph+query "
                                          f"({wall:.3f} s wall with ASCII I/O); {_cpu_model()}"}
        line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    eng.close()
    return line


def bench_c3(args):
    """3PCF of the 3PCF/4PCF configuration: 200k NS-clustered + 400k uniform in
    500^3, 10<r<30, 5 bins; work unit = apex-pair candidate sum_i C(k_i, 2)."""
    import numpy as np
    import torch
    from gramsci_amd import catalog as C
    from gramsci_amd.engine import Engine
    d = C.neyman_scott(100000, 5000, 20, 2.0, 500.0, 301)
    r = C.uniform(400000, 500.0, 302)
    xyz = np.concatenate([d, r])
    eng = Engine(0)
    dxyz = torch.from_numpy(xyz).cuda()
    res = {}
    for prec in (["fp32", "fp64"] if not args.no_fp64 else ["fp32"]):
        for _ in range(args.warmup):
            eng.set_catalogue(dxyz, None, len(d))
            eng.pcf3(10.0, 30.0, 5, precision=prec)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.set_catalogue(dxyz, None, len(d))
            out = eng.pcf3(10.0, 30.0, 5, precision=prec)
        res[prec] = ((time.perf_counter() - t0) / args.steps, out)
        print(f"[bench] 3pcf {prec}: {res[prec][0] * 1e3:.1f} ms, kernel {out['stats']['ms_count']:.1f} ms, "
              f"units {out['stats']['n_units']:.4g} binned {out['stats']['n_binned']:.4g} "
              f"rechecks {out['stats']['n_recheck']}", file=sys.stderr, flush=True)
    dt, out = res["fp32"]
    U = out["stats"]["n_units"]
    line = line_base("apex-pair candidates/s, 3PCF (config 3)", "triplet evals/s", U / dt, args.steps, args.warmup,
                     dt, "f32", "3PCF: 200k NS + 400k uniform, 500^3, 10<r<30, 5 bins (35 configs x 4 classes)",
                     units_per_step=int(U), binned_triplets=int(out["stats"]["n_binned"]))
    line["roofline"] = {"bound": "valu", "achieved": U * 8 / (out["stats"]["ms_count"] * 1e-3) / 1e12,
                        "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                        "frac": U * 8 / (out["stats"]["ms_count"] * 1e-3) / PEAK_FP32, "traffic": None,
                        "kernel": "k_triplet3<false, false, 0, false, false>", "kernel_ms": out["stats"]["ms_count"],
                        "flops_per_unit": 8}
    if "fp64" in res:
        line["fp64"] = {"value": res["fp64"][1]["stats"]["n_units"] / res["fp64"][0],
                        "ms_per_step": res["fp64"][0] * 1e3}
        assert np.array_equal(res["fp64"][1]["cnt"], out["cnt"])
    if not args.no_cpu:
        # density-matched 150^3 sub-box (the reference needs ~10 s for it on 8 cores)
        side, lo = 150.0, 175.0
        md = np.all((d >= lo) & (d < lo + side), axis=1)
        mr = np.all((r >= lo) & (r < lo + side), axis=1)
        ref = run_reference(d[md], None, r[mr], None, ["-rmin", "10", "-rmax", "30", "-nbins", "5", "-out", "o.txt",
                                                       "-3pcf"])
        if ref:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            sub = np.concatenate([d[md], r[mr]])
            Us = O.pcf3(sub, None, int(md.sum()), O.Params(10.0, 30.0, 5))["n_cand"]
            t, wall, cores = ref
            line["cpu_baseline"] = {"value": Us / t, "unit": "triplet evals/s", "cores": cores, "kind": "reference",
                                    "sample": f"density-matched {side:.0f}^3 sub-box ({int(md.sum())} + "
                                              f"{int(mr.sum())} points): {Us} apex-pair candidates in {t:.2f} s "
                                              f"graph+query ({wall:.2f} s wall); {_cpu_model()}"}
            line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    eng.close()
    return line


def bench_c3_4pcf(args):
    """4PCF on the 50k-galaxy + 100k-random subset of the 3PCF configuration,
    10<r<20, 3 bins; work unit = evaluated quad."""
    import numpy as np
    import torch
    from gramsci_amd import catalog as C
    from gramsci_amd.engine import Engine
    d = C.neyman_scott(100000, 5000, 20, 2.0, 500.0, 301)
    r = C.uniform(400000, 500.0, 302)
    rng = np.random.default_rng(303)
    d = d[np.sort(rng.choice(len(d), 50000, replace=False))]
    r = r[np.sort(rng.choice(len(r), 100000, replace=False))]
    xyz = np.concatenate([d, r])
    eng = Engine(0)
    dxyz = torch.from_numpy(xyz).cuda()
    for _ in range(args.warmup):
        eng.set_catalogue(dxyz, None, len(d))
        eng.pcf4(10.0, 20.0, 3)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.set_catalogue(dxyz, None, len(d))
        out = eng.pcf4(10.0, 20.0, 3)
    dt = (time.perf_counter() - t0) / args.steps
    U = out["stats"]["n_units"]
    line = line_base("evaluated quads/s, 4PCF (config 3 subset)", "quad evals/s", U / dt, args.steps, args.warmup,
                     dt, "f64", "4PCF: 50k NS + 100k uniform subset, 500^3, 10<r<20, 3 bins",
                     units_per_step=int(U), binned_quads=int(out["stats"]["n_binned"]))
    if not args.no_cpu:
        ref = run_reference(d, None, r, None, ["-rmin", "10", "-rmax", "20", "-nbins", "3", "-out", "o.txt", "-4pcf"])
        if ref:
            t, wall, cores = ref
            line["cpu_baseline"] = {"value": U / t, "unit": "quad evals/s", "cores": cores, "kind": "reference",
                                    "sample": f"whole workload: {U} quads in {t:.2f} s graph+query "
                                              f"({wall:.2f} s wall); {_cpu_model()}"}
            line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    eng.close()
    return line


def bench_c4(args):
    """64 uneven domain-decomposed regions with rmax buffer shells: one
    hipGraph-captured persistent launch vs one plain launch per region."""
    import numpy as np
    from gramsci_amd import catalog as C
    from gramsci_amd.engine import Engine
    xyz, buf, off, ndata = C.regions(64, 401, rmax=100.0)
    eng = Engine(0)
    p = dict(rmin=5.0, rmax=100.0, nbins=20, nmu=10, log_bins=False)
    eng.set_catalogue_batched(xyz, None, buf, off, ndata)
    t0 = time.perf_counter()
    eng.pcf2_batched(**p)          # includes capture and autotuning
    t_first = time.perf_counter() - t0
    for _ in range(args.warmup):
        eng.pcf2_batched(**p)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = eng.pcf2_batched(**p)
    dt = (time.perf_counter() - t0) / args.steps
    U = out["stats"]["n_binned"]
    # 64 plain launches (catalogue already on the device for each region)
    import torch
    regs = [(torch.from_numpy(xyz[off[k]:off[k + 1]]).cuda(), torch.from_numpy(buf[off[k]:off[k + 1]]).cuda(),
             int(ndata[k])) for k in range(64)]
    sep_cnt = np.zeros_like(out["cnt"])
    t0 = time.perf_counter()
    for k, (x, b, nd) in enumerate(regs):
        eng.set_catalogue(x, None, nd, b)
        sep_cnt[k] = eng.pcf2(**p, cell_div=out["stats"]["cell_div"])["cnt"]
    t_sep = time.perf_counter() - t0
    assert np.array_equal(sep_cnt, out["cnt"]), "batched counts differ from separate launches"
    line = line_base("binned pair evals/s, 64 batched regions (config 4, 2PCF)", "pair evals/s", U / dt, args.steps,
                     args.warmup, dt, "f32", "64 regions, lognormal sizes 5k-400k, 2x randoms, rmax shells; "
                     "2PCF 5<r<100, 20 bins, nmu=10", n_points=int(len(xyz)), binned_pairs_per_step=int(U),
                     cells_per_rmax=int(out["stats"]["cell_div"]))
    line["separate_launches"] = {"value": U / t_sep, "seconds": t_sep, "speedup_batched": t_sep / dt}
    line["first_call_s"] = t_first
    # the configuration's 3PCF leg: 5 < r < 25, 4 bins, all regions in one captured triplet launch
    p3 = dict(rmin=5.0, rmax=25.0, nbins=4)
    eng.pcf3_batched(**p3)         # capture and autotuning
    for _ in range(args.warmup):
        eng.pcf3_batched(**p3)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        o3 = eng.pcf3_batched(**p3)
    dt3 = (time.perf_counter() - t0) / args.steps
    sep3 = np.zeros_like(o3["cnt"][:, :, 0, :])
    t0 = time.perf_counter()
    for k, (x, b, nd) in enumerate(regs):
        eng.set_catalogue(x, None, nd, b)
        sep3[k] = eng.pcf3(**p3, cell_div=o3["stats"]["cell_div"])["cnt"]
    t_sep3 = time.perf_counter() - t0
    assert np.array_equal(sep3, o3["cnt"][:, :, 0, :]), "batched 3PCF counts differ from separate launches"
    U3 = o3["stats"]["n_units"]
    line["3pcf"] = {"value": U3 / dt3, "unit": "apex-pair candidates/s", "ms_per_step": dt3 * 1e3,
                    "binned_triplets_per_step": int(o3["stats"]["n_binned"]), "units_per_step": int(U3),
                    "config": "3PCF 5<r<25, 4 bins", "cells_per_rmax": int(o3["stats"]["cell_div"]),
                    "separate_launches": {"value": U3 / t_sep3, "seconds": t_sep3, "speedup_batched": t_sep3 / dt3}}
    eng.close()
    return line


def bench_c5(args):
    """Differentiable weighted 2PCF: 500k NS galaxies (w = exp(theta.t), 8 GRF
    templates) + 1.5M randoms, 20 bins to r=120, nmu=10: forward + adjoint."""
    import numpy as np
    import torch
    from gramsci_amd import catalog as C
    from gramsci_amd.autograd import weighted_xi
    d = C.neyman_scott(250000, 12500, 20, 2.0, 1000.0, 501)
    r = C.uniform(1500000, 1000.0, 502)
    fields = C.grf_templates(8, 128, 1000.0, 50.0, 503)
    t = C.sample_trilinear(fields, d, 1000.0)
    theta0 = np.random.default_rng(504).normal(0, 0.05, 8)
    xyz = torch.from_numpy(np.concatenate([d, r])).cuda()
    tt = torch.from_numpy(t).cuda()
    ones = torch.ones(len(r), dtype=torch.float64, device="cuda")
    target = torch.zeros((20, 10), dtype=torch.float64, device="cuda")

    def step():
        theta = torch.from_numpy(theta0).cuda().requires_grad_(True)
        # theta . t per galaxy as a broadcast product and row sum: rocBLAS's gemv for this
        # 500k x 8 shape took 9 ms per call, the reduction takes microseconds
        w = torch.cat([torch.exp((tt * theta).sum(dim=1)), ones])
        xi = weighted_xi(w, xyz, len(d), 0.0, 120.0, 20, 10)
        L = ((xi - target) ** 2).sum()
        L.backward()
        torch.cuda.synchronize()
        return theta.grad

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = step()
    dt = (time.perf_counter() - t0) / args.steps
    from gramsci_amd.autograd import _engine
    eng = _engine(0)
    eng.set_catalogue(xyz, torch.cat([torch.exp((tt * torch.from_numpy(theta0).cuda()).sum(dim=1)), ones]), len(d))
    U = eng.pcf2(0.0, 120.0, 20, 10)["stats"]["n_binned"]
    line = line_base("binned pair evals/s, forward + adjoint (config 5)", "pair evals/s", 2 * U / dt, args.steps,
                     args.warmup, dt, "f32", "weighted 2PCF + adjoint: 500k NS (w=exp(theta.t)) + 1.5M randoms, "
                     "0<r<120, 20 bins, nmu=10", binned_pairs=int(U), dL_dtheta=[float(x) for x in g.cpu()])
    return line


def bench_copy(args):
    """Plain 1 GiB device-to-device copy (calibrates the HBM byte counters)."""
    import torch
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return line_base("device copy bandwidth", "GB/s", 2 * 8 * n / dt / 1e9, args.steps, args.warmup, dt, "f64",
                     "copy of 2^27 float64 (1 GiB read + 1 GiB written)", bytes_read=8 * n, bytes_written=8 * n)


def valu_ceiling(steps, warmup):
    """Times the FP32 FMA issue-rate calibration kernel (gramsci_amd/libgs_calib.so: 8 independent FMA
    chains per lane, 8 work-groups of 256 threads per CU, no memory traffic). Returns a dict with the mean
    launch time (s), the VALU wave-instructions per launch and the resulting rates."""
    import ctypes
    import torch
    so = os.path.join(ROOT, "gramsci_amd", "libgs_calib.so")
    if not os.path.exists(so):
        raise RuntimeError(f"{so} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(so)
    lib.gs_calib_fma32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_float)]
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    threads, per_iter = lib.gs_calib_threads(), lib.gs_calib_fma_per_iter()
    blocks, iters = n_cu * 8 * 4, 4096
    out = torch.zeros(blocks * threads, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ms = ctypes.c_float(0.0)
    times = []
    for k in range(warmup + steps):
        rc = lib.gs_calib_fma32(blocks, iters, ctypes.c_void_p(out.data_ptr()), out.numel(), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(f"gs_calib_fma32 failed: {rc}")
        if k >= warmup:
            times.append(ms.value * 1e-3)
    dt = sum(times) / len(times)
    wave_instr = blocks * (threads // 64) * iters * per_iter
    return {"seconds": dt, "wave_instr": wave_instr, "wave_instr_per_s": wave_instr / dt,
            "tflops": 2 * 64 * wave_instr / dt / 1e12, "n_cu": n_cu, "blocks": blocks, "iters": iters,
            "per_iter": per_iter}


def bench_valu(args):
    """FP32 FMA issue-rate calibration: the practical VALU ceiling that the counting kernels'
    instruction issue is read against."""
    c = valu_ceiling(args.steps, args.warmup)
    line = line_base("FP32 FMA issue rate (VALU calibration)", "TFLOP/s", c["tflops"], args.steps, args.warmup,
                     c["seconds"], "f32", f"{c['blocks']} work-groups x 256 threads, {c['iters']} x {c['per_iter']} "
                     "dependent-chain FMAs per thread (8 chains), packed FP32 off", n_cu=c["n_cu"])
    line["valu_wave_instr_per_launch"] = c["wave_instr"]
    line["valu_wave_instr_per_s"] = c["wave_instr_per_s"]
    line["data"] = "none (register-only kernel)"
    return line


WORKLOADS = {"copy": bench_copy, "valu": bench_valu, "rmu": bench_rmu, "c1": bench_c1, "c3": bench_c3, "c3_4pcf": bench_c3_4pcf, "c4": bench_c4,
             "c5": bench_c5}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="rmu",
                    help="rmu (default): the (r, mu) 2PCF configuration the metric is quoted on")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--fp64-only", action="store_true", help="time only the plain FP64 path (profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (N=1); multi-GPU sharding is a design note, not built.",
              file=sys.stderr)
        sys.exit(2)
    line = WORKLOADS[args.workload](args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
