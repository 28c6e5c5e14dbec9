"""Benchmark of the counting hot path on one MI355X.

Default line: the (r, mu) 2PCF GPU configuration, 1,000,000 clustered galaxies
in a 1000^3 box (500,000 Poisson points + 25,000 Neyman-Scott clusters of 20
members, Gaussian sigma = 2) with lognormal(0, 0.2) weights, plus 3,000,000
uniform randoms (unit weights); rmin = 1, rmax = 150, 30 log-spaced bins,
nmu = 20.

A step is one full pass of the hot path with the catalogue already resident
in HBM: device-to-device catalogue hand-over, bounding box, GPU cell grid
(counting sort), and the persistent pair kernel, ending with the histograms
on the host.  value = binned ordered pairs (the reference's N2 increments)
per second of the FP32 + FP64-guard path; the plain FP64 path is timed on the
same input and reported beside it.  The same line carries the triplet leg
(3PCF configuration, "triplet") and the batched-regions legs (64 regions,
2PCF and 3PCF, "batched"), each with its own step count.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import shutil
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "binned pair/triplet evals/s on 1 MI355X: batched FP32+guard vs plain FP64 vs CPU"
# Counter readings of the pair kernel on the default workload (rocprofv3 --pmc runs whose commands
# the file lists): HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE (scaled by the counters'
# readings for a copy of known size) and VALU wave-instructions per dispatch (SQ_INSTS_VALU)
PAIR_RMU_COUNTERS = os.path.join(ROOT, "profiles", "k_pair2_rmu_counters.json")
# the sources the pair kernel's code object is built from: the profile records their hash
PAIR_KERNEL_SOURCES = ["gramsci_amd/csrc/pair2.hip", "gramsci_amd/csrc/gs_device.h", "gramsci_amd/csrc/gs_common.h",
                       "gramsci_amd/csrc/gs_kernels.h"]
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


def kernel_source_hash(paths=PAIR_KERNEL_SOURCES):
    import hashlib
    h = hashlib.sha256()
    for p in paths:
        with open(os.path.join(ROOT, p), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def pair_counters(kernel_ms_live):
    """The pair kernel's counter readings from PAIR_RMU_COUNTERS, with 'stale' set when they
    describe another build: the kernel sources' hash differs from the one recorded with the
    profile, or the kernel time timed now differs from the profiled dispatch time by more than
    5 %.  Stale readings are not used for this line."""
    try:
        with open(PAIR_RMU_COUNTERS) as f:
            c = json.load(f)
    except (OSError, ValueError):
        return None
    c["stale"] = (c.get("source_sha256") != kernel_source_hash()
                  or abs(kernel_ms_live - c["kernel_ms"]) > 0.05 * c["kernel_ms"])
    c["source"] = os.path.relpath(PAIR_RMU_COUNTERS, ROOT)
    return c


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


def _host_threads():
    """OpenMP threads for the reference: OMP_NUM_THREADS where the environment sets it (the host's
    CPU share), else every core this process may run on."""
    visible = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else visible
    return min(threads, visible), visible


def run_reference(d, wd, r, wr, args, runs=3, warmup=1, timeout=900):
    """Runs the reference program (oracle/_ref/gramsci_fast: the reference sources built with its own
    -O3 -ffast-math) on the host: `warmup` untimed runs, then `runs` timed runs on the same files.
    Returns a dict with the median graph+query time (the reference's own phase prints), the
    per-run times, the wall times, the threads used and the host's core count, or None."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.REF_BIN_FAST if os.path.exists(O.REF_BIN_FAST) else O.REF_BIN
    if not os.path.exists(ref):
        return None
    threads, visible = _host_threads()
    times, walls = [], []
    with tempfile.TemporaryDirectory() as td:
        O.write_catalogue(os.path.join(td, "d.txt"), d, wd)
        a = ["-gal", "d.txt"]
        if r is not None:
            O.write_catalogue(os.path.join(td, "r.txt"), r, wr)
            a += ["-ran", "r.txt"]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        for k in range(warmup + runs):
            t0 = time.perf_counter()
            res = subprocess.run([ref] + a + args, cwd=td, env=env, capture_output=True, text=True,
                                 timeout=timeout)
            wall = time.perf_counter() - t0
            if res.returncode != 0:
                return None
            tg, tq = _ref_phase_times(res.stdout)
            t = (tg or 0.0) + (tq or 0.0)
            if k >= warmup:
                times.append(t if t > 0 else wall)
                walls.append(wall)
    return {"median_s": statistics.median(times), "times_s": times, "wall_median_s": statistics.median(walls),
            "threads": threads, "host_cores": visible, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "runs": runs, "warmup": warmup}


def _baseline(value, unit, ref, sample):
    return {"value": value, "unit": unit, "cores": ref["threads"], "kind": "reference",
            "sample": f"{sample}; median of {ref['runs']} timed runs after {ref['warmup']} warm-up "
                      f"(graph+query {', '.join(f'{t:.2f}' for t in ref['times_s'])} s; "
                      f"wall {ref['wall_median_s']:.2f} s with ASCII I/O); {ref['threads']} OpenMP threads "
                      f"of {ref['host_cores']} cores available ({ref['nproc']} in the machine), {ref['cpu_model']}"}


def cpu_baseline(xyz, w, nd):
    """The reference program on a density-matched 150^3 sub-box of the same catalogue, same bins.
    Throughput = binned ordered pairs / median graph + query time.  The reference's cost per pair
    grows with the neighbour count k (per-point neighbour lists), and k is several times smaller in
    the sub-box than in the full box (rmax = 150 is the sub-box side), so this over-states its
    throughput on the full workload."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    side, lo = 150.0, 425.0
    inb = np.all((xyz >= lo) & (xyz < lo + side), axis=1)
    idx = np.nonzero(inb)[0]
    d_idx, r_idx = idx[idx < nd], idx[idx >= nd]
    ref = run_reference(xyz[d_idx], w[d_idx], xyz[r_idx], w[r_idx],
                        ["-rmin", "1.0", "-rmax", "150.0", "-nbins", "30", "-nmu", "20", "-log", "-out", "o.txt",
                         "-2pcf"])
    if not ref:
        return None
    sub = np.concatenate([xyz[d_idx], xyz[r_idx]])
    pairs = int(O.pcf2(sub, None, len(d_idx), O.Params(**PARAMS))["cnt"].sum())
    return _baseline(pairs / ref["median_s"], "pair evals/s", ref,
                     f"reference gramsci (-O3 -ffast-math build) on a density-matched {side:.0f}^3 sub-box "
                     f"({len(d_idx)} galaxies + {len(r_idx)} randoms) of the same catalogue, same bins; "
                     f"{pairs} binned ordered pairs per run")


def line_base(metric, unit, value, steps, warmup, t_step, dtype, workload, **cfg):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (seeded Neyman-Scott + uniform, numpy PCG64)",
            "config": dict(workload=workload, **cfg)}


def bench_rmu(args, legs=True):
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
              f"binned {out['stats']['n_binned']:.4g}, candidates {out['stats']['n_candidates']:.4g}, "
              f"rechecks {out['stats']['n_recheck']}", file=sys.stderr,
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
                     candidate_pairs_per_step=int(R["out"]["stats"]["n_candidates"]),
                     precision="FP32 + FP64 guard-band recheck" if main_prec == "fp32" else "plain FP64")
    line["roofline"] = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / peak, "traffic": None,
                        # the FP32 kernel on this many bins runs two apexes per lane (the fifth argument),
                        # which the library picks where it costs no resident waves
                        "kernel": ("k_pair2<true, true, true, true, false>" if main_prec == "fp64"
                                   else "k_pair2<false, true, true, true, true>"),
                        "kernel_ms": R["t_kern"], "flops_per_pair": FLOPS_PER_PAIR_RMU}
    if main_prec == "fp32":
        c = pair_counters(R["t_kern"])
        if c is not None:
            line["roofline"]["traffic_source"] = c["source"]
            line["roofline"]["traffic_stale"] = c["stale"]
            if not c["stale"]:
                line["roofline"]["traffic"] = (c["FETCH_SIZE_KiB"] * c["FETCH_SIZE_scale"] +
                                               c["WRITE_SIZE_KiB"] * c["WRITE_SIZE_scale"]) * 1024
                # instruction issue of the same launch against a pure FP32 FMA stream timed now: a
                # diagnostic of how busy the vector ALUs are, not a roofline
                fma = valu_ceiling(3, 1)
                rate = c["SQ_INSTS_VALU"] / (R["t_kern"] * 1e-3)
                line["valu_issue"] = {"wave_instr_per_launch": c["SQ_INSTS_VALU"], "rate": rate,
                                      "fma_stream_rate": fma["wave_instr_per_s"],
                                      "ratio": rate / fma["wave_instr_per_s"], "unit": "VALU wave-instructions/s",
                                      "lane_instr_per_binned_pair": 64.0 * c["SQ_INSTS_VALU"] / U}
    if "fp64" in res and main_prec != "fp64":
        F = res["fp64"]
        line["fp64"] = {"value": F["out"]["stats"]["n_binned"] / F["t_step"], "ms_per_step": F["t_step"] * 1e3,
                        "kernel_ms": F["t_kern"], "speedup_fp32_vs_fp64": F["t_step"] / R["t_step"]}
    eng.close()
    del dxyz, dw
    if legs and not args.fp64_only:
        steps = max(1, min(args.steps, 10))
        line["triplet"] = triplet_leg(steps, min(args.warmup, 2), not args.no_fp64)
        line["batched"] = batched_legs(steps, min(args.warmup, 2))
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(xyz, w, nd)
    if line["cpu_baseline"]:
        line["gpu_vs_cpu"] = value / line["cpu_baseline"]["value"]
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
        line["cpu_baseline"] = _baseline(U / ref["median_s"], "pair evals/s", ref,
                                         f"whole workload: {U} binned ordered pairs per run")
        line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    eng.close()
    return line


def _c3_catalogue():
    from gramsci_amd import catalog as C
    return C.neyman_scott(100000, 5000, 20, 2.0, 500.0, 301), C.uniform(400000, 500.0, 302)


def triplet_leg(steps, warmup, with_fp64=True):
    """3PCF of the 3PCF/4PCF configuration (200k NS-clustered + 400k uniform in 500^3, 10<r<30, 5 bins):
    FP32 + guard and plain FP64; unit = apex-pair candidate sum_i C(k_i, 2)."""
    import numpy as np
    import torch
    from gramsci_amd.engine import Engine
    d, r = _c3_catalogue()
    xyz = np.concatenate([d, r])
    eng = Engine(0)
    dxyz = torch.from_numpy(xyz).cuda()
    res = {}
    for prec in (["fp32", "fp64"] if with_fp64 else ["fp32"]):
        for _ in range(warmup):
            eng.set_catalogue(dxyz, None, len(d))
            eng.pcf3(10.0, 30.0, 5, precision=prec)
        kern = []
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.set_catalogue(dxyz, None, len(d))
            out = eng.pcf3(10.0, 30.0, 5, precision=prec)
            kern.append(out["stats"]["ms_count"])
        res[prec] = ((time.perf_counter() - t0) / steps, out, sum(kern) / len(kern))
        print(f"[bench] 3pcf {prec}: {res[prec][0] * 1e3:.1f} ms, kernel {res[prec][2]:.1f} ms, "
              f"units {out['stats']['n_units']:.4g} evaluated {out['stats']['n_evaluated']:.4g} "
              f"binned {out['stats']['n_binned']:.4g} "
              f"rechecks {out['stats']['n_recheck']}", file=sys.stderr, flush=True)
    eng.close()
    dt, out, kms = res["fp32"]
    U = out["stats"]["n_units"]
    E = out["stats"]["n_evaluated"]
    # value: the reference's work (apex-pair candidates, sum of C(k, 2)) per second; the roofline counts
    # the pairs the kernel evaluated, about a quarter of them (each triangle is visited once)
    leg = {"metric": "apex-pair candidates/s, 3PCF", "value": U / dt, "unit": "triplet evals/s",
           "steps": steps, "warmup": warmup, "ms_per_step": dt * 1e3, "kernel_ms": kms,
           "workload": "3PCF: 200k NS + 400k uniform, 500^3, 10<r<30, 5 bins (35 configs x 4 classes)",
           "units_per_step": int(U), "evaluated_pairs_per_step": int(E),
           "binned_triplets": int(out["stats"]["n_binned"]),
           "roofline": {"bound": "valu", "achieved": E * 8 / (kms * 1e-3) / 1e12, "peak": PEAK_FP32 / 1e12,
                        "unit": "TFLOP/s", "frac": E * 8 / (kms * 1e-3) / PEAK_FP32, "flops_per_unit": 8,
                        "units": "evaluated neighbour pairs",
                        "kernel": "k_triplet3 (FP32 + guard, two passes)"}}
    if "fp64" in res:
        assert np.array_equal(res["fp64"][1]["cnt"], out["cnt"]), "3PCF fp32/fp64 counts differ"
        leg["fp64"] = {"value": res["fp64"][1]["stats"]["n_units"] / res["fp64"][0],
                       "ms_per_step": res["fp64"][0] * 1e3, "kernel_ms": res["fp64"][2],
                       "speedup_fp32_vs_fp64": res["fp64"][0] / dt}
    return leg


def bench_c3(args):
    """3PCF of the 3PCF/4PCF configuration: 200k NS-clustered + 400k uniform in
    500^3, 10<r<30, 5 bins; work unit = apex-pair candidate sum_i C(k_i, 2)."""
    import numpy as np
    leg = triplet_leg(args.steps, args.warmup, not args.no_fp64)
    line = line_base("apex-pair candidates/s, 3PCF (config 3)", "triplet evals/s", leg["value"], args.steps,
                     args.warmup, leg["ms_per_step"] * 1e-3, "f32", leg["workload"],
                     units_per_step=leg["units_per_step"],
                     evaluated_pairs_per_step=leg["evaluated_pairs_per_step"],
                     binned_triplets=leg["binned_triplets"])
    line["roofline"] = dict(leg["roofline"], kernel_ms=leg["kernel_ms"], traffic=None)
    if "fp64" in leg:
        line["fp64"] = leg["fp64"]
    if not args.no_cpu:
        # density-matched 150^3 sub-box (the reference needs ~10 s for it on 8 cores)
        d, r = _c3_catalogue()
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
            line["cpu_baseline"] = _baseline(Us / ref["median_s"], "triplet evals/s", ref,
                                             f"density-matched {side:.0f}^3 sub-box ({int(md.sum())} + "
                                             f"{int(mr.sum())} points): {Us} apex-pair candidates per run")
            line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    return line


def bench_c3_4pcf(args):
    """4PCF on the 50k-galaxy + 100k-random subset of the 3PCF configuration,
    10<r<20, 3 bins; work unit = evaluated quad."""
    import numpy as np
    import torch
    from gramsci_amd.engine import Engine
    d, r = _c3_catalogue()
    rng = np.random.default_rng(303)
    d = d[np.sort(rng.choice(len(d), 50000, replace=False))]
    r = r[np.sort(rng.choice(len(r), 100000, replace=False))]
    xyz = np.concatenate([d, r])
    eng = Engine(0)
    dxyz = torch.from_numpy(xyz).cuda()
    for _ in range(args.warmup):
        eng.set_catalogue(dxyz, None, len(d))
        eng.pcf4(10.0, 20.0, 3)
    kern = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.set_catalogue(dxyz, None, len(d))
        out = eng.pcf4(10.0, 20.0, 3)
        kern.append(out["stats"]["ms_count"])
    dt = (time.perf_counter() - t0) / args.steps
    kms = sum(kern) / len(kern)
    U = out["stats"]["n_units"]
    line = line_base("evaluated quads/s, 4PCF (config 3 subset)", "quad evals/s", U / dt, args.steps, args.warmup,
                     dt, "f64", "4PCF: 50k NS + 100k uniform subset, 500^3, 10<r<20, 3 bins",
                     units_per_step=int(U), binned_quads=int(out["stats"]["n_binned"]))
    line["roofline"] = {"bound": "valu", "achieved": U * 16 / (kms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                        "unit": "TFLOP/s", "frac": U * 16 / (kms * 1e-3) / PEAK_FP64, "traffic": None,
                        "kernel": "k_quad4 (FP64, two passes)", "kernel_ms": kms, "flops_per_unit": 16}
    if not args.no_cpu:
        ref = run_reference(d, None, r, None, ["-rmin", "10", "-rmax", "20", "-nbins", "3", "-out", "o.txt", "-4pcf"])
        if ref:
            line["cpu_baseline"] = _baseline(U / ref["median_s"], "quad evals/s", ref,
                                             f"whole workload: {U} quads per run")
            line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    eng.close()
    return line


def batched_legs(steps, warmup):
    """64 uneven domain-decomposed regions with rmax buffer shells (config 4): the 2PCF leg (5<r<100,
    20 bins, nmu = 10) and the 3PCF leg (5<r<25, 4 bins), each as one hipGraph-captured persistent
    launch over all regions, against one plain launch per region (counts must agree)."""
    import numpy as np
    import torch
    from gramsci_amd import catalog as C
    from gramsci_amd.engine import Engine
    xyz, buf, off, ndata = C.regions(64, 401, rmax=100.0)
    eng = Engine(0)
    p = dict(rmin=5.0, rmax=100.0, nbins=20, nmu=10, log_bins=False)
    eng.set_catalogue_batched(xyz, None, buf, off, ndata)
    t0 = time.perf_counter()
    eng.pcf2_batched(**p)          # includes capture and autotuning
    t_first = time.perf_counter() - t0
    for _ in range(warmup):
        eng.pcf2_batched(**p)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = eng.pcf2_batched(**p)
    dt = (time.perf_counter() - t0) / steps
    U = out["stats"]["n_binned"]
    regs = [(torch.from_numpy(xyz[off[k]:off[k + 1]]).cuda(), torch.from_numpy(buf[off[k]:off[k + 1]]).cuda(),
             int(ndata[k])) for k in range(64)]
    sep_cnt = np.zeros_like(out["cnt"])
    t0 = time.perf_counter()
    for k, (x, b, nd) in enumerate(regs):
        eng.set_catalogue(x, None, nd, b)
        sep_cnt[k] = eng.pcf2(**p, cell_div=out["stats"]["cell_div"])["cnt"]
    t_sep = time.perf_counter() - t0
    assert np.array_equal(sep_cnt, out["cnt"]), "batched counts differ from separate launches"
    legs = {"workload": "64 regions, lognormal sizes 5k-400k, 2x randoms, rmax shells", "n_points": int(len(xyz)),
            "steps": steps, "warmup": warmup,
            "2pcf": {"value": U / dt, "unit": "pair evals/s", "ms_per_step": dt * 1e3, "config": "5<r<100, 20 bins, nmu=10",
                     "binned_pairs_per_step": int(U), "cells_per_rmax": int(out["stats"]["cell_div"]),
                     "first_call_s": t_first,
                     "separate_launches": {"value": U / t_sep, "seconds": t_sep, "speedup_batched": t_sep / dt}}}
    print(f"[bench] batched 2pcf: {dt * 1e3:.1f} ms vs {t_sep * 1e3:.1f} ms in 64 launches", file=sys.stderr,
          flush=True)
    p3 = dict(rmin=5.0, rmax=25.0, nbins=4)
    eng.pcf3_batched(**p3)         # capture and autotuning
    for _ in range(warmup):
        eng.pcf3_batched(**p3)
    t0 = time.perf_counter()
    for _ in range(steps):
        o3 = eng.pcf3_batched(**p3)
    dt3 = (time.perf_counter() - t0) / steps
    sep3 = np.zeros_like(o3["cnt"][:, :, 0, :])
    t0 = time.perf_counter()
    for k, (x, b, nd) in enumerate(regs):
        eng.set_catalogue(x, None, nd, b)
        sep3[k] = eng.pcf3(**p3, cell_div=o3["stats"]["cell_div"])["cnt"]
    t_sep3 = time.perf_counter() - t0
    assert np.array_equal(sep3, o3["cnt"][:, :, 0, :]), "batched 3PCF counts differ from separate launches"
    U3 = o3["stats"]["n_units"]
    legs["3pcf"] = {"value": U3 / dt3, "unit": "apex-pair candidates/s", "ms_per_step": dt3 * 1e3,
                    "binned_triplets_per_step": int(o3["stats"]["n_binned"]), "units_per_step": int(U3),
                    "config": "5<r<25, 4 bins", "cells_per_rmax": int(o3["stats"]["cell_div"]),
                    "separate_launches": {"value": U3 / t_sep3, "seconds": t_sep3, "speedup_batched": t_sep3 / dt3}}
    print(f"[bench] batched 3pcf: {dt3 * 1e3:.1f} ms vs {t_sep3 * 1e3:.1f} ms in 64 launches", file=sys.stderr,
          flush=True)
    eng.close()
    return legs


def bench_c4(args):
    """64 uneven domain-decomposed regions with rmax buffer shells: one
    hipGraph-captured persistent launch vs one plain launch per region."""
    legs = batched_legs(args.steps, args.warmup)
    L2 = legs["2pcf"]
    line = line_base("binned pair evals/s, 64 batched regions (config 4, 2PCF)", "pair evals/s", L2["value"],
                     args.steps, args.warmup, L2["ms_per_step"] * 1e-3, "f32",
                     legs["workload"] + "; 2PCF 5<r<100, 20 bins, nmu=10", n_points=legs["n_points"],
                     binned_pairs_per_step=L2["binned_pairs_per_step"], cells_per_rmax=L2["cells_per_rmax"])
    line["separate_launches"] = L2["separate_launches"]
    line["first_call_s"] = L2["first_call_s"]
    line["3pcf"] = legs["3pcf"]
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
    ap.add_argument("--no-legs", action="store_true",
                    help="default workload only: skip the triplet and batched legs (profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (N=1); multi-GPU sharding is a design note, not built.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "rmu":
        line = bench_rmu(args, legs=not args.no_legs)
    else:
        line = WORKLOADS[args.workload](args)
    _end_children()
    print(json.dumps(line), flush=True)


def _end_children():
    """Every process this run started (the reference program runs) has ended when its subprocess call
    returned; any descendant still alive here is reported and ended, so that nothing outlives the run."""
    try:
        import psutil
    except ImportError:
        return
    left = psutil.Process().children(recursive=True)
    for ch in left:
        print(f"[bench] ending leftover child process {ch.pid} ({ch.name()})", file=sys.stderr, flush=True)
        try:
            ch.terminate()
        except psutil.Error:
            pass
    if left:
        psutil.wait_procs(left, timeout=5)


if __name__ == "__main__":
    main()
