"""The Fortran host program (host/gramsci) as a drop-in for the reference's
command line: same flags, same output files.

CPU tests exercise option handling only (the engine is never reached).  GPU
tests run the host on the golden fixtures' catalogues with the reference's
arguments and compare the output file with the reference's own output:
header lines and bin-edge / mu columns byte for byte, N2 / N3 / ratio columns
to the last printed digit.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from golden_util import close_e147, load_input, manifest, parse_rows, read_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "host", "gramsci")
MAN = manifest()

pytestmark = pytest.mark.skipif(not os.path.exists(HOST), reason="host/gramsci not built")


def run(args, cwd):
    return subprocess.run([HOST] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_too_few_arguments_prints_help(tmp_path):
    r = run(["-gal", "x"], tmp_path)
    assert r.returncode == 0
    assert "Not enough input parameters" in r.stdout


def test_unknown_option_stops(tmp_path):
    r = run(["-gal", "a", "-rmin", "1", "-rmax", "2", "-bogus"], tmp_path)
    assert "unknown option -bogus ignored" in r.stdout


def test_options_match_on_six_characters(tmp_path):
    # the reference reads options into a character(len=6): "-nbinsXYZ" is "-nbins",
    # so parsing goes on to the (missing) input file instead of stopping on the option
    r = run(["-gal", "missing.txt", "-rmin", "1", "-rmax", "2", "-nbinsXYZ", "3", "-equi"], tmp_path)
    assert "unknown option" not in r.stdout
    assert r.returncode != 0


def _args(c, gal, ran, out):
    p = c["params"]
    a = ["-gal", gal]
    if ran:
        a += ["-ran", ran]
    a += ["-rmin", repr(float(p["rmin"])), "-rmax", repr(float(p["rmax"])), "-nbins", str(p["nbins"]),
          "-nmu", str(p["nmu"]), "-out", out]
    if p["log_bins"]:
        a.append("-log")
    a.append(c["mode"])
    return a


def _compare(name, run_name, text):
    ref = read_text(name, run_name)
    rl, gl = ref.splitlines(), text.splitlines()
    assert len(rl) == len(gl)
    ncols_fixed = {"-2pcf": 4, "-3pcf": 6, "-4pcf": 8, "-equi": 4, "-tetra": 4}[MAN[name]["mode"]]
    if MAN[name]["mode"] == "-3pcf" and MAN[name]["params"]["nmu"] > 1:
        ncols_fixed = 8            # RSD rows add mu min, mu max
    identical = 0
    for a, b in zip(rl, gl):
        ta, tb = a.split(), b.split()
        try:
            [float(t) for t in ta]
        except ValueError:
            assert a == b          # header line, byte for byte
            identical += 1
            continue
        assert ta[:ncols_fixed] == tb[:ncols_fixed]
        assert len(a) == len(b)
        identical += a == b
    rows_r, rows_g = parse_rows(ref), parse_rows(text)
    ok = close_e147(rows_r[:, ncols_fixed:ncols_fixed + 2], rows_g[:, ncols_fixed:ncols_fixed + 2])
    assert ok.all(), np.argwhere(~ok)[:5]
    return identical / len(rl)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MAN))
def test_host_output_matches_reference_file(name, tmp_path):
    c = MAN[name]
    d, r, wd, wr = load_input(name)
    O.write_catalogue(tmp_path / "d.txt", d, wd)
    ran = None
    if len(r):
        O.write_catalogue(tmp_path / "r.txt", r, wr)
        ran = "r.txt"
    run_name = "DR" if "DR" in c["runs"] else "D"
    if run_name == "D":
        ran = None
    res = run(_args(c, "d.txt", ran, "out.txt"), tmp_path)
    print(res.stdout[-400:], flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    frac = _compare(name, run_name, (tmp_path / "out.txt").read_text())
    print(f"{name}: {100 * frac:.1f}% of lines byte-identical", flush=True)


@pytest.mark.gpu
def test_host_fp64_flag_and_cut_mode(tmp_path):
    name = "c1_lin"
    c = MAN[name]
    d, r, _, _ = load_input(name)
    O.write_catalogue(tmp_path / "d.txt", d)
    O.write_catalogue(tmp_path / "r.txt", r)
    a = _args(c, "d.txt", "r.txt", "o32.txt")
    assert run(a, tmp_path).returncode == 0
    b = [x if x != "o32.txt" else "o64.txt" for x in a[:-1]] + ["-fp64", "-2pcf"]
    assert run(b, tmp_path).returncode == 0
    assert (tmp_path / "o32.txt").read_text() == (tmp_path / "o64.txt").read_text()
    # -cut: x y z w buffer, no normalisation, buffer points are never apexes; the
    # result is written over the input file (gramsci.F90:1224-1231)
    rng = np.random.default_rng(3)
    buf = (rng.random(len(d)) < 0.3).astype(int)
    O.write_catalogue(tmp_path / "cut.txt", d, None, extra=buf)
    p = O.Params(**c["params"])
    res = run(["-cut", "cut.txt", "-rmin", "2.0", "-rmax", "10.0", "-nbins", "10", "-2pcf"], tmp_path)
    assert res.returncode == 0, res.stdout
    rows = parse_rows((tmp_path / "cut.txt").read_text())
    ref = O.pcf2(d, None, len(d), p, buf=buf.astype(np.int32))
    # un-normalised: N2 is the plain ordered-pair count with unit weights
    assert np.array_equal(np.rint(rows[:, 4]).astype(np.int64), ref["cnt"][0, :, 0].astype(np.int64))
