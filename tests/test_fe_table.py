"""The FP32 force pass's interval table (hostpot::build_fe_table), on the host:
a probe (tests/probe/fe_table_probe.cpp, built with hipcc) prints the table
for a knot set, and every bin record is checked against a brute-force
interval search -- the interval a radius r falls in is the number of
breakpoints (pair knots, density knots, the density maximum r*) above r --
at radii spread over each bin, and every interval's polynomials against the
oracle's FP64 knot sums V and phi' within the FP32 rounding of their
coefficients (5e-6 eV, eV/A).  Knot sets the bin grid cannot separate (two
breakpoints 0.003 A apart, closer than the finest 2^-8 A bin) or that put a
breakpoint on every grid's bin edge are refused, so the force pass falls
back to the knot sums."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from moldy_amd import potential

PROBE_SRC = os.path.join(ROOT, "tests", "probe", "fe_table_probe.cpp")
NB, NI = 64, 10   # MDK_FE_NB bins, MDK_FE_NI intervals


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("probe") / "fe_table_probe")
    subprocess.run([hipcc, "-O1", "-std=c++17", "--offload-arch=gfx950", "-x", "hip", PROBE_SRC, "-o", out],
                   check=True, capture_output=True)
    return out


def build(probe, c):
    txt = "%d %s %s %d %s %s\n" % (len(c["av"]), " ".join(repr(float(v)) for v in c["av"]),
                                   " ".join(repr(float(v)) for v in c["rv"]), len(c["ap"]),
                                   " ".join(repr(float(v)) for v in c["ap"]),
                                   " ".join(repr(float(v)) for v in c["rp"]))
    res = subprocess.run([probe], input=txt, capture_output=True, text=True, check=True)
    lines = res.stdout.split("\n")
    ok, binh, binc, pcr = lines[0].split()
    words = [ln.split() for ln in lines[1:1 + 4 * NB + 8 * NI]]
    f = np.array([float(w[0]) for w in words], dtype=np.float32)
    bits = np.array([int(w[1]) for w in words], dtype=np.int64)
    B = sorted(list(c["rv"]) + list(c["rp"]) + ([float(pcr)] if float(pcr) > 0 else []), reverse=True)
    return int(ok) == 1, np.float32(float(binh)), np.float32(float(binc)), f, bits, np.array(B)


def lookup(r, binh, binc, f, bits):
    """The force pass's interval choice: bin by fmaf(r, binh, binc), then the
    record's split radius picks its lower or upper interval offset."""
    r32 = np.float32(r)
    b = int(np.float32(np.float64(r32) * np.float64(binh) + np.float64(binc)))
    b = min(max(b, 0), NB - 1)
    split = f[4 * b]
    off = bits[4 * b + 2] if r32 >= split else bits[4 * b + 1]
    return int(off) // 8


def check_table(oracle, c, binh, binc, f, bits, B):
    nb = len(B)
    h = 1.0 / float(binh)
    lo = -float(binc) * h
    rng = np.random.default_rng(0)
    # bin records: radii spread over every bin (away from breakpoints by 1e-6)
    for b in range(NB):
        for r in lo + (b + rng.uniform(0.0, 1.0, 9)) * h:
            if np.min(np.abs(B - np.float32(r))) < 1e-6:
                continue
            want = int(np.sum(B > np.float32(r)))
            assert lookup(r, binh, binc, f, bits) == min(want, nb), (b, r)
    # interval polynomials against the oracle's knot sums (r above the join at 1.9 A)
    pot = oracle.Potential(c)
    iv = f[4 * NB:].astype(np.float64).reshape(NI, 8)
    edges = np.concatenate([[B[0] + 0.5], B, [1.9]])
    for m in range(nb + 1):
        hi_r, lo_r = edges[m], max(edges[m + 1], 1.9)
        if hi_r <= lo_r:
            continue
        r = np.linspace(lo_r, hi_r, 23)[1:-1]
        u, c0, c1, c2, c3, q0, q1, q2 = iv[m]
        x = u - r
        v = c0 + x * (c1 + x * (c2 + x * c3))
        dph = q0 + x * (q1 + x * q2)
        assert np.max(np.abs(v - pot.f("vee", r))) < 5e-6, m
        assert np.max(np.abs(dph - pot.f("dphi", r))) < 5e-6, m


def test_fe_table_matches_bruteforce_for_the_shipped_potential(probe, oracle):
    ok, binh, binc, f, bits, B = build(probe, potential.FE_POTENTIAL)
    assert ok
    check_table(oracle, potential.FE_POTENTIAL, binh, binc, f, bits, B)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fe_table_matches_bruteforce_for_perturbed_knots(probe, oracle, seed):
    # amplitudes +-5 %, knot radii +-0.2 % (keeps the breakpoints 0.04 A apart)
    rng = np.random.default_rng(seed)
    th = potential.theta_vector()
    scale = np.where(np.isin(np.arange(16), [6, 7, 8, 9, 10, 11, 14, 15]), 0.002, 0.05)
    th = th * (1 + scale * rng.uniform(-1.0, 1.0, 16))
    c = potential.coeffs_from_theta(th)
    ok, binh, binc, f, bits, B = build(probe, c)
    assert ok
    check_table(oracle, c, binh, binc, f, bits, B)


def test_fe_table_refuses_breakpoints_closer_than_the_finest_bin(probe):
    c = potential.coeffs_from_theta(potential.theta_vector())
    c["rv"][1] = c["rv"][0] - 0.003      # two pair knots 0.003 A apart
    ok = build(probe, c)[0]
    assert not ok


def test_fe_table_refuses_a_breakpoint_on_every_bin_edge(probe):
    c = potential.coeffs_from_theta(potential.theta_vector())
    c["rv"][2] = 3.0                     # a multiple of 2^-4 .. 2^-8: on an edge of every grid
    ok = build(probe, c)[0]
    assert not ok
