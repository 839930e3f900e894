#!/usr/bin/env python
"""bench.py -- atom-steps/s of MOLDY's Fe force hot path on one MI355X.

Workload (the headline configuration): bcc Fe, 64^3 conventional cells =
524,288 atoms, a = 2.8665 A, seeded Gaussian displacements sigma = 0.05 A,
Maxwell-Boltzmann velocities at 300 K, Parrinello-Rahman NPT at 0 GPa
(box mass 50 amu/atom), dt = 1 fs, Finnis-Sinclair Fe potential.  A "step" is
one MD step of the reference's moldin loop (velocity Verlet with force:
neighbour maintenance, density + embedding, force + virial, PR box update,
thermodynamic sums).  Inputs are resident in HBM before timing starts.

value = atom-steps/s of the FP32-pair/FP64-accumulate form; the plain FP64
form is run on the same input and reported beside it.  The CPU baseline is
the reference Fortran (built from its own sources into oracle/_ref) timed on
this host on a bounded sample of the same workload.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(seed=2026):
    from moldy_amd import lattice
    return lattice.make_system(nc=64, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=seed)


def run_form(sd, precision, steps, warmup, breakdown_steps=50):
    """Times `steps` MD steps with HIP events around every 10th launch of the
    force kernel (an event pair idles the GPU for a few microseconds, so only a
    sample of launches is bracketed); then a separate pass of `breakdown_steps`
    with events around every kernel gives the per-kernel breakdown."""
    from moldy_amd import engine
    eng = engine.Engine(sd, ivol=0, press_gpa=0.0, boxmass=50.0, ntcm=10, precision=precision,
                        rebuild=1, rpad=0.3)
    eng.linkup()
    eng.md_init()
    if warmup:
        eng.md_steps(warmup, trace=False)
    eng.synchronize()
    reb0, _ = eng.stats()
    eng.timing(True, classes=("force",), every=10)
    t0 = time.perf_counter()
    eng.md_steps(steps, trace=False)
    eng.synchronize()
    t1 = time.perf_counter()
    kforce = eng.timing_read()["force"]
    eng.timing(False)
    reb1, full_pairs = eng.stats()
    eng.timing(True, every=1)
    eng.md_steps(breakdown_steps, trace=False)
    eng.synchronize()
    kt = eng.timing_read()
    eng.timing(False)
    res = {
        "seconds": t1 - t0,
        "ms_per_step": (t1 - t0) / steps * 1e3,
        "atom_steps_per_s": sd.n * steps / (t1 - t0),
        "force_kernel": {"ms_total": kforce[0], "launches": kforce[1],
                         "us_avg": (1e3 * kforce[0] / kforce[1]) if kforce[1] else None},
        "kernels": {k: {"ms_total": v[0], "launches": v[1],
                        "us_avg": (1e3 * v[0] / v[1]) if v[1] else None} for k, v in kt.items()},
        "rebuilds_in_timed_region": reb1 - reb0,
        "full_pairs_per_atom": full_pairs / sd.n,
        "pe": eng.scalar("pe"),
    }
    return eng, res


def batch_legs(reps):
    """Batched evaluation of 4096 uneven periodic cells (one persistent
    hipGraph-replayed launch, autotuned) and the fused loss + adjoint gradient
    over the same batch (targets from knots perturbed by a seeded +-5 %)."""
    from moldy_amd import engine, lattice, potential
    cells = lattice.make_batch(ncell=4096, seed=7)
    be = engine.BatchEngine(cells, precision=engine.PREC_FP32)
    be.evaluate()
    be.synchronize()
    block, wg, _ = be.tune(reps=3)
    ms_eval = be.time(reps)
    info = be.info()
    rng = np.random.default_rng(11)
    th = potential.theta_vector() * (1 + rng.uniform(-0.05, 0.05, 16))
    be.set_potential(potential.coeffs_from_theta(th))
    be.evaluate()
    e_t, _, f_t, _ = be.results()
    be.set_potential(potential.FE_POTENTIAL)
    be.set_targets(e_t, f_t)
    loss, grad = be.loss_grad()
    ms_grad = be.time_loss_grad(reps)
    be.close()
    na, nf = info["natom"], info["pairs"] / info["natom"]
    # per-atom stream models: evaluation = positions 24 + forces 24 + F' 8 +
    # the list written once (with the densities) and read by the force phase
    # (8 n_f), plus energy and virial per cell; loss + gradient = positions,
    # force targets, forces, force seeds written and read (120) + the list
    # written once and read by the force and adjoint phases (12 n_f)
    b_eval = na * (56.0 + 8.0 * nf) + 56.0 * len(cells)
    b_grad = na * (120.0 + 12.0 * nf) + 56.0 * len(cells)

    def roof(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"bytes_per_launch": nbytes, "ms_per_launch": ms, "achieved_GBs": gbs,
                "frac": gbs / PEAK_HBM_GBS,
                "bound": "VALU issue and latency (far from the HBM roof; per-type VALU counts "
                         "are in profiles/)"}
    return {
        "workload": "4096 periodic Fe cells, bcc/fcc n^3 supercells n in 3..10, strain +-3%, "
                    "0-2% vacancies, sigma 0.1 A (seed 7)",
        "cells": len(cells), "atoms": info["natom"],
        "full_list_neighbours_per_atom": info["pairs"] / info["natom"],
        "tuned_block": block, "tuned_workgroups_per_cu": wg,
        "ms_per_evaluation": ms_eval, "atoms_per_s": info["natom"] / ms_eval * 1e3,
        "cells_per_s": len(cells) / ms_eval * 1e3,
        "roofline": roof(b_eval, ms_eval),
        "adjoint": {"ms_per_loss_grad": ms_grad, "grads_per_s": 1e3 / ms_grad, "loss": loss,
                    "n_parameters": len(grad), "roofline": roof(b_grad, ms_grad)},
        "dtype": "f32 pair/spline arithmetic, f64 accumulation",
    }


def copy_calibration():
    """Plain device copies of known size (1 GiB read + 1 GiB written per
    dispatch, 3 dispatches each of an elementwise kernel and of the runtime's
    copy kernel), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale
    those counters to bytes."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_dispatch": 2 * 4 * (1 << 28)}))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_f32_bcc524288_traffic.md")


def measured_traffic():
    """HBM bytes per force launch from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (scaled by the copy calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for ln in fh:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ref_step_seconds(refrun, sd, threads, n1, n2, repeats, **kw):
    """Per-step wall seconds of the reference's MD loop: (loop(n2) - loop(n1))
    / (n2 - n1) for runs of n1 and n2 steps, median of `repeats` pairs."""
    per = []
    for _ in range(repeats):
        t1 = refrun.time_md_loop(sd, n1, threads, timeout=900, **kw)
        t2 = refrun.time_md_loop(sd, n2, threads, timeout=900, **kw)
        per.append((t2 - t1) / (n2 - n1))
    return float(np.median(per)), per


def cpu_baseline(sd, threads, with_c2=True):
    """The reference Fortran (its modules built with amdflang -O2 -fopenmp) on
    the host cores, OMP_PROC_BIND=close and OMP_PLACES=cores.  Per-step wall
    time by step differencing of the MD loop: runs of 1 and 11 steps, so the
    difference holds ten steps and exactly one list build (nbrupdate 10), as
    the reference's cadence; median of 5 pairs at `threads` threads and of 3
    pairs at 1 thread.  The 3.46M-atom liquid sample: runs of 1 and 3 steps
    (no list build in the difference), median of 3 pairs.  Falls back to the
    C port when the reference driver is not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    host = {"model": _cpu_model(), "host_cpus": len(os.sched_getaffinity(0)), "nproc": os.cpu_count(),
            "omp": "OMP_PROC_BIND=close OMP_PLACES=cores"}
    if os.path.exists(refrun.DRIVER_OMP):
        kw = dict(ivol=0, ntcm=10, press=0.0, boxmass=50.0)
        sec, per = _ref_step_seconds(refrun, sd, threads, 1, 11, 5, **kw)
        sec1, per1 = _ref_step_seconds(refrun, sd, 1, 1, 11, 3, **kw)
        out = {"value": sd.n / sec, "unit": "atom-steps/s", "cores": threads, "kind": "reference",
               "sample": f"{sd.n} atoms (the same NPT workload), per-step wall time of the MD loop from "
                         f"runs of 1 and 11 steps (ten steps incl. one nbrupdate=10 list build), "
                         f"median of 5 pairs",
               "per_step_s": per, **host,
               "single_thread": {"value": sd.n / sec1, "cores": 1, "per_step_s": per1,
                                 "sample": "same, median of 3 pairs"}}
        if with_c2:
            liq = c2_system()
            sec2, per2 = _ref_step_seconds(refrun, liq, threads, 1, 3, 3, ivol=1, ntcm=1000, press=0.0,
                                           boxmass=50.0)
            out["hot_liquid_3456000"] = {"value": liq.n / sec2, "cores": threads, "per_step_s": per2,
                                         "sample": f"{liq.n} atoms, runs of 1 and 3 steps (no list build "
                                                   f"in the difference), median of 3 pairs"}
        return out
    import pyoracle as po
    pot = po.Potential()
    po.lib().orc_set_threads(threads)
    S = po.OracleSystem(sd, pot, ivol=0, ntcm=10)
    S.call("linkup")
    S.call("md_init")
    steps = 10
    t0 = time.perf_counter()
    for _ in range(steps):
        S.call("md_step")
    secs = time.perf_counter() - t0
    return {"value": sd.n * steps / secs, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": f"{sd.n} atoms x {steps} MD steps, C restatement with OpenMP", **host}


def c0_leg(steps=1000):
    """NVE energy conservation on the 2000-atom cell (10^3 bcc cells, 300 K,
    dt = 1 fs, reference rebuild cadence): total energy per atom over the run."""
    from moldy_amd import engine, lattice
    sd = lattice.make_system(nc=10, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=1234)
    out = {"workload": f"bcc Fe 10^3 cells = {sd.n} atoms, 300 K, NVE, dt=1 fs, {steps} steps, "
                       "nbrupdate 10 (reference cadence)"}
    for prec, tag in ((engine.PREC_FP64, "fp64"), (engine.PREC_FP32, "fp32")):
        e = engine.Engine(sd, ivol=1, precision=prec, ntcm=10, rebuild=0)
        e.linkup()
        e.md_init()
        e.synchronize()
        t0 = time.perf_counter()
        tr = e.md_steps(steps, trace=True)
        e.synchronize()
        secs = time.perf_counter() - t0
        full = e.stats()[1]
        e.close()
        te = tr[:, 2] / sd.n
        t_ps = np.arange(steps) * 1e-3
        slope = float(np.polyfit(t_ps, te, 1)[0])
        out[tag] = {"te_first_eV_per_atom": float(te[0]), "te_last_eV_per_atom": float(te[-1]),
                    "max_abs_dev_eV_per_atom": float(np.max(np.abs(te - te[0]))),
                    "drift_eV_per_atom_per_ps": slope,
                    "ms_per_step": secs / steps * 1e3, "atom_steps_per_s": sd.n * steps / secs,
                    "roofline": step_roofline(sd.n, full / sd.n, secs / steps),
                    "bound": "launch latency (2000 atoms: each kernel is a few workgroups)"}
    return out


def step_roofline(n, nf, sec_per_step):
    """Whole MD step against the HBM roof: 304 + 8 n_f algorithmic bytes per
    atom-step (positions, velocities, accelerations and forces in FP64, the
    packed records, the density pass's and the force pass's list reads)."""
    b = n * (304.0 + 8.0 * nf)
    gbs = b / sec_per_step / 1e9
    return {"model": "whole MD step, 304 + 8 n_f bytes per atom-step", "bytes_per_step": b,
            "achieved_GBs": gbs, "frac": gbs / PEAK_HBM_GBS}


def c2_system():
    """120^3 bcc cells (3,456,000 atoms) at 7.0 g/cm^3, 2500 K, dt = 0.5 fs."""
    from moldy_amd import lattice
    a = lattice.liquid_lattice_parameter(7.0)
    return lattice.make_system(nc=120, a=a, sigma=0.05, temp=2500.0, deltat_fs=0.5, seed=2500)


def c2_leg(steps=1000):
    """Hot liquid: 120^3 bcc cells (3,456,000 atoms) at 7.0 g/cm^3, 2500 K,
    NVE, dt = 0.5 fs, skin 0.3 A, rebuilds on the displacement criterion only;
    rebuild count and neighbour-build time reported apart from the step rate."""
    from moldy_amd import engine
    sd = c2_system()
    a = sd.b0[0, 0] / 120
    out = {"workload": f"bcc Fe 120^3 cells = {sd.n} atoms, a={a:.5f} A (7.0 g/cm^3), 2500 K, NVE, "
                       f"dt=0.5 fs, skin 0.3 A, {steps} steps, rebuild when max displacement > skin/2"}
    for prec, tag in ((engine.PREC_FP32, "fp32"), (engine.PREC_FP64, "fp64")):
        e = engine.Engine(sd, ivol=1, precision=prec, rebuild=1, skin=0.3)
        e.linkup()
        e.md_init()
        e.md_steps(10, trace=False)
        e.synchronize()
        e.timing(True, classes=("rebuild", "force"), every=1)
        r0 = e.stats()[0]
        t0 = time.perf_counter()
        tr = e.md_steps(steps, trace=True)
        e.synchronize()
        secs = time.perf_counter() - t0
        reb, full = e.stats()
        reb -= r0
        kt = e.timing_read()
        e.timing(False)
        e.close()
        rb = kt.get("rebuild", (0.0, 0))
        kf = kt.get("force", (0.0, 0))
        us_f = 1e3 * kf[0] / max(kf[1], 1)
        nb = sd.n * (56.0 + 4.0 * full / sd.n)  # the same per-atom stream model as the main roofline
        out[tag] = {"ms_per_step": secs / steps * 1e3, "atom_steps_per_s": sd.n * steps / secs,
                    "rebuilds": reb, "rebuild_us_avg": 1e3 * rb[0] / max(rb[1], 1),
                    "rebuild_ms_total": rb[0],
                    "te_drift_eV_per_atom": float((tr[-1, 2] - tr[0, 2]) / sd.n),
                    "step_roofline": step_roofline(sd.n, full / sd.n, secs / steps),
                    "roofline": {"kernel": "force pass", "us_per_launch": us_f, "bytes_per_launch": nb,
                                 "achieved_GBs": nb / (us_f * 1e-6) / 1e9,
                                 "frac": nb / (us_f * 1e-6) / 1e9 / PEAK_HBM_GBS,
                                 "full_list_neighbours_per_atom": full / sd.n},
                    "timing_note": "HIP events around every force launch and every rebuild in this leg"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0))))
    ap.add_argument("--no-cpu-c2", action="store_true", help="skip the 3.46M-atom CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the NVE-drift (2000 atoms) and hot-liquid (3.46M atoms) legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the FETCH_SIZE/WRITE_SIZE calibration copies")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the force path of one periodic cell "
              "on one MI355X); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    from moldy_amd import engine

    sd = workload()
    log(f"workload n={sd.n} steps={args.steps} warmup={args.warmup}")
    eng32, r32 = run_form(sd, engine.PREC_FP32, args.steps, args.warmup)
    log(f"fp32 form: {r32['ms_per_step']:.4f} ms/step, {r32['atom_steps_per_s']:.4e} atom-steps/s")
    f32 = eng32.forces_cart()
    eng32.close()
    r64 = None
    force_err = None
    if not args.no_fp64:
        eng64, r64 = run_form(sd, engine.PREC_FP64, args.steps, args.warmup)
        log(f"fp64 form: {r64['ms_per_step']:.4f} ms/step, {r64['atom_steps_per_s']:.4e} atom-steps/s")
        eng64.close()
        # FP32-form force error on this configuration: both forms restarted
        # from the same input after md_init (FP64 form is oracle-verified).
        e32 = engine.Engine(sd, ivol=0, precision=engine.PREC_FP32)
        e32.linkup()
        e32.md_init()
        e64 = engine.Engine(sd, ivol=0, precision=engine.PREC_FP64)
        e64.linkup()
        e64.md_init()
        force_err = float(np.max(np.abs(e32.forces_cart() - e64.forces_cart())))
        pe_err_atom = abs(e32.scalar("pe") - e64.scalar("pe")) / sd.n
        e32.close()
        e64.close()
        log(f"fp32 vs fp64 max force error {force_err:.3e} eV/A, pe error {pe_err_atom:.3e} eV/atom")
    del f32

    # roofline of the dominant kernel (force pass): algorithmic bytes per atom
    # = positions (24) + F' (8) + neighbour indices (4 n_f) + forces (24)
    nf = r32["full_pairs_per_atom"]
    kf = r32["force_kernel"]
    bytes_per_launch = sd.n * (56.0 + 4.0 * nf)
    achieved = bytes_per_launch / (kf["us_avg"] * 1e-6) / 1e9 if kf["us_avg"] else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": (achieved / PEAK_HBM_GBS) if achieved else None, "traffic": measured_traffic(),
                "kernel": "k_force_f32 (FP32 pair terms from the interval table, 16-byte FP32 partner "
                          "records, FP64 sums across lanes)",
                "us_per_launch": kf["us_avg"], "bytes_per_launch": bytes_per_launch}

    batch = None
    if not args.no_batch:
        log("batch: 4096 cells ...")
        batch = batch_legs(reps=10)
        log(f"batch: {batch['ms_per_evaluation']:.3f} ms/evaluation, "
            f"{batch['adjoint']['grads_per_s']:.1f} grads/s")

    c0 = c2 = None
    if not args.no_workloads:
        log("nve drift, 2000 atoms ...")
        c0 = c0_leg()
        log(f"nve drift: fp64 {c0['fp64']['drift_eV_per_atom_per_ps']:.3e}, "
            f"fp32 {c0['fp32']['drift_eV_per_atom_per_ps']:.3e} eV/atom/ps")
        log("hot liquid, 3.46M atoms ...")
        c2 = c2_leg()
        log(f"hot liquid fp32: {c2['fp32']['ms_per_step']:.3f} ms/step, {c2['fp32']['rebuilds']} rebuilds, "
            f"{c2['fp32']['rebuild_us_avg']:.0f} us each")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(sd, args.cpu_threads, with_c2=not args.no_cpu_c2)
        log(f"cpu baseline {cpu['value']:.4e} atom-steps/s on {cpu['cores']} cores ({cpu['kind']}, "
            f"{cpu['model']})")
        if "single_thread" in cpu:
            log(f"cpu baseline, 1 thread: {cpu['single_thread']['value']:.4e} atom-steps/s")

    line = {
        "metric": "atom-steps/s, 0.5M-atom Fe EAM (FP32/FP64-acc vs FP64); batched grads/s",
        "value": r32["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair/spline arithmetic, f64 accumulation and integration",
        "data": "synthetic (seeded bcc Fe lattice, Gaussian displacements, MB velocities)",
        "config": {"workload": "bcc Fe 64x64x64 cells = 524288 atoms, a=2.8665 A, sigma=0.05 A, "
                               "300 K, Parrinello-Rahman NPT 0 GPa, dt=1 fs, Finnis-Sinclair Fe",
                   "n_atoms": sd.n, "neighbour_skin_A": 0.3,
                   "rebuilds_in_timed_region": r32["rebuilds_in_timed_region"],
                   "full_list_neighbours_per_atom": nf},
        "fp64_plain": None if r64 is None else {"value": r64["atom_steps_per_s"],
                                                "ms_per_step": r64["ms_per_step"],
                                                "force_kernel_us": r64["force_kernel"]["us_avg"],
                                                "kernels": r64["kernels"]},
        "fp32_force_error_vs_fp64_eV_per_A": force_err,
        "kernels": r32["kernels"],
        "kernels_note": "per-kernel averages from a separate 50-step pass with events around every "
                        "kernel class; in the timed region only every 10th force-kernel launch records "
                        "its own dispatch events",
        "roofline": roofline,
        "step_roofline": step_roofline(sd.n, nf, r32["ms_per_step"] * 1e-3),
        "nve_drift_2000": c0,
        "hot_liquid_3456000": c2,
        "batch": batch,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (r32["atom_steps_per_s"] / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
