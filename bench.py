#!/usr/bin/env python3
"""Benchmark of the MI355X force engine on the EAM aluminium workload.

Workload: fcc Al, a = 4.05 A, 40^3 cells = 256,000 atoms, periodic,
300 K Maxwell-Boltzmann velocities plus 0.02 A Gaussian displacements
(seeded), Johnson-type EAM tabulated to 5,000-point splines with cutoff
5.5 A, FP32 forces with compensated sums, FP64 velocity Verlet, dt = 2 fs.
One step = one velocity-Verlet step of all atoms (kick, drift, list check,
density pass, force pass, kick).  Inputs are resident on the GPU before the
timed region.

Prints one JSON line: atom-steps/s, the force error of the final state
against the FP64 oracle, the dominant kernel's HBM roofline fraction
(measured with HIP events on the engine's stream) and the oracle's CPU rate
on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import objectmd_amd as om  # noqa: E402
from objectmd_amd import lattice, tables  # noqa: E402

HBM_PEAK_GBS = 8000.0
# compulsory per-atom bytes of one launch of the force pass: read the FP32
# snapshot (16 B) and F'(rho) (4 B), write the FP32 force (12 B)
FORCE_PASS_BYTES = 32
# compulsory per-atom-step bytes of the whole EAM step (integrator 128,
# density pass 20, force pass 32)
STEP_BYTES = 180


def build_system(cells, seed, disp, temperature, a=4.05):
    pos, (lo, hi) = lattice.fcc100(cells, cells, cells, a)
    rng = np.random.default_rng(seed)
    x = pos + rng.normal(0.0, disp, pos.shape)
    L = hi - lo
    for d in range(3):
        below = x[:, d] < lo[d]
        x[below, d] = x[below, d] + L[d]
        above = x[:, d] >= lo[d] + L[d]
        x[above, d] = x[above, d] + (-L[d])
    m = tables.mass_sim(tables.MASS_AL)
    v = lattice.maxwell_boltzmann(len(x), m, temperature, seed + 1)
    return x, v, lo, hi, m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_probe(mbytes):
    """One device-to-device hipMemcpy of mbytes MiB: the plain copy kernel
    whose FETCH_SIZE / WRITE_SIZE counters calibrate the traffic figures."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = int(mbytes) << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:   # DeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"[bench] copy probe: {nbytes} bytes device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--skin", type=float, default=0.6)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="steps of the FP64 oracle timed on the host (0: skip)")
    ap.add_argument("--no-force-check", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB "
                         "(counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus above 1 is not supported "
              "(independent replicas and domain decomposition are described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    dt = 0.002
    seed = 12346
    x, v, lo, hi, m = build_system(args.cells, seed, 0.02, 300.0)
    n = len(x)
    model = tables.JohnsonEAM()
    tab = model.tables(5000)
    log(f"[bench] EAM Al n={n} skin={args.skin} steps={args.steps} warmup={args.warmup}")

    eng = om.Engine()
    eng.set_box(lo, hi, 7)
    eng.upload_atoms(x, v, mass=[m])
    eng.add_table(om.TAB_PAIR, 0, 0, tab["PAIR"].y, tab["PAIR"].dr, tab["PAIR"].r0)
    eng.add_table(om.TAB_EDENS, 0, 0, tab["EDENS"].y, tab["EDENS"].dr, tab["EDENS"].r0)
    eng.add_table(om.TAB_EMBED, 0, 0, tab["EMBED"].y, tab["EMBED"].dr, tab["EMBED"].r0)
    eng.set_potential(om.POT_EAM)
    eng.set_neighbor(args.skin, args.check_every, 0)
    base, _ = eng.compute()            # FirstRun: BasePotential, then f = 0
    eng.zero_forces()
    ek0 = 0.5 * m * float((v * v).sum())
    e_start = ek0 / n                   # Ep - BasePotential = 0 at t = 0
    if args.warmup:
        eng.step(args.warmup, dt)
    log("[bench] warmup done")

    eng.timing_start()
    t0 = time.perf_counter()
    ek, ep = eng.step(args.steps, dt)   # returns after the stream drained
    t1 = time.perf_counter()
    times = eng.timing_stop()
    elapsed = t1 - t0
    rate = n * args.steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed:.4f} s: {rate:.4e} atom-steps/s")
    e_end = (ek + ep - base) / n
    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_eam256k.json")
    if os.path.exists(tpath) and args.cells == 40:
        with open(tpath) as fl:
            traffic = json.load(fl).get("traffic_bytes_per_launch")
    force_ms, force_launches = times["force"]
    avg_force_s = force_ms / max(force_launches, 1) / 1e3
    achieved = FORCE_PASS_BYTES * n / avg_force_s / 1e9 if avg_force_s > 0 else 0.0

    # force error of the final state against the FP64 oracle
    force_err = None
    if not args.no_force_check:
        from oracle import oracle as O
        xf = eng.download(om.X)
        eng.compute()
        fg = eng.download(om.F)
        sp = [O.Spline.from_table(tab[k]) for k in ("PAIR", "EDENS", "EMBED")]
        nb = O.neighbors(xf, lo, hi - lo, 7, 5.5 + args.skin, maxn=96)
        fo = O.eam_forces(xf, lo, hi - lo, 7, nb, *sp)[0]
        force_err = float(np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum()))
        log(f"[bench] force relRMS vs FP64 oracle = {force_err:.3e}")

    cpu = None
    if args.cpu_steps > 0:
        from oracle import oracle as O
        sp = [O.Spline.from_table(tab[k]) for k in ("PAIR", "EDENS", "EMBED")]
        c0 = time.perf_counter()
        O.run(1, x, v, m, lo, hi - lo, 7, sp, 5.5, 0.3, 10, args.cpu_steps, dt, maxn=96)
        c1 = time.perf_counter()
        cpu = {"value": n * args.cpu_steps / (c1 - c0), "unit": "atom-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"same {n}-atom EAM system, {args.cpu_steps} velocity-Verlet steps "
                         "of the FP64 C oracle (half list, list rebuilt every 10 steps), "
                         "one thread"}
        log(f"[bench] cpu oracle {cpu['value']:.4e} atom-steps/s")

    out = {
        "metric": "atom-steps/sec at stated force error vs FP64 oracle; % HBM roofline",
        "value": rate,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "EAM Al fcc a=4.05 A, 40^3 cells = 256,000 atoms, PBC, 300 K + "
                        "0.02 A displacements, Johnson-type EAM 5,000-point splines, "
                        "rc 5.5 A, dt 2 fs",
            "atoms": n,
            "force_dtype": "f32 with compensated per-atom sums",
            "integrator_dtype": "f64",
            "force_rel_rms_vs_fp64": force_err,
            "energy_per_atom_start_end_eV": [e_start, e_end],
            "skin_A": args.skin,
            "step_roofline_frac": STEP_BYTES * rate / (HBM_PEAK_GBS * 1e9),
            "kernel_ms": {k: {"total": t, "launches": c} for k, (t, c) in times.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_eam_force",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
