#!/usr/bin/env python3
"""Headline benchmark: device-resident MD of 32,000 three-site waters
(96,000 atoms, cubic box 98.6 A) with Lennard-Jones (C2 switch 8-10 A) plus
Ewald real-space Coulomb (alpha 0.32 A^-1, Rc 10 A) and the harmonic water
model, leapfrog at dt = 1 fs, pair list rebuilt every 10 steps (and sooner if
an atom moves more than half the list buffer).  The start is a lattice with
seeded orientations, relieved of H-H contacts under 2 A and equilibrated for
a few hundred untimed steps with velocity rescaling to 300 K.

Prints one JSON line: pair interactions/s (whole job), ns/day, the pair
kernel's roofline figures (HIP events around every pair-kernel launch on
the kernel's own stream) and the CPU baseline (the FP64 oracle on this
host's cores, a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from protomol_amd import _lib as L  # noqa: E402
from protomol_amd import systems as S  # noqa: E402

N_MOL, BOX, DT_FS, REBUILD = 32000, 98.6, 1.0, 10
ALPHA, ACCURACY = 0.32, math.exp(-10.24)
# algorithmic work per within-cutoff pair of the LJ(C2) + Ewald-real kernel:
# distance 11, LJ 11, C2 switch 13, Ewald real 23, force combine/accumulate 10,
# energies 2, virial 12 (add/sub/mul = 1, FMA = 2, each rsqrt/rcp/exp = 1)
FLOP_PER_PAIR = 82
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def force_desc(skin, timing=True):
    d = L.ForceDesc()
    d.term_a = L.TERM_LJ
    d.switch_a = L.switch(L.SWITCH_C2, 8.0, 10.0)
    d.term_b = L.TERM_EWALD_REAL
    d.cutoff = 10.0
    d.ewald_alpha = ALPHA
    d.ewald_accuracy = ACCURACY
    d.ewald_correction = 1
    d.skin = skin
    d.rebuild_interval = REBUILD
    d.flags = L.FLAG_BONDED | (L.FLAG_TIMING if timing else 0)
    return d


def cpu_baseline(system, seconds):
    """The FP64 oracle (the restatement of the reference's algorithm) timed on
    this host: force evaluations of the 96k-atom system until `seconds` pass."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
    except Exception as e:  # the checker library is not built
        return {"value": None, "unit": "pair interactions/s", "cores": 0, "kind": "port",
                "sample": f"unavailable: {e}"}
    cores = max(1, min(os.cpu_count() or 1, 16))
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    n_eval, t_tot, pairs = 0, 0.0, 0
    while t_tot < seconds or n_eval == 0:
        t0 = time.perf_counter()
        _, _, cl = O.nonbonded_cutoff(system, O.TERM_LJ, sw, cellsize=5.0, nthreads=cores)
        _, _, ce, _ = O.ewald_real(system, ALPHA, ACCURACY, cellsize=5.0, nthreads=cores)
        O.bonded(system)
        t_tot += time.perf_counter() - t0
        n_eval += 1
        pairs = max(cl[0], ce[0])
        log(f"cpu baseline: evaluation {n_eval} done, {t_tot:.1f} s")
    per = t_tot / n_eval
    return {"value": pairs / per, "unit": "pair interactions/s", "cores": cores, "kind": "port",
            "sample": f"{n_eval} force evaluation(s) of the 96,000-atom system (LJ C2 + Ewald real + "
                      f"correction + bonded, cell list rebuilt per evaluation as the reference does), "
                      f"{per:.3f} s each on {cores} threads",
            "ns_per_day": (1.0 / per) * DT_FS * 0.0864}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--skin", type=float, default=1.2)
    ap.add_argument("--equil", type=int, default=300, help="untimed equilibration steps (rescaled to 300 K)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent systems shard by placement, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    log("building the 96,000-atom water box")
    system = S.water_box(N_MOL, BOX, seed=2024, min_contact=2.0)
    v0 = S.maxwell_velocities(system, 300.0, seed=2025)
    name, ncu, hbm = L.device_info()
    log(f"device: {name}, {ncu} CUs, {hbm:.0f} GiB")
    with L.Context(force_desc(args.skin), [system]) as ctx:
        ctx.md_set_state(system.x, v0)
        seg = 25
        for k in range(0, args.equil, seg):
            ctx.md_run(seg, DT_FS, energies=False)
            x, v, _, _ = ctx.md_get_state()
            t_now = S.temperature(system, v)
            ctx.md_set_state(x, v * math.sqrt(300.0 / max(t_now, 1e-6)))
        if args.equil:
            log(f"equilibrated {args.equil} steps (last T before rescale {t_now:.0f} K)")
        log(f"warmup: {args.warmup} steps")
        ctx.md_run(args.warmup, DT_FS, energies=False)
        log(f"timed: {args.steps} steps")
        t0 = time.perf_counter()
        ctx.md_run(args.steps, DT_FS, energies=False)
        t1 = time.perf_counter()
        st = ctx.stats()
        x, v, f, e = ctx.md_get_state()
        log(f"timed region done: {1e3 * (t1 - t0) / args.steps:.3f} ms/step; counting pairs")
        cnt, _ = ctx.census(0, x)
    ms_step = 1e3 * (t1 - t0) / args.steps
    steps_s = 1e3 / ms_step
    pairs = int(cnt)
    pair_ms = st["pair_ms"] / max(1.0, st["pair_launches"])
    achieved = pairs * FLOP_PER_PAIR / (pair_ms * 1e-3) / 1e12
    out = {
        "metric": "pair interactions/s + ns/day @96k atoms; batched systems/s + p99 latency",
        "value": pairs * steps_s,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 32,000-water lattice box, Maxwell 300 K velocities, equilibrated untimed)",
        "config": {"workload": "water 32,000 mol / 96,000 atoms, box 98.6 A, LJ C2 8-10 A + Ewald real "
                               "alpha 0.32 Rc 10 A + bond/angle, leapfrog dt 1 fs, pair list every 10 steps",
                   "atoms": system.n, "pairs_per_step": pairs, "skin_A": args.skin, "rebuild_interval": REBUILD,
                   "rebuilds": st["rebuilds"], "max_displacement_A": st["max_displacement"],
                   "equilibration_steps": args.equil},
        "ns_per_day": steps_s * DT_FS * 0.0864,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_pair<LJ, EWALD_REAL>", "kernel_ms": pair_ms,
                     "flop_per_pair": FLOP_PER_PAIR},
        "device": name,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(system, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
