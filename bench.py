#!/usr/bin/env python3
"""Benchmarks of the nonbonded force path on one MI355X.

Headline (the JSON line's ``value``): device-resident MD of 32,000 three-site
waters (96,000 atoms, cubic box 98.6 A) with Lennard-Jones (C2 switch
8-10 A) plus Ewald real-space Coulomb (alpha 0.32 A^-1, Rc 10 A) and the
harmonic water model, leapfrog at dt = 1 fs, pair list rebuilt every 10 steps
(and sooner if an atom moves more than half the list buffer), forces only on
the timed steps (energies at the last step; a variant sums them every 10th
step).  The start is a lattice with seeded orientations, relieved of H-H
contacts under 2 A and equilibrated for a few hundred untimed steps with
velocity rescaling to 300 K.

Beside it, in the same line:
  force_err     RMS and largest per-atom relative force error of the final
                state against the FP64 oracle (every form has one)
  roofline      the pair kernel's FP32 rate (HIP events around every
                pair-kernel launch on the kernel's own stream) against peak
  cpu_baseline  the FP64 oracle (the reference algorithm) on 1 thread and on
                the host cores of this GPU's share, and the reference's own
                classes (compiled from its sources) on 1 thread
  water1k       3,000 atoms: one evaluation and 100 steps at 0.5 fs
  batched       64 water boxes of 500-16,000 molecules (log-uniform), all in
                one context: graph-replayed MD with autotuned launch shapes,
                against the same systems run one per context (plain kernel)
  served        a Poisson stream of requests (force evaluation or 200-step
                MD segment, 50/50) on water boxes of 1,000-32,000 molecules:
                continuous batching with admission by HBM footprint, against
                the same requests served one at a time; p50/p99 latency
  gb            32 GB/OBC bead chains of 5,000 atoms, all pairs, batched
  ewald         the headline box with full Ewald (real + reciprocal +
                correction, ProtoMol's CoulombEwald): MD step time, the
                reciprocal-space passes' FP64 rate against peak, and the FP64
                oracle's reciprocal term on this host's cores
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
from protomol_amd import autotune  # noqa: E402
from protomol_amd import serve  # noqa: E402
from protomol_amd import systems as S  # noqa: E402
from protomol_amd import workloads as W  # noqa: E402

DT_FS, REBUILD = 1.0, 10
ALPHA, ACCURACY = W.EWALD_ALPHA, W.EWALD_ACCURACY
# algorithmic work per within-cutoff pair of the LJ(C2) + Ewald-real kernel:
# distance 11, LJ 11, C2 switch 13, Ewald real 23, force combine/accumulate 10,
# energies 2, virial 12 (add/sub/mul = 1, FMA = 2, each rsqrt/rcp/exp = 1)
FLOP_PER_PAIR = 82
# LJ + Coulomb, C2 on both: distance 8, 1/r^2 and 1/r 2, LJ 11, Coulomb 3,
# two C2 switches 26, combine/accumulate 10, energies 2, virial 12
FLOP_PER_PAIR_C2C2 = 74
# GB per unordered pair per step: Born radii pass ~65, partial sums ~26,
# GB + LJ + Coulomb force pass ~150
GB_FLOP_PER_PAIR = 240
PEAK_FP32_TFLOPS = 157.3
# reciprocal-space work per (k-vector, atom) pair, counted on
# NonbondedFullEwaldSystemForce::reciprocalTerm: structure factor 8 (sinA,
# cosA from the (h, k) products and z phases 6, sums 2), force 10 (the
# sin/cos cross term 3, times a 1, times k 3, accumulate 3)
EW_FLOP_PER_KATOM = 18
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------ CPU side

def box_threads():
    """Host cores of this GPU's share: the process's CPU affinity, capped by
    OMP_NUM_THREADS when the environment sets it."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "host_share_threads": box_threads()}


def load_oracle():
    """The FP64 oracle (test-side checker and CPU baseline), or None."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
        O.lib()
        return O
    except Exception as e:  # the checker library is not built
        log(f"oracle unavailable: {e}")
        return None


def ref_seconds(records):
    """Wall time the reference's own classes (oracle/_ref/ref_driver, built
    from the reference sources) report for the evaluations in `records`, or
    None when that build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import refio
        if not os.path.exists(refio.DRIVER):
            return None
        return float(refio.run(records)["seconds"][0])
    except Exception as e:
        log(f"reference timing unavailable: {e}")
        return None


def timed(fn, seconds, at_least=1):
    """Run fn until `seconds` of it have passed (at least `at_least` times);
    mean seconds per call."""
    n, tot = 0, 0.0
    while tot < seconds or n < at_least:
        t0 = time.perf_counter()
        fn()
        tot += time.perf_counter() - t0
        n += 1
    return tot / n


def force_error(f, fo):
    d = np.linalg.norm(f - fo, axis=1)
    return {"rms_rel": float(np.sqrt(np.sum((f - fo) ** 2) / np.sum(fo ** 2))),
            "max_rel_atom": float(np.max(d / np.maximum(np.linalg.norm(fo, axis=1), 1.0))),
            "against": "FP64 oracle at the same positions", "max_rel_floor": "1 kcal/mol/A"}


def oracle_water32k(O, w, x=None, nthreads=1):
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    fl, _, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, cellsize=5.0, x=x, nthreads=nthreads, census=False)
    fe, _, _, _ = O.ewald_real(w, ALPHA, ACCURACY, cellsize=5.0, x=x, nthreads=nthreads, census=False)
    fb, _ = O.bonded(w, x=x)
    return fl + fe + fb


def oracle_water1k(O, w, x=None, nthreads=1):
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    f, _, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw, cellsize=5.0, x=x, nthreads=nthreads,
                                 census=False)
    fb, _ = O.bonded(w, x=x)
    return f + fb


def ref_records_water32k(w):
    """Record sets for the reference's LJ C2 pair force, its Ewald real +
    correction force and the bonded terms on system w (census off)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refio
    base = refio.system_records(w)
    lj = dict(base, mode=np.int32(1), term_a=np.int32(1), term_b=np.int32(0), sw_a=np.array((2, 8.0, 10.0, 0, 0), float),
              alg_cutoff=10.0, cellsize=5.0, census=np.int32(0))
    ew = dict(base, mode=np.int32(2), ew_alpha=ALPHA, ew_accuracy=ACCURACY, ew_parts=np.int32(5), cellsize=5.0,
              census=np.int32(0))
    bd = dict(base, mode=np.int32(4), bonds=w.bonds.reshape(-1), bond_k=w.bond_k, bond_r0=w.bond_r0,
              angles=w.angles.reshape(-1), angle_k=w.angle_k, angle_t0=w.angle_t0, cellsize=5.0)
    return [lj, ew, bd]


def reference_eval_seconds(record_sets):
    total = 0.0
    for r in record_sets:
        s = ref_seconds(r)
        if s is None:
            return None
        total += s
    return total


def cpu_baseline(system, pairs, seconds):
    """The headline system's force evaluation (LJ C2 + Ewald real +
    correction + bonded, cell list rebuilt per evaluation as the reference
    does) on the host: the oracle on 1 thread and on this GPU's host share,
    and the reference's own classes on 1 thread."""
    O = load_oracle()
    info = cpu_info()
    if O is None:
        return {"value": None, "unit": "pair interactions/s", "cores": 0, "kind": "port", "sample": "unavailable"}
    nt = info["host_share_threads"]
    t1 = timed(lambda: oracle_water32k(O, system, nthreads=1), seconds * 0.3)
    log(f"cpu baseline: oracle 1 thread {t1:.3f} s per evaluation")
    ta = timed(lambda: oracle_water32k(O, system, nthreads=nt), seconds * 0.3, at_least=2)
    log(f"cpu baseline: oracle {nt} threads {ta:.3f} s per evaluation")
    tr = reference_eval_seconds(ref_records_water32k(system))
    log(f"cpu baseline: reference 1 thread {tr} s per evaluation")
    rate = lambda t: None if t is None else pairs / t
    nsd = lambda t: None if t is None else (1.0 / t) * DT_FS * 0.0864
    out = {"value": rate(ta), "unit": "pair interactions/s", "cores": nt, "kind": "port",
           "sample": f"force evaluations of the 96,000-atom system ({pairs} pairs within 10 A; LJ C2 + Ewald real "
                     f"+ correction + bond/angle, cell list rebuilt per evaluation): {ta:.3f} s each on {nt} threads, "
                     f"{t1:.3f} s on 1 thread",
           "ns_per_day": nsd(ta),
           "threads_1": {"value": rate(t1), "ns_per_day": nsd(t1), "s_per_eval": t1, "kind": "port"},
           "threads_all": {"value": rate(ta), "ns_per_day": nsd(ta), "s_per_eval": ta, "threads": nt, "kind": "port"},
           "reference_1thread": {"value": rate(tr), "ns_per_day": nsd(tr), "s_per_eval": tr, "kind": "reference",
                                 "what": "the reference's NonbondedCutoffSystemForce<OneAtomPair<PBC,C2,LJ>> + "
                                         "NonbondedFullEwaldSystemForce<real, correction> + Bond/Angle "
                                         "SystemForces, compiled from its sources (single-threaded, as it runs)"}}
    out.update(info)
    return out


# ------------------------------------------------------------ GPU forms

def headline(args):
    log("building the 96,000-atom water box")
    system = W.water32k()
    v0 = S.maxwell_velocities(system, 300.0, seed=2025)
    shape = {"cell_scale": args.cell_scale, "pair_waves": args.pair_waves, "pair_split": args.pair_split}

    def desc(flags):
        return W.desc_water32k(args.skin, REBUILD, flags, **shape)

    # the production path: MD steps replayed from captured 10-step graphs
    with L.Context(desc(0), [system]) as ctx:
        ctx.md_set_state(system.x, v0)
        seg, t_now = 25, 300.0
        for k in range(0, args.equil, seg):
            ctx.md_run(seg, DT_FS, energies=False)
            x, v, _, _ = ctx.md_get_state()
            t_now = S.temperature(system, v)
            ctx.md_set_state(x, v * math.sqrt(300.0 / max(t_now, 1e-6)))
        if args.equil:
            log(f"equilibrated {args.equil} steps (last T before rescale {t_now:.0f} K)")
        log(f"warmup: {args.warmup} steps")
        ctx.md_run(args.warmup, DT_FS, energies=False)
        x_start, v_start, _, _ = ctx.md_get_state()
        log(f"timed: {args.steps} steps")
        t0 = time.perf_counter()
        ctx.md_run(args.steps, DT_FS, energies=False)
        t1 = time.perf_counter()
        st = ctx.stats()
        x, v, f, e = ctx.md_get_state()
        log(f"timed region done: {1e3 * (t1 - t0) / args.steps:.3f} ms/step")
        # the same steps with energies and virial summed every 10th step
        n_en = max(10, args.steps // 10 * 10)
        t2 = time.perf_counter()
        ctx.md_run(n_en, DT_FS, energies=True, every=10)
        t3 = time.perf_counter()
        log(f"energies every 10th step: {1e3 * (t3 - t2) / n_en:.3f} ms/step; counting pairs")
        cnt, _ = ctx.census(0, x)
    # the drop-in call (SystemForce::evaluate through pm_evaluate): host
    # positions in, host forces and energies out, so PCIe is included; the
    # list is built and energies and virial are summed on every call
    with L.Context(desc(0), [system]) as ectx:
        fe = np.zeros_like(x)
        ee = np.zeros(L.NSLOTS)
        for _ in range(3):
            ectx.evaluate(0, x, fe, ee)
        n_ev = 20
        t6 = time.perf_counter()
        for _ in range(n_ev):
            ectx.evaluate(0, x, fe, ee)
        t7 = time.perf_counter()
    ev_ms = 1e3 * (t7 - t6) / n_ev
    log(f"drop-in pm_evaluate (host buffers, PCIe included): {ev_ms:.3f} ms per evaluation")
    # the pair kernel's launch duration: HIP events on the launch stream around
    # every pair launch, over the same number of steps from the timed region's
    # start state (run eagerly: a replayed graph holds no per-launch events)
    pair_ms, eager_ms = 0.0, None
    if not args.no_timing:
        with L.Context(desc(L.FLAG_TIMING), [system]) as tctx:
            tctx.md_set_state(x_start, v_start)
            tctx.md_run(args.warmup, DT_FS, energies=False)
            t4 = time.perf_counter()
            tctx.md_run(args.steps, DT_FS, energies=False)
            t5 = time.perf_counter()
            tst = tctx.stats()
        pair_ms = tst["pair_ms"] / max(1.0, tst["pair_launches"])
        eager_ms = 1e3 * (t5 - t4) / args.steps
        log(f"pair kernel {pair_ms:.4f} ms per launch over {int(tst['pair_launches'])} launches "
            f"(eager steps with events: {eager_ms:.3f} ms/step)")
    ms_step = 1e3 * (t1 - t0) / args.steps
    ms_en = 1e3 * (t3 - t2) / n_en
    steps_s = 1e3 / ms_step
    pairs = int(cnt)
    achieved = pairs * FLOP_PER_PAIR / (pair_ms * 1e-3) / 1e12 if pair_ms > 0 else 0.0
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
        "data": "synthetic (seeded 32,000-water lattice box, ProtoMol's 300 K velocity start, equilibrated untimed)",
        "config": {"workload": "water 32,000 mol / 96,000 atoms, box 98.6 A, LJ C2 8-10 A + Ewald real "
                               "alpha 0.32 Rc 10 A + bond/angle, leapfrog dt 1 fs, pair list every 10 steps",
                   "atoms": system.n, "pairs_per_step": pairs, "skin_A": args.skin, "rebuild_interval": REBUILD,
                   "md_path": "10-step graphs of kick/drift, list check, pair and local passes, replayed",
                   "energies": "forces only on the timed steps, energies and virial at the last step "
                               "(energies_every_10 sums them every 10th step)",
                   "rebuilds": st["rebuilds"], "max_displacement_A": st["max_displacement"],
                   "list_entries": st["list_entries"],
                   "list_efficiency": pairs / max(1.0, 64.0 * st["list_entries"]),
                   "equilibration_steps": args.equil, "cell_scale": args.cell_scale or 1.0,
                   "pair_waves": args.pair_waves or 4, "pair_split": args.pair_split or 12},
        "ns_per_day": steps_s * DT_FS * 0.0864,
        "energies_every_10": {"ms_per_step": ms_en, "ns_per_day": 1e3 / ms_en * DT_FS * 0.0864,
                              "pair_interactions_per_s": pairs * 1e3 / ms_en, "steps": n_en},
        "drop_in_evaluate": {"ms_per_evaluation": ev_ms, "pair_interactions_per_s": pairs * 1e3 / ev_ms,
                             "what": "pm_evaluate with host positions in and host forces and energies out "
                                     "(PCIe included), list built and energies summed on every call; not "
                                     "the headline"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": pair_kernel_traffic(),
                     "kernel": "k_pair<LJ, EWALD_REAL> (forces)", "kernel_ms": pair_ms,
                     "flop_per_pair": FLOP_PER_PAIR,
                     "kernel_timing": "HIP events around each pair launch, eager steps from the timed "
                                      "region's start state", "eager_ms_per_step": eager_ms},
    }
    O = load_oracle()
    if O is not None:
        fo = oracle_water32k(O, system, x=x, nthreads=box_threads())
        out["force_err"] = force_error(f, fo)
        log(f"force error vs FP64: {out['force_err']}")
    return out, system


def pair_kernel_traffic():
    """HBM bytes per launch of the pair kernel at this workload: FETCH_SIZE and
    WRITE_SIZE from two rocprofv3 --pmc runs, each scaled by the counter's
    reading for a device copy of known size (profiles/pmc_k_pair_water96k.txt)."""
    path = os.path.join(ROOT, "profiles", "pmc_k_pair_water96k.txt")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split()[1])
    except OSError:
        pass
    return None


def water1k(args):
    """3,000-atom water: one force/energy/virial evaluation and 100 leapfrog
    steps at 0.5 fs from ProtoMol's 300 K velocity start."""
    w = W.water1k()
    v0 = S.maxwell_velocities(w, 300.0, seed=1001)
    steps, dt = 100, 0.5
    with L.Context(W.desc_water1k(flags=L.FLAG_VIRIAL), [w]) as ctx:
        f = np.zeros((w.n, 3))
        e = np.zeros(L.NSLOTS)
        ctx.evaluate(0, w.x, f, e)
        t_eval = timed(lambda: ctx.evaluate(0, w.x, np.zeros((w.n, 3)), np.zeros(L.NSLOTS)), 0.2, at_least=5)
        pairs = int(ctx.census(0, w.x)[0])
        ctx.md_set_state(w.x, v0)
        ctx.md_run(steps, dt, energies=True)   # warm: graph capture
        ctx.md_set_state(w.x, v0)
        t0 = time.perf_counter()
        ctx.md_run(steps, dt, energies=True)
        t_md = time.perf_counter() - t0
    ms = 1e3 * t_md / steps
    out = {"workload": "water 1,000 mol / 3,000 atoms, box 31.04 A, LJ + Coulomb both C2 8-10 A + bond/angle; "
                       "1 evaluation + 100 leapfrog steps at 0.5 fs from ProtoMol's 300 K start",
           "pairs": pairs, "evaluate_ms": 1e3 * t_eval, "ms_per_step": ms,
           "ns_per_day": 1e3 / ms * dt * 0.0864, "pair_interactions_per_s": pairs * 1e3 / ms,
           "run_ms": 1e3 * (t_eval + t_md)}
    O = load_oracle()
    if O is not None:
        out["force_err"] = force_error(f, oracle_water1k(O, w, nthreads=box_threads()))
        nt = box_threads()

        def cpu_run(threads):
            v = v0.copy()
            x = w.x.copy()
            F = oracle_water1k(O, w, x=x, nthreads=threads)
            h = dt * S.INV_TIMEFACTOR
            m = w.mass[:, None]
            for i in range(steps):
                v = v + F * h * (0.5 if i == 0 else 1.0) / m
                x = x + v * h
                F = oracle_water1k(O, w, x=x, nthreads=threads)
            return v + F * 0.5 * h / m

        t1 = timed(lambda: cpu_run(1), 0.0)
        ta = timed(lambda: cpu_run(nt), 0.0)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        try:
            import refio
            rec = dict(refio.system_records(w), mode=np.int32(1), term_a=np.int32(1), term_b=np.int32(2),
                       sw_a=np.array((2, 8.0, 10.0, 0, 0), float), sw_b=np.array((2, 8.0, 10.0, 0, 0), float),
                       alg_cutoff=10.0, cellsize=5.0, census=np.int32(0), repeat=np.int32(20))
            tr = ref_seconds(rec)
            tr = None if tr is None else tr / 20.0
        except Exception:
            tr = None
        per = lambda t: None if t is None else t / (steps + 1)
        out["cpu_baseline"] = {
            "value": pairs / per(ta), "unit": "pair interactions/s", "cores": nt, "kind": "port",
            "sample": f"the whole run (1 + 100 force evaluations and the leapfrog updates) on the host: "
                      f"{ta:.2f} s on {nt} threads, {t1:.2f} s on 1 thread",
            "ns_per_day": 1.0 / per(ta) * dt * 0.0864,
            "threads_1": {"value": pairs / per(t1), "ns_per_day": 1.0 / per(t1) * dt * 0.0864, "run_s": t1},
            "reference_1thread": {"value": None if tr is None else pairs / tr,
                                  "ns_per_day": None if tr is None else 1.0 / tr * dt * 0.0864, "s_per_eval": tr,
                                  "kind": "reference", "what": "OneAtomPairTwo<PBC,C2,LJ,C2,Coulomb> cutoff force "
                                                              "of the reference's classes, pairs only"},
            **cpu_info()}
    return out


def batched(args):
    """64 water boxes (log-uniform 500-16,000 molecules), MD in one context vs one context each."""
    sizes = W.ensemble64_sizes(args.batch_systems)
    log(f"batched: generating {len(sizes)} water boxes ({sum(sizes) * 3} atoms)")
    systems = W.waters_at_density(sizes, 500, workers=box_threads())
    v0 = [S.maxwell_velocities(s, 300.0, seed=900 + k) for k, s in enumerate(systems)]
    base = W.desc_water32k(args.skin)
    log("batched: autotuning launch shapes on the batch")
    best, table = autotune.tune(base, systems, v0, DT_FS, steps=20, warm=10, log=log)
    tuned = autotune.apply(base, best)
    steps = args.batch_steps
    with L.Context(tuned, systems) as ctx:
        ctx.md_set_state(np.concatenate([s.x for s in systems]), np.concatenate(v0))
        ctx.md_run(20, DT_FS, energies=False)
        t0 = time.perf_counter()
        ctx.md_run(steps, DT_FS, energies=False)
        t_batch = time.perf_counter() - t0
        rebuilds = ctx.stats()["rebuilds"]
        xs, _, fs, _ = ctx.md_get_state()
    log(f"batched: {steps} steps of {len(systems)} systems in {t_batch:.2f} s")
    # the plain kernel: each system in its own context, one after another
    t_plain = 0.0
    for k, s in enumerate(systems):
        with L.Context(base, [s]) as ctx:
            ctx.md_set_state(s.x, v0[k])
            ctx.md_run(20, DT_FS, energies=False)
            t0 = time.perf_counter()
            ctx.md_run(steps, DT_FS, energies=False)
            t_plain += time.perf_counter() - t0
    log(f"batched: plain one-system-per-context total {t_plain:.2f} s")
    n = len(systems)
    out = {"workload": f"{n} water boxes, {min(sizes)}-{max(sizes)} molecules (log-uniform 500-16,000), "
                       f"{sum(sizes) * 3} atoms, LJ C2 + Ewald real + bond/angle, {steps} steps at 1 fs",
           "system_steps_per_s": n * steps / t_batch, "plain_system_steps_per_s": n * steps / t_plain,
           "speedup_vs_plain": t_plain / t_batch, "ms_per_step": 1e3 * t_batch / steps,
           "atom_steps_per_s": sum(sizes) * 3 * steps / t_batch, "rebuilds": rebuilds,
           "tuned": best, "tuning": [[p, ms] for p, ms in table]}
    O = load_oracle()
    if O is not None:
        offs = np.cumsum([0] + [s.n for s in systems])
        nt = box_threads()
        errs = {}
        for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
            sl = slice(offs[k], offs[k + 1])
            errs[f"box{k}_{systems[k].n}_atoms"] = force_error(fs[sl], oracle_water32k(O, systems[k], x=xs[sl],
                                                                                      nthreads=nt))
        out["force_err"] = errs
        # CPU: a seeded sample of 4 boxes, one evaluation each, scaled to
        # the ensemble by atom count (the work per atom is the same at one density)
        rng = np.random.default_rng(5)
        sample = sorted(rng.choice(n, size=min(4, n), replace=False).tolist())
        res = {}
        for tag, threads in (("threads_1", 1), ("threads_all", nt)):
            t = sum(timed(lambda k=k: oracle_water32k(O, systems[k], nthreads=threads), 0.0) for k in sample)
            atoms = sum(systems[k].n for k in sample)
            step_s = t / atoms * offs[-1]      # one step of all 64 systems
            res[tag] = {"system_steps_per_s": n / step_s, "threads": threads, "sample_s": t}
        out["cpu_baseline"] = {"value": res["threads_all"]["system_steps_per_s"], "unit": "system steps/s",
                               "cores": nt, "kind": "port",
                               "sample": f"one force evaluation of each of boxes {sample} "
                                         f"({sum(systems[k].n for k in sample)} atoms), scaled by atom count to "
                                         f"all {n} systems", **res, **cpu_info()}
    return out


def _requests(templates, n_req, rate_hz, steps, seed):
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for k in range(n_req):
        s = templates[int(rng.integers(len(templates)))]
        kind = "md" if rng.random() < 0.5 else "eval"
        shift = rng.uniform(-0.5, 0.5, size=3) * s.box
        x = s.x + shift
        v = S.maxwell_velocities(s, 300.0, seed=10_000 + k) if kind == "md" else None
        reqs.append(serve.Request(rid=k, kind=kind, system=s, x=x, v=v, steps=steps if kind == "md" else 0,
                                  arrival=t))
        if rate_hz:
            t += rng.exponential(1.0 / rate_hz)
    return reqs


def served(args):
    """Poisson request stream on water boxes of 1,000-32,000 molecules."""
    n_templates = 16
    sizes = W.served_sizes(n_templates)
    log(f"served: generating {n_templates} water boxes for the request stream")
    templates = W.waters_at_density(sizes, 700, workers=box_threads())
    desc = W.desc_water32k(args.skin, rebuild=0)
    n_req, steps = args.serve_requests, 200
    reserve = int(args.serve_reserve_gib * (1 << 30))
    # plain: each request on its own through the drop-in API (a context per
    # request: evaluate, or set state + LeapfrogIntegrator::run), one at a time
    reqs = _requests(templates, n_req, 0.0, steps, seed=5)
    log(f"served: plain API, {n_req} requests one at a time")
    t0 = time.perf_counter()
    for r in reqs:
        with L.Context(desc, [r.system]) as ctx:
            if r.kind == "eval":
                f = np.zeros_like(r.x)
                e = np.zeros(L.NSLOTS)
                ctx.evaluate(0, r.x, f, e)
            else:
                ctx.md_set_state(r.x, r.v)
                ctx.md_run(steps, DT_FS, energies=True)
                ctx.md_get_state()
    t_plain = time.perf_counter() - t0
    plain_rate = n_req / t_plain
    log(f"served: plain API {plain_rate:.1f} requests/s")
    # the same requests one at a time inside the server (no batching)
    reqs = _requests(templates, n_req, 0.0, steps, seed=5)
    with L.Context(desc, [], reserve_bytes=reserve) as ctx:
        srv = serve.Server(ctx, dt_fs=DT_FS, chunk=steps)
        t0 = time.perf_counter()
        for r in reqs:
            srv.submit(r)
            while not srv.idle():
                srv.tick()
        t_serial = time.perf_counter() - t0
    serial_rate = n_req / t_serial
    log(f"served: one-at-a-time server {serial_rate:.1f} requests/s")
    # continuous batching, every request waiting at t = 0 (capacity)
    reqs = _requests(templates, n_req, 0.0, steps, seed=5)
    with L.Context(desc, [], reserve_bytes=reserve) as ctx:
        srv = serve.Server(ctx, dt_fs=DT_FS, chunk=args.serve_chunk)
        t0 = time.perf_counter()
        done_sat = srv.run(reqs, max_seconds=900)
        t_sat = time.perf_counter() - t0
    sat_rate = n_req / t_sat
    log(f"served: batched capacity {sat_rate:.1f} requests/s")
    # Poisson arrivals at fixed fractions of that capacity: latency
    loads = {}
    for load in args.serve_loads:
        rate = load * sat_rate
        reqs = _requests(templates, n_req, rate, steps, seed=6)
        with L.Context(desc, [], reserve_bytes=reserve) as ctx:
            srv = serve.Server(ctx, dt_fs=DT_FS, chunk=args.serve_chunk)
            t0 = time.perf_counter()
            done = srv.run(reqs, max_seconds=900)
            t_poi = time.perf_counter() - t0
        lat = serve.latency_summary(done)
        log(f"served: Poisson {rate:.1f}/s -> p50 {lat['p50_ms']:.1f} ms, p99 {lat['p99_ms']:.1f} ms")
        loads[f"{load:g}"] = {"rate_per_s": rate, "requests_per_s": n_req / t_poi, "latency": lat,
                              "ticks": srv.ticks}
    out = {"workload": f"{n_req} requests, 50/50 force evaluation / {steps}-step MD segment, water boxes "
                       f"{min(sizes)}-{max(sizes)} molecules ({n_templates} seeded boxes, each request "
                       f"shifted and with its own velocities), LJ C2 + Ewald real + bond/angle",
           "plain_requests_per_s": plain_rate, "one_at_a_time_requests_per_s": serial_rate,
           "batched_capacity_requests_per_s": sat_rate, "speedup_vs_plain": sat_rate / plain_rate,
           "speedup_vs_one_at_a_time": sat_rate / serial_rate, "poisson": loads,
           "chunk_steps": args.serve_chunk, "reserve_gib": args.serve_reserve_gib}
    O = load_oracle()
    if O is not None:
        # one evaluation request (the largest box among them) against the oracle
        ev = max((r for r in done_sat if r.kind == "eval"), key=lambda r: r.system.n)
        out["force_err"] = force_error(ev.f, oracle_water32k(O, ev.system, x=ev.x, nthreads=box_threads()))
        # CPU serving the same stream one request at a time (FIFO): the
        # oracle's time per atom-evaluation on the smallest and largest box,
        # applied to each request (1 evaluation, or 201 for a 200-step segment)
        nt = box_threads()
        small, large = min(templates, key=lambda s: s.n), max(templates, key=lambda s: s.n)
        reqs = _requests(templates, n_req, 0.0, steps, seed=5)
        res = {}
        for tag, threads in (("threads_1", 1), ("threads_all", nt)):
            per_atom = np.mean([timed(lambda s=s: oracle_water32k(O, s, nthreads=threads), 0.0) / s.n
                                for s in (small, large)])
            total = sum(per_atom * r.system.n * (1 if r.kind == "eval" else steps + 1) for r in reqs)
            res[tag] = {"requests_per_s": n_req / total, "threads": threads}
        out["cpu_baseline"] = {"value": res["threads_all"]["requests_per_s"], "unit": "requests/s", "cores": nt,
                               "kind": "port", "sample": "one force evaluation of the smallest and the largest box; "
                               "the stream's 2,000 requests served one at a time at that per-atom cost", **res,
                               **cpu_info()}
    return out


def gb_batch(args):
    """32 GB/OBC bead chains of 5,000 atoms, all pairs, MD in one context."""
    n_chain, n_atoms = args.gb_chains, 5000
    log(f"gb: generating {n_chain} chains of {n_atoms} atoms")
    chains = W.chains32(n_chain, n_atoms, workers=box_threads())
    v0 = [S.maxwell_velocities(c, 300.0, seed=4000 + k) for k, c in enumerate(chains)]
    steps = args.gb_steps
    pairs = n_chain * n_atoms * (n_atoms - 1) // 2
    with L.Context(W.desc_chains(), chains) as ctx:
        ctx.md_set_state(np.concatenate([c.x for c in chains]), np.concatenate(v0))
        ctx.md_run(5, DT_FS, energies=False)
        t0 = time.perf_counter()
        ctx.md_run(steps, DT_FS, energies=False)
        t = time.perf_counter() - t0
        xs, _, fs, _ = ctx.md_get_state()
    # the GB passes' own time: HIP events around them on their stream
    with L.Context(W.desc_chains(L.FLAG_TIMING), chains) as ctx:
        ctx.md_set_state(np.concatenate([c.x for c in chains]), np.concatenate(v0))
        ctx.md_run(10, DT_FS, energies=False)
        st = ctx.stats()
    gb_ms = st["pair_ms"] / max(1.0, st["pair_launches"])
    achieved = pairs * GB_FLOP_PER_PAIR / (gb_ms * 1e-3) / 1e12
    log(f"gb: {steps} steps in {t:.2f} s; GB passes {gb_ms:.3f} ms per evaluation")
    out = {"workload": f"{n_chain} chains x {n_atoms} atoms, GB/OBC + LJ + Coulomb, all pairs, + bonds/angles, "
                       f"{steps} steps at 1 fs",
           "ms_per_step": 1e3 * t / steps, "system_steps_per_s": n_chain * steps / t,
           "pair_interactions_per_s": pairs * steps / t,
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "kernel": "k_gb_sym<1> + k_gb_sums<1> + k_gb_sym<2> + "
                        "k_gb_sums<2> + k_gb_pairs + k_gb_fsum", "kernel_ms": gb_ms,
                        "flop_per_pair": GB_FLOP_PER_PAIR}}
    O = load_oracle()
    if O is not None:
        c0 = chains[0]
        fo, _, _ = O.gb_evaluate(c0, x=xs[:n_atoms])
        fl, _ = O.vacuum_lj_coulomb(c0, x=xs[:n_atoms])
        fbd, _ = O.bonded(c0, x=xs[:n_atoms])
        out["force_err"] = force_error(fs[:n_atoms], fo + fl + fbd)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        tr = None
        try:
            import refio
            rec = dict(refio.system_records(c0), mode=np.int32(3), gb_rho=c0.gb_radius, gb_S=c0.gb_scaling,
                       gb_params=np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3]), cellsize=np.int32(5),
                       gb_ljcoul=np.int32(1))
            tr = ref_seconds(rec)
        except Exception as e:
            log(f"gb reference timing unavailable: {e}")
        if tr:
            out["cpu_baseline"] = {"value": n_chain * (n_atoms * (n_atoms - 1) // 2) / (n_chain * tr),
                                   "unit": "pair interactions/s", "cores": 1, "kind": "reference",
                                   "sample": f"one force evaluation of one chain by the reference's own classes "
                                             f"(GBBornRadii, GBPartialSum, GBForce, LJ, Coulomb through "
                                             f"NonbondedSimpleFullSystemForce; the O(N) bonds and angles "
                                             f"left out), {tr:.2f} s, times {n_chain} chains",
                                   "system_steps_per_s": 1.0 / tr, **cpu_info()}
    return out


def ewald_full(args):
    """The 96,000-atom water box with the reciprocal-space sum added."""
    system = getattr(args, "_system", None)
    if system is None:
        log("ewald: building the 96,000-atom water box")
        system = W.water32k()
    v0 = S.maxwell_velocities(system, 300.0, seed=2025)
    d = W.desc_water32k(args.skin, cell_scale=args.cell_scale)
    d.flags |= L.FLAG_EWALD_RECIPROCAL
    rec = []
    with L.Context(d, [system]) as ctx:
        f = np.zeros((system.n, 3))
        e = np.zeros(L.NSLOTS)
        for k in range(args.ewald_evals):
            f[:] = 0.0
            ctx.evaluate(0, system.x, f, e)
            st = ctx.stats()
            rec.append(st["reciprocal_ms"])
            log(f"ewald: evaluation {k + 1}: reciprocal {st['reciprocal_ms']:.3f} ms, device {st['device_ms']:.3f} ms")
        nk = int(st["kvectors"])
        ctx.md_set_state(system.x, v0)
        ctx.md_run(3, DT_FS, energies=False)
        t0 = time.perf_counter()
        ctx.md_run(args.ewald_steps, DT_FS, energies=False)
        t = time.perf_counter() - t0
    log(f"ewald: {args.ewald_steps} MD steps in {t:.2f} s")
    rms = float(np.median(rec[1:] if len(rec) > 1 else rec))
    work = float(nk) * system.n
    achieved = work * EW_FLOP_PER_KATOM / (rms * 1e-3) / 1e12
    out = {"workload": "water 32,000 mol / 96,000 atoms, box 98.6 A, LJ C2 8-10 A + full Ewald alpha 0.32 "
                       "(Rc 10 A, kc 2.05 A^-1) + bond/angle, leapfrog dt 1 fs",
           "kvectors": nk, "ms_per_step": 1e3 * t / args.ewald_steps, "reciprocal_ms": rms,
           "katom_pairs_per_s": work / (rms * 1e-3),
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS,
                        "kernel": "k_ew_tables + k_ew_sf + k_ew_ssum + k_ew_force + k_ew_fsum (FP64)",
                        "flop_per_katom": EW_FLOP_PER_KATOM}}
    O = load_oracle()
    if O is not None:
        fr, _ = O.ewald_reciprocal(system, ALPHA, ACCURACY, nthreads=box_threads())
        out["force_err"] = force_error(f, oracle_water32k(O, system, nthreads=box_threads()) + fr)
        if not args.no_cpu:
            out["cpu_baseline"] = ewald_cpu_baseline(O, args.cpu_seconds)
    return out


def ewald_cpu_baseline(O, seconds):
    """The oracle's reciprocal term (the reference's algorithm) on this host,
    on a bounded sample: 8,000 waters (24,000 atoms, same alpha and kc)."""
    nt = box_threads()
    w = S.water_box(8000, (8000 / W.DENSITY) ** (1.0 / 3.0), seed=11, min_contact=2.0)
    nk = len(O.ewald_kvectors(w.box, ALPHA, ACCURACY)[0])
    per = timed(lambda: O.ewald_reciprocal(w, ALPHA, ACCURACY, nthreads=nt), seconds * 0.5)
    log(f"ewald cpu baseline: {per:.3f} s per evaluation on {nt} threads")
    return {"value": nk * w.n / per, "unit": "(k-vector, atom) pairs/s", "cores": nt, "kind": "port",
            "sample": f"reciprocal-term evaluations of 8,000 waters ({w.n} atoms, {nk} k-vectors), "
                      f"{per:.3f} s each on {nt} threads", **cpu_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--skin", type=float, default=1.2)
    ap.add_argument("--equil", type=int, default=300, help="untimed equilibration steps (rescaled to 300 K)")
    ap.add_argument("--cell-scale", type=float, default=1.25,
                    help="column width factor (1.25: the autotuner's choice for water at 0.1 atoms/A^3)")
    ap.add_argument("--pair-waves", type=int, default=0, help="pair-kernel waves per workgroup (0 = default 4)")
    ap.add_argument("--pair-split", type=int, default=0, help="waves per super-cluster list (0 = default 12)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true",
                    help="skip the eager pass that times each pair launch with HIP events (no roofline figure)")
    ap.add_argument("--forms", default="water1k,batched,served,gb,ewald",
                    help="comma list of extra forms beside the headline (water1k, batched, served, gb, ewald) "
                         "or 'none'")
    ap.add_argument("--only", default="", help="run one extra form alone: water1k, batched, served, gb or ewald")
    ap.add_argument("--batch-systems", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests per served phase (the stream's 2,000)")
    ap.add_argument("--serve-chunk", type=int, default=20)
    ap.add_argument("--serve-loads", type=float, nargs="+", default=[0.5, 0.8])
    ap.add_argument("--serve-reserve-gib", type=float, default=24.0)
    ap.add_argument("--gb-chains", type=int, default=32)
    ap.add_argument("--gb-steps", type=int, default=200)
    ap.add_argument("--ewald-evals", type=int, default=4)
    ap.add_argument("--ewald-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent systems shard by placement, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    name, ncu, hbm = L.device_info()
    log(f"device: {name}, {ncu} CUs, {hbm:.0f} GiB")
    forms = {"water1k": water1k, "batched": batched, "served": served, "gb": gb_batch, "ewald": ewald_full}
    if args.only:
        print(json.dumps({args.only: forms[args.only](args), "device": name}), flush=True)
        return
    out, system = headline(args)
    out["device"] = name
    args._system = system
    for f in [f for f in args.forms.split(",") if f and f != "none"]:
        out[f] = forms[f](args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(system, out["config"]["pairs_per_step"], args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
