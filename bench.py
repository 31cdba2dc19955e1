#!/usr/bin/env python3
"""Benchmarks of the nonbonded force path on one MI355X.

Headline (the JSON line's ``value``): device-resident MD of 32,000 three-site
waters (96,000 atoms, cubic box 98.6 A) with Lennard-Jones (C2 switch
8-10 A) plus Ewald real-space Coulomb (alpha 0.32 A^-1, Rc 10 A) and the
harmonic water model, leapfrog at dt = 1 fs, pair list rebuilt every 10 steps
(and sooner if an atom moves more than half the list buffer).  The start is a
lattice with seeded orientations, relieved of H-H contacts under 2 A and
equilibrated for a few hundred untimed steps with velocity rescaling to 300 K.

Beside it, in the same line:
  roofline      the pair kernel's FP32 rate (HIP events around every
                pair-kernel launch on the kernel's own stream) against peak
  cpu_baseline  the FP64 oracle (the reference algorithm) on this host's cores
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
from multiprocessing import get_context

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from protomol_amd import _lib as L  # noqa: E402
from protomol_amd import autotune  # noqa: E402
from protomol_amd import serve  # noqa: E402
from protomol_amd import systems as S  # noqa: E402

N_MOL, BOX, DT_FS, REBUILD = 32000, 98.6, 1.0, 10
ALPHA, ACCURACY = 0.32, math.exp(-10.24)
DENSITY = 0.0334  # molecules / A^3
# algorithmic work per within-cutoff pair of the LJ(C2) + Ewald-real kernel:
# distance 11, LJ 11, C2 switch 13, Ewald real 23, force combine/accumulate 10,
# energies 2, virial 12 (add/sub/mul = 1, FMA = 2, each rsqrt/rcp/exp = 1)
FLOP_PER_PAIR = 82
PEAK_FP32_TFLOPS = 157.3
# reciprocal-space work per (k-vector, atom) pair, counted on
# NonbondedFullEwaldSystemForce::reciprocalTerm: structure factor 8 (sinA,
# cosA from the (h, k) products and z phases 6, sums 2), force 10 (the
# sin/cos cross term 3, times a 1, times k 3, accumulate 3)
EW_FLOP_PER_KATOM = 18
PEAK_FP64_TFLOPS = 78.6
WORKERS = 16


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def force_desc(skin, timing=False, rebuild=REBUILD, **shape):
    d = L.ForceDesc()
    d.term_a = L.TERM_LJ
    d.switch_a = L.switch(L.SWITCH_C2, 8.0, 10.0)
    d.term_b = L.TERM_EWALD_REAL
    d.cutoff = 10.0
    d.ewald_alpha = ALPHA
    d.ewald_accuracy = ACCURACY
    d.ewald_correction = 1
    d.skin = skin
    d.rebuild_interval = rebuild
    d.flags = L.FLAG_BONDED | (L.FLAG_TIMING if timing else 0)
    for k, v in shape.items():
        setattr(d, k, v)
    return d


def _water(args):
    n, seed = args
    return S.water_box(n, (n / DENSITY) ** (1.0 / 3.0), seed=seed, min_contact=2.0)


def make_waters(sizes, seed0):
    """Water boxes at 0.0334 molecules/A^3, generated in parallel."""
    jobs = [(n, seed0 + k) for k, n in enumerate(sizes)]
    with get_context("fork").Pool(min(WORKERS, len(jobs))) as pool:
        return pool.map(_water, jobs)


def cpu_baseline(system, seconds):
    """The FP64 oracle (the restatement of the reference's algorithm) timed on
    this host: force evaluations of the 96k-atom system until `seconds` pass."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
    except Exception as e:  # the checker library is not built
        return {"value": None, "unit": "pair interactions/s", "cores": 0, "kind": "port",
                "sample": f"unavailable: {e}"}
    cores = max(1, min(os.cpu_count() or 1, WORKERS))
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


def headline(args):
    log("building the 96,000-atom water box")
    system = S.water_box(N_MOL, BOX, seed=2024, min_contact=2.0)
    v0 = S.maxwell_velocities(system, 300.0, seed=2025)
    shape = {"cell_scale": args.cell_scale, "pair_waves": args.pair_waves, "pair_split": args.pair_split}
    with L.Context(force_desc(args.skin, timing=not args.no_timing, **shape), [system]) as ctx:
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
                   "list_entries": st["list_entries"],
                   "list_efficiency": pairs / max(1.0, 64.0 * st["list_entries"]),
                   "equilibration_steps": args.equil, "cell_scale": args.cell_scale or 1.0,
                   "pair_waves": args.pair_waves or 4, "pair_split": args.pair_split or 12},
        "ns_per_day": steps_s * DT_FS * 0.0864,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": pair_kernel_traffic(),
                     "kernel": "k_pair<LJ, EWALD_REAL>", "kernel_ms": pair_ms,
                     "flop_per_pair": FLOP_PER_PAIR},
    }
    return out, system


def batched(args):
    """64 water boxes (log-uniform 500-16,000 molecules), MD in one context vs one context each."""
    sizes = S.ensemble_sizes(args.batch_systems, 500, 16000, seed=77)
    log(f"batched: generating {len(sizes)} water boxes ({sum(sizes) * 3} atoms)")
    systems = make_waters(sizes, 500)
    v0 = [S.maxwell_velocities(s, 300.0, seed=900 + k) for k, s in enumerate(systems)]
    base = force_desc(args.skin)
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
    return {"workload": f"{n} water boxes, {min(sizes)}-{max(sizes)} molecules (log-uniform 500-16,000), "
                        f"{sum(sizes) * 3} atoms, LJ C2 + Ewald real + bond/angle, {steps} steps at 1 fs",
            "system_steps_per_s": n * steps / t_batch, "plain_system_steps_per_s": n * steps / t_plain,
            "speedup_vs_plain": t_plain / t_batch, "ms_per_step": 1e3 * t_batch / steps,
            "atom_steps_per_s": sum(sizes) * 3 * steps / t_batch, "rebuilds": rebuilds,
            "tuned": best, "tuning": [[p, ms] for p, ms in table]}


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
    sizes = S.ensemble_sizes(n_templates, 1000, 32000, seed=91)
    log(f"served: generating {n_templates} water boxes for the request stream")
    templates = make_waters(sizes, 700)
    desc = force_desc(args.skin, rebuild=0)
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
        srv.run(reqs, max_seconds=900)
        t_sat = time.perf_counter() - t0
    sat_rate = n_req / t_sat
    log(f"served: batched capacity {sat_rate:.1f} requests/s")
    # Poisson arrivals at a fixed fraction of that capacity: latency
    rate = args.serve_load * sat_rate
    reqs = _requests(templates, n_req, rate, steps, seed=6)
    with L.Context(desc, [], reserve_bytes=reserve) as ctx:
        srv = serve.Server(ctx, dt_fs=DT_FS, chunk=args.serve_chunk)
        t0 = time.perf_counter()
        done = srv.run(reqs, max_seconds=900)
        t_poi = time.perf_counter() - t0
        ticks = srv.ticks
    lat = serve.latency_summary(done)
    log(f"served: Poisson {rate:.1f}/s -> p50 {lat['p50_ms']:.1f} ms, p99 {lat['p99_ms']:.1f} ms")
    return {"workload": f"{n_req} requests, 50/50 force evaluation / {steps}-step MD segment, water boxes "
                        f"{min(sizes)}-{max(sizes)} molecules ({n_templates} seeded boxes, each request "
                        f"shifted and with its own velocities), LJ C2 + Ewald real + bond/angle",
            "plain_requests_per_s": plain_rate, "one_at_a_time_requests_per_s": serial_rate,
            "batched_capacity_requests_per_s": sat_rate, "speedup_vs_plain": sat_rate / plain_rate,
            "speedup_vs_one_at_a_time": sat_rate / serial_rate, "poisson_rate_per_s": rate,
            "poisson_requests_per_s": n_req / t_poi, "latency": lat, "ticks": ticks,
            "chunk_steps": args.serve_chunk, "reserve_gib": args.serve_reserve_gib}


def gb_batch(args):
    """32 GB/OBC bead chains of 5,000 atoms, all pairs, MD in one context."""
    n_chain, n_atoms = args.gb_chains, 5000
    log(f"gb: generating {n_chain} chains of {n_atoms} atoms")
    with get_context("fork").Pool(min(WORKERS, n_chain)) as pool:
        chains = pool.starmap(S.gb_chain, [(n_atoms, 3000 + k) for k in range(n_chain)])
    d = L.ForceDesc()
    d.gb_offset, d.gb_alpha, d.gb_beta, d.gb_gamma = 0.09, 1.0, 0.8, 4.85
    d.gb_solute_dielec, d.gb_solvent_dielec = 1.0, 78.3
    d.flags = L.FLAG_GB
    v0 = [S.maxwell_velocities(c, 300.0, seed=4000 + k) for k, c in enumerate(chains)]
    steps = args.gb_steps
    with L.Context(d, chains) as ctx:
        ctx.md_set_state(np.concatenate([c.x for c in chains]), np.concatenate(v0))
        ctx.md_run(5, DT_FS, energies=False)
        t0 = time.perf_counter()
        ctx.md_run(steps, DT_FS, energies=False)
        t = time.perf_counter() - t0
    pairs = n_chain * n_atoms * (n_atoms - 1) // 2
    log(f"gb: {steps} steps in {t:.2f} s")
    return {"workload": f"{n_chain} chains x {n_atoms} atoms, GB/OBC + LJ + Coulomb, all pairs, {steps} steps at 1 fs",
            "ms_per_step": 1e3 * t / steps, "system_steps_per_s": n_chain * steps / t,
            "pair_interactions_per_s": pairs * steps / t}


def ewald_full(args):
    """The 96,000-atom water box with the reciprocal-space sum added."""
    system = getattr(args, "_system", None)
    if system is None:
        log("ewald: building the 96,000-atom water box")
        system = S.water_box(N_MOL, BOX, seed=2024, min_contact=2.0)
    v0 = S.maxwell_velocities(system, 300.0, seed=2025)
    d = force_desc(args.skin, cell_scale=args.cell_scale)
    d.flags |= L.FLAG_EWALD_RECIPROCAL
    rec = []
    with L.Context(d, [system]) as ctx:
        f = np.zeros((system.n, 3))
        e = np.zeros(L.NSLOTS)
        for k in range(args.ewald_evals):
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
    if not args.no_cpu:
        out["cpu_baseline"] = ewald_cpu_baseline(args.cpu_seconds)
    return out


def ewald_cpu_baseline(seconds):
    """The oracle's reciprocal term (the reference's algorithm) on this host,
    on a bounded sample: 8,000 waters (24,000 atoms, same alpha and kc)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
    except Exception as e:
        return {"value": None, "unit": "(k-vector, atom) pairs/s", "cores": 0, "kind": "port",
                "sample": f"unavailable: {e}"}
    cores = max(1, min(os.cpu_count() or 1, WORKERS))
    w = S.water_box(8000, (8000 / DENSITY) ** (1.0 / 3.0), seed=11, min_contact=2.0)
    nk = len(O.ewald_kvectors(w.box, ALPHA, ACCURACY)[0])
    n_eval, t_tot = 0, 0.0
    while t_tot < seconds or n_eval == 0:
        t0 = time.perf_counter()
        O.ewald_reciprocal(w, ALPHA, ACCURACY, nthreads=cores)
        t_tot += time.perf_counter() - t0
        n_eval += 1
        log(f"ewald cpu baseline: evaluation {n_eval} done, {t_tot:.1f} s")
    per = t_tot / n_eval
    return {"value": nk * w.n / per, "unit": "(k-vector, atom) pairs/s", "cores": cores, "kind": "port",
            "sample": f"{n_eval} reciprocal-term evaluation(s) of 8,000 waters ({w.n} atoms, {nk} k-vectors), "
                      f"{per:.3f} s each on {cores} threads"}


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
                    help="no per-launch HIP events (MD steps replay a captured graph; no roofline figure)")
    ap.add_argument("--forms", default="batched,served,gb,ewald",
                    help="comma list of extra forms beside the headline (batched, served, gb, ewald) or 'none'")
    ap.add_argument("--only", default="", help="run one extra form alone: batched, served, gb or ewald")
    ap.add_argument("--batch-systems", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests per served phase (the stream's 2,000)")
    ap.add_argument("--serve-chunk", type=int, default=20)
    ap.add_argument("--serve-load", type=float, default=0.5)
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
    forms = {"batched": batched, "served": served, "gb": gb_batch, "ewald": ewald_full}
    if args.only:
        print(json.dumps({args.only: forms[args.only](args), "device": name}), flush=True)
        return
    out, system = headline(args)
    out["device"] = name
    args._system = system
    for f in [f for f in args.forms.split(",") if f and f != "none"]:
        out[f] = forms[f](args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(system, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
