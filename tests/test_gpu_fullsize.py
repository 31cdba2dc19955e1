"""Parity at the benchmarked sizes: every workload bench.py measures, at its
real size, through the C ABI, against the FP64 oracle (itself pinned to the
reference's classes) and against the same systems run alone.

  water1k     one evaluation (pair set bit-exact, forces, energies, virial)
              and 100 leapfrog steps at 0.5 fs from ProtoMol's own 300 K
              velocity start against the oracle's FP64 leapfrog
  water32k    one evaluation, and the MD state after 100 steps at 1 fs: the
              live pair list's set bit-exact and the forces the step loop
              holds against the oracle at those positions
  ensemble64  the 64 boxes evaluated and advanced in one context, bitwise
              the same as each box alone; the smallest and largest box
              against the oracle
  served      requests on boxes of 1,000-32,000 molecules, served together,
              bitwise the same as each request alone
  chains32    32 GB chains of 5,000 atoms evaluated and advanced in one
              context, bitwise the same as each alone; two chains against
              the oracle

Forces: RMS relative error <= 1e-5 against FP64 (the stated tolerance of
FP32 pair arithmetic); the largest per-atom error (relative, floor
1 kcal/mol/A) is printed and held to 1e-3.
"""
import os

import numpy as np
import pytest

import oracle as O
import pmtest as T
from protomol_amd import _lib as L
from protomol_amd import serve
from protomol_amd import systems as S
from protomol_amd import workloads as W

pytestmark = pytest.mark.gpu

RTOL_F = 1e-5
MAX_ATOM = 1e-3


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def oracle_water1k(w, x=None, virial=True):
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    f, e, cen = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw, cellsize=5.0, virial=virial, x=x,
                                   nthreads=_threads())
    fb, eb = O.bonded(w, x=x, virial=virial)
    return f + fb, e + eb, cen


def oracle_water32k(w, x=None, virial=True):
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    fl, el, cl = O.nonbonded_cutoff(w, O.TERM_LJ, sw, cellsize=5.0, virial=virial, x=x, nthreads=_threads())
    fe, ee, ce, _ = O.ewald_real(w, W.EWALD_ALPHA, W.EWALD_ACCURACY, cellsize=5.0, virial=virial, x=x,
                                 nthreads=_threads())
    fb, eb = O.bonded(w, x=x, virial=virial)
    assert cl == ce  # LJ (10 A switch cutoff) and Ewald (Rc 10 A) see one pair set
    return fl + fe + fb, el + ee + eb, ce


def check_forces(tag, f, fo):
    rms, mx = T.rms_rel(f, fo), T.max_rel_atom(f, fo)
    print(f"  {tag}: force error vs FP64 oracle: RMS relative {rms:.2e}, max per atom {mx:.2e}", flush=True)
    assert rms <= RTOL_F, rms
    assert mx <= MAX_ATOM, mx


def check_energies(e, eo, slots, npairs):
    for s in slots:
        tol = 1e-5 * max(1.0, abs(eo[s])) + 1e-7 * npairs
        assert abs(e[s] - eo[s]) <= tol, (s, e[s], eo[s], tol)


def check_virial(e, eo, npairs):
    tol = 1e-5 * max(1.0, np.max(np.abs(eo[7:16]))) + 1e-7 * npairs
    assert np.max(np.abs(e[7:16] - eo[7:16])) <= tol, (e[7:16], eo[7:16], tol)


def evaluate(ctx, sid, x):
    f = np.zeros((x.shape[0], 3))
    e = np.zeros(L.NSLOTS)
    ctx.evaluate(sid, x, f, e)
    return f, e


@pytest.fixture(scope="module")
def water1k():
    return W.water1k()


@pytest.fixture(scope="module")
def water32k():
    return W.water32k()


# ------------------------------------------------------------------ water1k

def test_water1k_evaluate_vs_oracle(water1k):
    print("case water1k evaluate (3,000 atoms)", flush=True)
    w = water1k
    with L.Context(W.desc_water1k(flags=L.FLAG_VIRIAL), [w]) as ctx:
        f, e = evaluate(ctx, 0, w.x)
        cen = ctx.census(0, w.x)
    fo, eo, co = oracle_water1k(w)
    assert cen == co
    check_forces("water1k", f, fo)
    check_energies(e, eo, (0, 1, 2, 3), cen[0])
    check_virial(e, eo, cen[0])


def test_water1k_md_100_steps_vs_oracle_leapfrog(water1k):
    print("case water1k md (100 steps at 0.5 fs)", flush=True)
    w = water1k
    v0 = S.maxwell_velocities(w, 300.0, seed=1001)   # ProtoMol's velocity start, randomtype 0

    def forces(x):
        f, e, _ = oracle_water1k(w, x=x, virial=False)
        return f, e

    xo, vo, Fo, Eo = T.leapfrog_oracle(forces, w.x.copy(), v0.copy(), w.mass, 0.5, 100)
    with L.Context(W.desc_water1k(), [w]) as ctx:
        ctx.md_set_state(w.x, v0)
        ctx.md_run(100, 0.5, energies=True)
        x, v, F, E = ctx.md_get_state()
        cen = ctx.census(0, None)
        st = ctx.stats()
    # the trajectory: FP32 pair forces against FP64 ones, 100 steps apart
    dx = float(np.max(np.abs(x - xo)))
    print(f"  after 100 steps: max |dx| {dx:.2e} A, v RMS rel {T.rms_rel(v, vo):.2e}, "
          f"rebuilds {st['rebuilds']:.0f}", flush=True)
    assert dx < 5e-4
    assert T.rms_rel(v, vo) < 1e-3
    for s in (0, 1, 2, 3):
        assert abs(E[0][s] - Eo[s]) <= 1e-4 * max(1.0, abs(Eo[s])), (s, E[0][s], Eo[s])
    # the forces the step loop holds, against FP64 at the same positions
    fo, _, co = oracle_water1k(w, x=x, virial=False)
    assert cen == co
    check_forces("water1k after 100 steps", F, fo)


# ----------------------------------------------------------------- water32k

def test_water32k_evaluate_vs_oracle(water32k):
    print("case water32k evaluate (96,000 atoms)", flush=True)
    w = water32k
    with L.Context(W.desc_water32k(flags=L.FLAG_VIRIAL), [w]) as ctx:
        f, e = evaluate(ctx, 0, w.x)
        cen = ctx.census(0, w.x)
    fo, eo, co = oracle_water32k(w)
    print(f"  pairs within the cutoff: {cen[0]}", flush=True)
    assert cen == co
    check_forces("water32k", f, fo)
    check_energies(e, eo, (0, 1, 2, 3), cen[0])
    check_virial(e, eo, cen[0])


def test_water32k_md_state_vs_oracle(water32k):
    """After 100 steps at 1 fs (list every 10 steps or on the displacement
    test): the pair set the live list yields is the FP64 one, and the forces
    the step loop holds are the oracle's at those positions."""
    print("case water32k md state (100 steps at 1 fs)", flush=True)
    w = water32k
    v0 = S.maxwell_velocities(w, 300.0, seed=2025)
    with L.Context(W.desc_water32k(), [w]) as ctx:
        ctx.md_set_state(w.x, v0)
        ctx.md_run(100, 1.0, energies=True)
        x, v, F, E = ctx.md_get_state()
        cen = ctx.census(0, None)
        st = ctx.stats()
    fo, eo, co = oracle_water32k(w, x=x, virial=False)
    print(f"  rebuilds {st['rebuilds']:.0f}, largest displacement {st['max_displacement']:.3f} A", flush=True)
    assert cen == co
    check_forces("water32k after 100 steps", F, fo)
    check_energies(E[0], eo, (0, 1, 2, 3), cen[0])


# --------------------------------------------------------------- ensemble64

def test_ensemble64_batch_vs_alone_bitwise_and_oracle():
    systems = W.ensemble64()
    sizes = [s.n for s in systems]
    print(f"case ensemble64 ({len(systems)} boxes, {sum(sizes)} atoms, {min(sizes)}-{max(sizes)} atoms each)",
          flush=True)
    desc = W.desc_water32k(flags=L.FLAG_VIRIAL)
    v0s = [S.maxwell_velocities(s, 300.0, seed=900 + k) for k, s in enumerate(systems)]
    steps = 20
    with L.Context(desc, systems) as ctx:
        x = np.concatenate([s.x for s in systems])
        fb = np.zeros_like(x)
        eb = np.zeros((len(systems), L.NSLOTS))
        ctx.evaluate_batch(x, fb, eb)
        ctx.md_set_state(x, np.concatenate(v0s))
        ctx.md_run(steps, 1.0, energies=True)
        xm, vm, Fm, Em = ctx.md_get_state()
    off = 0
    for k, s in enumerate(systems):
        with L.Context(desc, [s]) as ctx:
            f1, e1 = evaluate(ctx, 0, s.x)
            ctx.md_set_state(s.x, v0s[k])
            ctx.md_run(steps, 1.0, energies=True)
            x1, v1, F1, E1 = ctx.md_get_state()
        sl = slice(off, off + s.n)
        assert np.array_equal(fb[sl], f1) and np.array_equal(eb[k], e1), k
        assert np.array_equal(xm[sl], x1) and np.array_equal(vm[sl], v1), k
        assert np.array_equal(Fm[sl], F1) and np.array_equal(Em[k], E1[0]), k
        off += s.n
    offs = np.cumsum([0] + sizes)
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        fo, eo, _ = oracle_water32k(systems[k])
        check_forces(f"ensemble box {k} ({sizes[k]} atoms)", fb[offs[k]:offs[k + 1]], fo)
        check_energies(eb[k], eo, (0, 1, 2, 3), 300 * sizes[k])


# ------------------------------------------------------------------- served

def test_served_requests_alone_vs_mix_bitwise():
    """Evaluation and 200-step MD requests on boxes across the served range
    (1,000-32,000 molecules) served together: each result is bitwise the
    result of the same request alone."""
    sizes = [1000, 2100, 4400, 9000, 19000, 32000]
    systems = W.waters_at_density(sizes, 7100)
    print(f"case served ({len(systems)} boxes, {sizes[0]}-{sizes[-1]} molecules)", flush=True)
    desc = W.desc_water32k(rebuild=0)
    reqs = []
    for k, s in enumerate(systems):
        for kind in ("eval", "md"):
            rid = len(reqs)
            v = S.maxwell_velocities(s, 300.0, seed=20 + rid) if kind == "md" else None
            reqs.append(serve.Request(rid=rid, kind=kind, system=s, x=s.x + 0.37 * rid, v=v,
                                      steps=200 if kind == "md" else 0, arrival=0.002 * rid))
    with L.Context(desc, [], reserve_bytes=24 << 30) as ctx:
        done = serve.Server(ctx, dt_fs=1.0, chunk=20).run(reqs, max_seconds=600)
    assert len(done) == len(reqs)
    for r in sorted(done, key=lambda q: q.rid):
        with L.Context(desc, [r.system]) as ctx:
            if r.kind == "eval":
                f, e = evaluate(ctx, 0, r.x)
                assert np.array_equal(r.f, f) and np.array_equal(r.e, e), r.rid
            else:
                ctx.md_set_state(r.x, r.v)
                ctx.md_run(r.steps, 1.0, energies=True)
                x, v, F, E = ctx.md_get_state()
                assert np.array_equal(r.x_out, x) and np.array_equal(r.v_out, v), r.rid
                assert np.array_equal(r.f, F) and np.array_equal(r.e, E[0]), r.rid


# ----------------------------------------------------------------- chains32

def test_chains32_batch_vs_alone_bitwise_and_oracle():
    chains = W.chains32()
    print(f"case chains32 ({len(chains)} GB chains x {chains[0].n} atoms)", flush=True)
    desc = W.desc_chains()
    v0s = [S.maxwell_velocities(c, 300.0, seed=4000 + k) for k, c in enumerate(chains)]
    steps = 5
    with L.Context(desc, chains) as ctx:
        x = np.concatenate([c.x for c in chains])
        fb = np.zeros_like(x)
        eb = np.zeros((len(chains), L.NSLOTS))
        ctx.evaluate_batch(x, fb, eb)
        ctx.md_set_state(x, np.concatenate(v0s))
        ctx.md_run(steps, 1.0, energies=True)
        xm, vm, Fm, Em = ctx.md_get_state()
    n = chains[0].n
    for k, c in enumerate(chains):
        with L.Context(desc, [c]) as ctx:
            f1, e1 = evaluate(ctx, 0, c.x)
            ctx.md_set_state(c.x, v0s[k])
            ctx.md_run(steps, 1.0, energies=True)
            x1, v1, F1, E1 = ctx.md_get_state()
        sl = slice(k * n, (k + 1) * n)
        assert np.array_equal(fb[sl], f1) and np.array_equal(eb[k], e1), k
        assert np.array_equal(xm[sl], x1) and np.array_equal(vm[sl], v1) and np.array_equal(Fm[sl], F1), k
        assert np.array_equal(Em[k], E1[0]), k
    npairs = 3 * n * (n - 1) // 2
    for k in (0, len(chains) - 1):
        fo, eo, _ = O.gb_evaluate(chains[k])
        fl, el = O.vacuum_lj_coulomb(chains[k])
        fbd, ebd = O.bonded(chains[k])
        check_forces(f"chain {k}", fb[k * n:(k + 1) * n], fo + fl + fbd)
        assert abs(eb[k][0] - (eo[0] + el[0])) <= 1e-5 * abs(eo[0] + el[0]) + 1e-7 * npairs
        assert abs(eb[k][1] - el[1]) <= 1e-5 * max(1.0, abs(el[1])) + 1e-7 * npairs
        assert abs(eb[k][2] - ebd[2]) <= 1e-9 * max(1.0, abs(ebd[2])) and abs(eb[k][3] - ebd[3]) <= 1e-6


def test_chains_md_200_steps_forces_vs_oracle():
    """The forces the step loop holds after 200 steps at 1 fs (the benchmarked
    run) against FP64 at those positions: the chain stays a chain (no
    overlapping beads), where FP32 pair arithmetic is accurate."""
    print("case chains md state (200 steps)", flush=True)
    c = S.gb_chain(5000, 3000)
    v0 = S.maxwell_velocities(c, 300.0, seed=4000)
    with L.Context(W.desc_chains(), [c]) as ctx:
        ctx.md_set_state(c.x, v0)
        ctx.md_run(200, 1.0, energies=True)
        x, v, F, E = ctx.md_get_state()
    fo, _, _ = O.gb_evaluate(c, x=x)
    fl, _ = O.vacuum_lj_coulomb(c, x=x)
    fbd, _ = O.bonded(c, x=x)
    d = np.linalg.norm(x[1:] - x[:-1], axis=1)
    print(f"  bond lengths after 200 steps: {d.min():.3f}-{d.max():.3f} A", flush=True)
    check_forces("chain after 200 steps", F, fo + fl + fbd)
