"""Shared helpers for the tests: force models, fixture loading, error metrics."""
from __future__ import annotations

import math
import os

import numpy as np

from protomol_amd import systems as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

EWALD_ALPHA = 0.32
EWALD_ACCURACY = math.exp(-10.24)   # Rc = sqrt(10.24) / 0.32 = 10 A

# switch tuples: (kind, switchon, cutoff, switchoff, order) in PM_SWITCH_* numbering
C2_8_10 = (3, 8.0, 10.0, 0.0, 0.0)

# golden-fixture driver numbering (1 C1, 2 C2, 3 Cn, 4 CmpCnCn, 5 Cutoff) -> PM_SWITCH_*
DRIVER_TO_PM = {1: 2, 2: 3, 3: 4, 4: 5, 5: 1}
DRIVER_TO_ORC = {1: 2, 2: 3, 3: 4, 4: 5, 5: 1}


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def fixture_water():
    """The seeded system the ref_cutoff_* / ref_ewald / ref_bonded fixtures were written for."""
    return S.water_box(343, 22.5, seed=5, jitter=0.3, atom_noise=0.03)


def rms_rel(f, ref):
    return float(np.sqrt(np.sum((f - ref) ** 2) / np.sum(ref ** 2)))


def max_rel_atom(f, ref, floor=1.0):
    """max over atoms of |dF| / max(|F_ref|, floor)."""
    d = np.linalg.norm(f - ref, axis=1)
    r = np.maximum(np.linalg.norm(ref, axis=1), floor)
    return float(np.max(d / r))


def water423_system():
    d = golden("water423.npz")
    ex = np.concatenate([
        np.concatenate([d["bonds"], np.full((len(d["bonds"]), 1), S.EXCL_FULL, np.int32)], 1),
        np.concatenate([d["angles"][:, [0, 2]], np.full((len(d["angles"]), 1), S.EXCL_FULL, np.int32)], 1),
    ]).astype(np.int32)
    sysw = S.System(x=d["x0"].copy(), charge=d["charge"] * S.SQRT_COULOMB, mass=d["mass"], molecule=d["molecule"],
                    box=d["box"], origin=d["origin"], type=d["type"], lj_table=d["lj_table"], exclusions=ex,
                    bonds=d["bonds"], bond_k=d["bond_k"], bond_r0=d["bond_r0"], angles=d["angles"],
                    angle_k=d["angle_k"], angle_t0=d["angle_t0"], name="water423")
    return sysw, d


def leapfrog_oracle(force_fn, x, v, mass, dt_fs, n_steps):
    """LeapfrogIntegrator::run (LeapfrogIntegrator.cpp:89-104) in numpy FP64,
    with the reference's expression order for each kick."""
    h = dt_fs * S.INV_TIMEFACTOR
    m = mass[:, None]
    F, E = force_fn(x)
    for i in range(n_steps):
        if i == 0:
            v = v + F * h * 0.5 / m
        else:
            v = v + F * h / m
        x = x + v * h
        F, E = force_fn(x)
    v = v + F * (0.5 * dt_fs * S.INV_TIMEFACTOR) / m
    return x, v, F, E


# ------------------------------------------------ Langevin noise and oracle

_PH_M0, _PH_M1, _PH_W0, _PH_W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox-4x32-10 (Salmon et al., SC 2011) on uint64 numpy arrays holding
    32-bit words (vectorised over the counter)."""
    c = [np.asarray(v, np.uint64) & 0xFFFFFFFF for v in (c0, c1, c2, c3)]
    k = [np.uint64(k0), np.uint64(k1)]
    m32 = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(_PH_M0) * c[0]
        p1 = np.uint64(_PH_M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k[0], p1 & m32, (p0 >> np.uint64(32)) ^ c[3] ^ k[1], p0 & m32]
        k = [(k[0] + np.uint64(_PH_W0)) & m32, (k[1] + np.uint64(_PH_W1)) & m32]
    return c


def langevin_gauss(seed, n, ctr):
    """The three N(0, 1) deviates per atom of one Langevin half kick: Box-Muller
    on Philox-4x32-10 of (atom, ctr, 'LANG', 0) under key (seed, 'PROT')."""
    a = np.arange(n, dtype=np.uint64)
    r = philox4x32_10(a, np.full(n, ctr, np.uint64), np.full(n, 0x4C414E47, np.uint64), np.zeros(n, np.uint64),
                      seed, 0x50524F54)
    u = [(x.astype(np.float64) + 0.5) * 2.0 ** -32 for x in r]
    m0, m1 = np.sqrt(-2.0 * np.log(u[0])), np.sqrt(-2.0 * np.log(u[2]))
    return np.stack([m0 * np.cos(2 * np.pi * u[1]), m0 * np.sin(2 * np.pi * u[1]), m1 * np.cos(2 * np.pi * u[3])], 1)


def langevin_oracle(force_fn, x, v, mass, dt_fs, n_steps, temperature, gamma_ps, seed, ctr0=1):
    """LangevinLeapfrogIntegrator (STSIntegrator::run: half kick, drift,
    forces, half kick; doHalfKick LangevinLeapfrogIntegrator.cpp:46-70) in
    numpy FP64, with the device's noise stream; ctr0 = the system's force
    evaluations before the first kick."""
    h = dt_fs * S.INV_TIMEFACTOR
    g = gamma_ps / (1000 * S.INV_TIMEFACTOR)
    fdt = (1.0 - math.exp(-0.5 * g * h)) / g
    vdt = math.exp(-0.5 * g * h)
    ndt = math.sqrt((1.0 - math.exp(-g * h)) / (2.0 * g))
    fc = 2 * S.BOLTZMANN * temperature * g
    m = mass[:, None]
    n = len(mass)

    def half(v, F, k):
        return v * vdt + F * fdt / m + langevin_gauss(seed, n, k) * np.sqrt(fc / m) * ndt

    F, E = force_fn(x)
    c = ctr0
    for i in range(n_steps):
        if i > 0:
            v = half(v, F, 2 * c)
        v = half(v, F, 2 * c + 1)
        x = x + v * h
        F, E = force_fn(x)
        c += 1
    v = half(v, F, 2 * c)
    return x, v, F, E


def vacuum_lj_coulomb(c, virial=False, molvirial=False, x=None):
    """All-pairs LJ (per-atom GROMACS mixing) + Coulomb in vacuum, as
    OneAtomPair<Vacuum, Universal, ...> adds them (OneAtomPair.h:47-107):
    exclusions only between atoms of one molecule; with the molecular virial
    on, pairs of two molecules add f (x) (R_J - R_I) to MOLVIRIAL and f (x) d
    to VIRIAL, other pairs add f (x) d only when the atomic virial is on.
    Returns forces (n, 3) and 36 slots."""
    import oracle as O
    x = c.x if x is None else x
    n = c.n
    mol = np.asarray(c.molecule)
    d = x[None, :, :] - x[:, None, :]
    r2 = np.sum(d * d, axis=2)
    iu = np.triu_indices(n, 1)
    excl = {}
    for a, b, cls in c.exclusions:
        if mol[a] == mol[b]:
            excl[(min(a, b), max(a, b))] = cls
    cls = np.array([excl.get((i, j), 0) for i, j in zip(*iu)])
    keep = cls != 1
    i, j, cls = iu[0][keep], iu[1][keep], cls[keep]
    rr = 1.0 / r2[i, j]
    sg = 0.5 * (c.lj_sigma[i] + c.lj_sigma[j])
    ep = np.sqrt(c.lj_epsilon[i] * c.lj_epsilon[j])
    A, B = 4 * ep * sg ** 12, 4 * ep * sg ** 6
    A = np.where(cls == 2, A * c.lj_scaling, A)
    B = np.where(cls == 2, B * c.lj_scaling, B)
    r6 = rr ** 3
    elj = A * r6 * r6 - B * r6
    flj = (12 * A * r6 * r6 - 6 * B * r6) * rr
    ec = c.charge[i] * c.charge[j] * np.sqrt(rr)
    ec = np.where(cls == 2, ec * c.coulomb_scaling, ec)
    fc = ec * rr
    fij = d[i, j] * (flj + fc)[:, None]
    f = np.zeros((n, 3))
    np.add.at(f, i, -fij)
    np.add.at(f, j, fij)
    e = np.zeros(36)
    e[0] = ec.sum()
    e[1] = elj.sum()
    other = mol[i] != mol[j]
    if molvirial:
        com = O.molecule_com(c, x=x)
        cd = com[mol[j]] - com[mol[i]]
        w = np.einsum("pa,pb->ab", fij[other], cd[other])
        # ScalarStructure::addVirial fills the lower triangle with the upper one
        w[1, 0], w[2, 0], w[2, 1] = w[0, 1], w[0, 2], w[1, 2]
        e[16:25] += w.reshape(-1)
    atomic = (other & molvirial) | (np.ones_like(other) if virial else np.zeros_like(other))
    e[7:16] += np.einsum("pa,pb->ab", fij[atomic], d[i, j][atomic]).reshape(-1)
    return f, e

