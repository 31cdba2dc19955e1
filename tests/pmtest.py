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

# golden-fixture driver numbering (1 C1, 2 C2, 3 Cn, 4 CmpCnCn) -> PM_SWITCH_*
DRIVER_TO_PM = {1: 2, 2: 3, 3: 4, 4: 5}
DRIVER_TO_ORC = {1: 2, 2: 3, 3: 4, 4: 5}


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
