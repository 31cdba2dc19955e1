"""The five benchmark workloads of the nonbonded force path: their seeded
synthetic systems and force descriptions, shared by bench.py and the
full-size parity tests so both run the same inputs.

  water1k     1,000 waters (3,000 atoms) in a 31.04 A box, cutoff LJ and
              Coulomb, both C2-switched 8 -> 10 A (OneAtomPairTwo<PBC, C2, LJ,
              C2, Coulomb>), bond + angle water model; 100 leapfrog steps at
              0.5 fs from ProtoMol's 300 K velocity start
  water32k    32,000 waters (96,000 atoms) in a 98.6 A box, LJ C2 8 -> 10 A +
              Ewald real space (alpha 0.32 /A, Rc 10 A) with the excluded-pair
              correction, bond + angle; leapfrog at 1 fs, pair list rebuilt
              every 10 steps
  ensemble64  64 water boxes of 500-16,000 molecules (log-uniform by seed) at
              0.0334 molecules / A^3, the water32k force model
  served      water boxes of 1,000-32,000 molecules (log-uniform), the
              water32k force model, as a request stream's systems
  chains32    32 GB/OBC bead chains of 5,000 atoms, all pairs (Born radii,
              partial sums, GB + LJ + Coulomb) plus the chains' bonds and
              angles

Water boxes are lattice starts with seeded orientations whose H-H contacts
under 2 A are relieved (systems.water_box, min_contact=2.0).
"""
from __future__ import annotations

import math
from multiprocessing import get_context

from . import _lib as L
from . import systems as S

DENSITY = 0.0334                 # molecules / A^3
EWALD_ALPHA = 0.32
EWALD_ACCURACY = math.exp(-10.24)  # Rc = sqrt(10.24) / 0.32 = 10 A
WORKERS = 8                      # generator processes (CPU only)


def water1k():
    return S.water_box(1000, 31.04, seed=1000, min_contact=2.0)


def water32k():
    return S.water_box(32000, 98.6, seed=2024, min_contact=2.0)


def _water_at_density(args):
    n, seed = args
    return S.water_box(n, (n / DENSITY) ** (1.0 / 3.0), seed=seed, min_contact=2.0)


def waters_at_density(sizes, seed0, workers=WORKERS):
    """Water boxes of the given molecule counts at 0.0334 molecules/A^3,
    seeds seed0, seed0 + 1, ...; generated in worker processes (spawned, so
    nothing of a GPU context is inherited)."""
    jobs = [(int(n), seed0 + k) for k, n in enumerate(sizes)]
    if workers <= 1 or len(jobs) == 1:
        return [_water_at_density(j) for j in jobs]
    with get_context("spawn").Pool(min(workers, len(jobs))) as pool:
        return pool.map(_water_at_density, jobs)


def ensemble64_sizes(count=64):
    return S.ensemble_sizes(count, 500, 16000, seed=77)


def ensemble64(count=64, workers=WORKERS):
    return waters_at_density(ensemble64_sizes(count), 500, workers)


def served_sizes(count=16):
    return S.ensemble_sizes(count, 1000, 32000, seed=91)


def served_templates(count=16, workers=WORKERS):
    return waters_at_density(served_sizes(count), 700, workers)


def _chain(args):
    n, seed = args
    return S.gb_chain(n, seed)


def chains32(count=32, n_atoms=5000, workers=WORKERS):
    jobs = [(n_atoms, 3000 + k) for k in range(count)]
    if workers <= 1:
        return [_chain(j) for j in jobs]
    with get_context("spawn").Pool(min(workers, len(jobs))) as pool:
        return pool.map(_chain, jobs)


def desc_water1k(skin=1.2, rebuild=10, flags=0) -> L.ForceDesc:
    """OneAtomPairTwo<PBC, C2, LJ, C2, Coulomb> 8 -> 10 A + bond + angle."""
    d = L.ForceDesc()
    d.term_a = L.TERM_LJ
    d.switch_a = L.switch(L.SWITCH_C2, 8.0, 10.0)
    d.term_b = L.TERM_COULOMB
    d.switch_b = L.switch(L.SWITCH_C2, 8.0, 10.0)
    d.cutoff = 10.0
    d.skin = skin
    d.rebuild_interval = rebuild
    d.flags = L.FLAG_BONDED | flags
    return d


def desc_water32k(skin=1.2, rebuild=10, flags=0, **shape) -> L.ForceDesc:
    """LJ C2 8 -> 10 A + Ewald real space (alpha 0.32, Rc 10 A, correction) + bond + angle."""
    d = L.ForceDesc()
    d.term_a = L.TERM_LJ
    d.switch_a = L.switch(L.SWITCH_C2, 8.0, 10.0)
    d.term_b = L.TERM_EWALD_REAL
    d.cutoff = 10.0
    d.ewald_alpha = EWALD_ALPHA
    d.ewald_accuracy = EWALD_ACCURACY
    d.ewald_correction = 1
    d.skin = skin
    d.rebuild_interval = rebuild
    d.flags = L.FLAG_BONDED | flags
    for k, v in shape.items():
        if v:
            setattr(d, k, v)
    return d


def desc_chains(flags=0) -> L.ForceDesc:
    """GB/OBC (offset 0.09, OBC2 alpha 1, beta 0.8, gamma 4.85, dielectrics
    1 / 78.3) + LJ + Coulomb, all pairs, and the chains' harmonic bonds and
    angles."""
    d = L.ForceDesc()
    d.gb_offset, d.gb_alpha, d.gb_beta, d.gb_gamma = 0.09, 1.0, 0.8, 4.85
    d.gb_solute_dielec, d.gb_solvent_dielec = 1.0, 78.3
    d.flags = L.FLAG_GB | L.FLAG_BONDED | flags
    return d
