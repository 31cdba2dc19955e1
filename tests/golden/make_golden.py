"""Write the golden fixtures under tests/golden/.

Run in a container that holds the ProtoMol sources (never on the GPU box):

    make -C oracle lib ref && python tests/golden/make_golden.py

Two kinds of fixtures:
  ref_*.npz        inputs + outputs of the reference's own force classes
                   (oracle/_ref/ref_driver, compiled from the reference
                   sources) on seeded synthetic systems: forces, the 36
                   ScalarStructure slots, the evaluated pair-set census and,
                   for GB, the Born radii.
  ref_velocity.npz the reference's velocity start (random number streams of
                   both generator types, randomVelocity, kineticEnergy)
  water423.npz     the data files of the reference's regression test
                   test/tests/water_CHARMM_PERIODIC_L.conf: PSF/PAR/PDB
                   inputs and the committed expected outputs (forces at step 0
                   and 200, final positions/velocities, energies), converted
                   to arrays.
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refio  # noqa: E402

from protomol_amd import systems as S  # noqa: E402

REFTESTS = "/root/reference/test/tests"

# switch vectors for the driver: kind (1 C1, 2 C2, 3 Cn, 4 CmpCnCn), switchon, cutoff, switchoff, order
CUTOFF_CASES = {
    "c2c2": (1, 2, (2, 8.0, 10.0, 0.0, 0.0), (2, 8.0, 10.0, 0.0, 0.0)),
    "c2c1": (1, 2, (2, 0.1, 10.0, 0.0, 0.0), (1, 0.0, 10.0, 0.0, 0.0)),
    "cncn": (1, 2, (3, 8.0, 10.0, 0.0, 6.0), (3, 8.0, 10.0, 0.0, 6.0)),
    "cmpcncn": (1, 2, (4, 7.0, 10.0, 9.0, 2.0), (4, 7.0, 10.0, 9.0, 2.0)),
    "lj_c2": (1, 0, (2, 8.0, 10.0, 0.0, 0.0), None),
    "coul_c1": (2, 0, (1, 0.0, 10.0, 0.0, 0.0), None),
    # OneAtomPairTwo with two cutoffs: LJ switched off at 8 A, Coulomb at 10 A
    "c2c2_two": (1, 2, (2, 6.0, 8.0, 0.0, 0.0), (2, 8.0, 10.0, 0.0, 0.0)),
    # two cutoffs through the run-time switch path: LJ C2 ending at 8 A, Coulomb C1 at 10 A
    "c2c1_two": (1, 2, (2, 6.0, 8.0, 0.0, 0.0), (1, 0.0, 10.0, 0.0, 0.0)),
    # CutoffSwitchingFunction (does not modify values): alone, both terms with
    # two cutoffs (no switch modifies: both out to 10 A), and beside a C2
    # switch (both switches applied: the Cutoff term stops at its own cutoff)
    "coul_cut": (2, 0, (5, 0.0, 10.0, 0.0, 0.0), None),
    "cut_cut_two": (1, 2, (5, 0.0, 8.0, 0.0, 0.0), (5, 0.0, 10.0, 0.0, 0.0)),
    "cut_c2_two": (1, 2, (5, 0.0, 8.0, 0.0, 0.0), (2, 8.0, 10.0, 0.0, 0.0)),
    "c2_cut_two": (1, 2, (2, 6.0, 8.0, 0.0, 0.0), (5, 0.0, 10.0, 0.0, 0.0)),
}


def water_fixture_system():
    return S.water_box(343, 22.5, seed=5, jitter=0.3, atom_noise=0.03)


def make_cutoff(name, sysobj, case):
    ta, tb, sa, sb = case
    r = refio.system_records(sysobj)
    r.update(mode=np.int32(1), term_a=np.int32(ta), term_b=np.int32(tb), sw_a=np.array(sa, float),
             alg_cutoff=10.0, cellsize=5.0, virial=np.int32(1))
    if sb is not None:
        r["sw_b"] = np.array(sb, float)
    out = refio.run(r)
    np.savez_compressed(os.path.join(HERE, f"ref_cutoff_{name}.npz"), x=sysobj.x, f=out["f"].reshape(-1, 3),
                        e=out["e"], census=np.array(out["census"], dtype=np.uint64),
                        term_a=ta, term_b=tb, sw_a=np.array(sa, float),
                        sw_b=np.array(sb if sb is not None else sa, float))


# vacuum cutoff cases (NonbondedCutoffForceModule.cpp:110-140): name -> (term_a, term_b, sw_a, sw_b)
VACUUM_CASES = {
    # OneAtomPairTwo: LJ switched off at 8 A, Coulomb at 10 A
    "c2c2_two": (1, 2, (2, 6.0, 8.0, 0.0, 0.0), (2, 8.0, 10.0, 0.0, 0.0)),
    "lj_c2": (1, 0, (2, 6.0, 8.0, 0.0, 0.0), None),
    "coul_c1": (2, 0, (1, 0.0, 10.0, 0.0, 0.0), None),
}


def droplet_fixture_system():
    return S.water_droplet(14.0, seed=3)


def make_vacuum(name, sysobj, case):
    """NonbondedCutoffSystemForce<CubicCellManager, OneAtomPair[Two]<VacuumBoundaryConditions, ...>>,
    with the atomic and the molecular virial."""
    ta, tb, sa, sb = case
    r = refio.system_records(sysobj)
    r.update(mode=np.int32(6), term_a=np.int32(ta), term_b=np.int32(tb), sw_a=np.array(sa, float),
             alg_cutoff=max(sa[2], sb[2] if sb is not None else 0.0), cellsize=5.0, virial=np.int32(1),
             molvirial=np.int32(1))
    if sb is not None:
        r["sw_b"] = np.array(sb, float)
    out = refio.run(r)
    np.savez_compressed(os.path.join(HERE, f"ref_vacuum_{name}.npz"), x=sysobj.x, f=out["f"].reshape(-1, 3),
                        e=out["e"], census=np.array(out["census"], dtype=np.uint64),
                        term_a=ta, term_b=tb, sw_a=np.array(sa, float),
                        sw_b=np.array(sb if sb is not None else sa, float))


def make_ewald(sysobj):
    acc = math.exp(-10.24)
    r = refio.system_records(sysobj)
    r.update(mode=np.int32(2), ew_alpha=0.32, ew_accuracy=acc, ew_correction=np.int32(1), cellsize=5.0,
             virial=np.int32(1))
    out = refio.run(r)
    np.savez_compressed(os.path.join(HERE, "ref_ewald.npz"), x=sysobj.x, f=out["f"].reshape(-1, 3), e=out["e"],
                        census=np.array(out["census"], dtype=np.uint64), alpha=0.32, accuracy=acc)


def make_ewald_reciprocal(sysobj):
    """NonbondedFullEwaldSystemForce with the reciprocal term: the full sum
    (real + reciprocal + correction, ProtoMol's "CoulombEwald") and the
    reciprocal term alone, both with the virial."""
    acc = math.exp(-10.24)
    out = {}
    for parts, tag in ((7, "full"), (2, "recip")):
        r = refio.system_records(sysobj)
        r.update(mode=np.int32(2), ew_alpha=0.32, ew_accuracy=acc, ew_parts=np.int32(parts), cellsize=5.0,
                 virial=np.int32(1))
        o = refio.run(r)
        out[f"f_{tag}"] = o["f"].reshape(-1, 3)
        out[f"e_{tag}"] = o["e"]
    np.savez_compressed(os.path.join(HERE, "ref_ewald_recip.npz"), x=sysobj.x, alpha=0.32, accuracy=acc, **out)


def make_bonded(sysobj):
    r = refio.system_records(sysobj)
    r.update(mode=np.int32(4), bonds=sysobj.bonds.reshape(-1), bond_k=sysobj.bond_k, bond_r0=sysobj.bond_r0,
             angles=sysobj.angles.reshape(-1), angle_k=sysobj.angle_k, angle_t0=sysobj.angle_t0, cellsize=5.0)
    out = refio.run(r)
    np.savez_compressed(os.path.join(HERE, "ref_bonded.npz"), x=sysobj.x, f=out["f"].reshape(-1, 3), e=out["e"])


def make_gb(n, seed):
    g = S.gb_chain(n, seed=seed)
    params = np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3])
    r = refio.system_records(g)
    r.update(mode=np.int32(3), gb_rho=g.gb_radius, gb_S=g.gb_scaling, gb_params=params, cellsize=np.int32(5),
             gb_ljcoul=np.int32(1))
    out = refio.run(r)
    np.savez_compressed(os.path.join(HERE, f"ref_gb{n}.npz"), x=g.x, f=out["f"].reshape(-1, 3), e=out["e"],
                        born=out["born"], params=params, seed=seed)


def make_gb_ace(n, seed):
    """GB + GBACEForce (ACE non-polar solvation) on a chain, without the
    LJ/Coulomb terms: GBBornRadii, GBPartialSum, GBForce, GBACEForce."""
    g = S.gb_chain(n, seed=seed)
    params = np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3])
    r = refio.system_records(g)
    r.update(mode=np.int32(3), gb_rho=g.gb_radius, gb_S=g.gb_scaling, gb_params=params, cellsize=np.int32(5),
             gb_ljcoul=np.int32(0), gb_ace=np.int32(1), ace_sigma=2.26 / 418.4, ace_rho_s=1.4)
    out = refio.run(r)
    r["gb_ace"] = np.int32(0)
    base = refio.run(r)
    np.savez_compressed(os.path.join(HERE, f"ref_gb_ace{n}.npz"), x=g.x, f=out["f"].reshape(-1, 3), e=out["e"],
                        f_noace=base["f"].reshape(-1, 3), e_noace=base["e"], born=out["born"], params=params,
                        seed=seed, ace_sigma=2.26 / 418.4, ace_rho_s=1.4)


def gb_multi_chain(n, seed, per):
    """A GB chain cut into molecules of `per` atoms (the pair terms see
    pairs of different molecules, so the molecular virial has work)."""
    g = S.gb_chain(n, seed=seed)
    g.molecule = (np.arange(n) // per).astype(np.int32)
    return g


def chain_fixture_system(per_atom=False):
    return S.chain_box(512, 52.0, seed=3, per_atom=per_atom)


def make_modified():
    """Systems with MODIFIED (1-4) exclusions: united-atom chains, once with
    the type table (its own A14 / B14) and once with per-atom LJ and
    lj_scaling, both with coulomb_scaling 1/1.2.  LJ + Coulomb C2 cutoff pairs
    (forces, energies, virial, pair census) and the Ewald real-space sum with
    its correction term."""
    acc = math.exp(-10.24)
    ta, tb, sa, sb = CUTOFF_CASES["c2c2"]
    out = {"alpha": 0.32, "accuracy": acc}
    for tag, per_atom in (("table", False), ("atom", True)):
        c = chain_fixture_system(per_atom)
        out[f"x_{tag}"] = c.x
        r = refio.system_records(c)
        r.update(mode=np.int32(1), term_a=np.int32(ta), term_b=np.int32(tb), sw_a=np.array(sa, float),
                 sw_b=np.array(sb, float), alg_cutoff=10.0, cellsize=5.0, virial=np.int32(1))
        o = refio.run(r)
        out[f"f_cut_{tag}"], out[f"e_cut_{tag}"] = o["f"].reshape(-1, 3), o["e"]
        out[f"census_{tag}"] = np.array(o["census"], dtype=np.uint64)
        r = refio.system_records(c)
        r.update(mode=np.int32(2), ew_alpha=0.32, ew_accuracy=acc, ew_correction=np.int32(1), cellsize=5.0,
                 virial=np.int32(1))
        o = refio.run(r)
        out[f"f_ew_{tag}"], out[f"e_ew_{tag}"] = o["f"].reshape(-1, 3), o["e"]
    np.savez_compressed(os.path.join(HERE, "ref_modified.npz"), **out)


def gb_chain_modified(n=200, seed=7):
    """A GB chain whose 1-4 pairs are MODIFIED, with lj_scaling 0.5 and
    coulomb_scaling 1/1.2 (the scaled LJ and Coulomb of the GB systems' pair
    terms)."""
    g = S.gb_chain(n, seed=seed)
    ex14 = np.stack([np.arange(n - 3), np.arange(3, n), np.full(n - 3, S.EXCL_MODIFIED)], 1)
    g.exclusions = np.concatenate([g.exclusions, ex14]).astype(np.int32)
    g.lj_scaling, g.coulomb_scaling = 0.5, 1.0 / 1.2
    return g


def gb_chain_table(n=200, seed=7):
    """The MODIFIED-pair GB chain with a 4-type LJ table (types i mod 4, own
    A14 / B14) in place of per-atom LJ."""
    g = gb_chain_modified(n, seed)
    sig, eps = np.array([3.0, 3.4, 3.8, 2.6]), np.array([0.08, 0.12, 0.18, 0.05])
    tab = np.zeros((4, 4, 4))
    for a in range(4):
        for b in range(4):
            sg, ep = 0.5 * (sig[a] + sig[b]), math.sqrt(eps[a] * eps[b])
            tab[a, b] = (4 * ep * sg ** 12, 4 * ep * sg ** 6, 2 * ep * (0.9 * sg) ** 12, 2 * ep * (0.9 * sg) ** 6)
    g.type = (np.arange(n) % 4).astype(np.int32)
    g.lj_table, g.lj_sigma, g.lj_epsilon = tab, None, None
    return g


def make_gb_modified(n=200, seed=7):
    g = gb_chain_modified(n, seed)
    params = np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3])
    r = refio.system_records(g)
    r.update(mode=np.int32(3), gb_rho=g.gb_radius, gb_S=g.gb_scaling, gb_params=params, cellsize=np.int32(5),
             gb_ljcoul=np.int32(1))
    out = refio.run(r)
    res = {"x": g.x, "f": out["f"].reshape(-1, 3), "e": out["e"], "params": params, "n": n, "seed": seed}
    t = gb_chain_table(n, seed)
    r = refio.system_records(t)
    r.update(mode=np.int32(3), gb_rho=t.gb_radius, gb_S=t.gb_scaling, gb_params=params, cellsize=np.int32(5),
             gb_ljcoul=np.int32(1))
    out = refio.run(r)
    res.update(f_table=out["f"].reshape(-1, 3), e_table=out["e"])
    np.savez_compressed(os.path.join(HERE, "ref_gb_modified.npz"), **res)


def make_molvirial(w):
    """The atomic and molecular virials (ScalarStructure::molecularVirial on)
    of the reference's classes: LJ + Coulomb C2 cutoff pairs and the full
    Ewald sum on the water box, bonded terms, and GB + LJ + Coulomb on a chain
    cut into molecules, each with the atomic virial on and off."""
    acc = math.exp(-10.24)
    out = {"x": w.x, "alpha": 0.32, "accuracy": acc}
    ta, tb, sa, sb = CUTOFF_CASES["c2c2"]
    for vir in (1, 0):
        r = refio.system_records(w)
        r.update(mode=np.int32(1), term_a=np.int32(ta), term_b=np.int32(tb), sw_a=np.array(sa, float),
                 sw_b=np.array(sb, float), alg_cutoff=10.0, cellsize=5.0, virial=np.int32(vir),
                 molvirial=np.int32(1))
        o = refio.run(r)
        out[f"f_cut_v{vir}"], out[f"e_cut_v{vir}"] = o["f"].reshape(-1, 3), o["e"]
        for parts, tag in ((7, "ew"), (2, "rec")):
            r = refio.system_records(w)
            r.update(mode=np.int32(2), ew_alpha=0.32, ew_accuracy=acc, ew_parts=np.int32(parts), cellsize=5.0,
                     virial=np.int32(vir), molvirial=np.int32(1))
            o = refio.run(r)
            out[f"f_{tag}_v{vir}"], out[f"e_{tag}_v{vir}"] = o["f"].reshape(-1, 3), o["e"]
    r = refio.system_records(w)
    r.update(mode=np.int32(4), bonds=w.bonds.reshape(-1), bond_k=w.bond_k, bond_r0=w.bond_r0,
             angles=w.angles.reshape(-1), angle_k=w.angle_k, angle_t0=w.angle_t0, cellsize=5.0,
             virial=np.int32(1))
    o = refio.run(r)
    out["e_bonded_v1"] = o["e"]
    g = gb_multi_chain(200, 17, 50)
    params = np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3])
    for vir in (1, 0):
        r = refio.system_records(g)
        r.update(mode=np.int32(3), gb_rho=g.gb_radius, gb_S=g.gb_scaling, gb_params=params, cellsize=np.int32(5),
                 gb_ljcoul=np.int32(1), virial=np.int32(vir), molvirial=np.int32(1))
        o = refio.run(r)
        out[f"f_gb_v{vir}"], out[f"e_gb_v{vir}"] = o["f"].reshape(-1, 3), o["e"]
    out.update(gb_x=g.x, gb_seed=17, gb_per=50, gb_params=params)
    np.savez_compressed(os.path.join(HERE, "ref_molvirial.npz"), **out)


# ------------------------------------------------------------- water_423

def _pdb_xyz(path):
    out = []
    for line in open(path):
        if line.startswith("ATOM") or line.startswith("HETATM"):
            out.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    return np.array(out)


def _xyz_frames(path):
    # ProtoMol XYZ output: an optional frame-count line, then per frame an
    # atom-count line and "name x y z" rows (io/XYZWriter.cpp:41-57)
    lines = [l for l in open(path).read().split("\n") if l.strip()]
    frames, k = [], 0
    while k < len(lines):
        if len(lines[k].split()) == 1 and k + 1 < len(lines) and len(lines[k + 1].split()) == 1:
            k += 1
            continue
        n = int(lines[k].split()[0])
        rows = [[float(v) for v in lines[j].split()[1:4]] for j in range(k + 1, k + 1 + n)]
        frames.append(np.array(rows))
        k += 1 + n
    return frames


def make_water423():
    psf = open(os.path.join(REFTESTS, "data/water_423.psf")).read().split("\n")
    i = next(k for k, l in enumerate(psf) if "!NATOM" in l)
    nat = int(psf[i].split()[0])
    types, q, m, resid = [], [], [], []
    for l in psf[i + 1:i + 1 + nat]:
        p = l.split()
        resid.append(int(p[2]))
        types.append(p[5])
        q.append(float(p[6]))
        m.append(float(p[7]))

    def section(tag, width):
        j = next(k for k, l in enumerate(psf) if tag in l)
        cnt = int(psf[j].split()[0])
        vals = []
        k = j + 1
        while len(vals) < cnt * width:
            vals += [int(v) - 1 for v in psf[k].split()]
            k += 1
        return np.array(vals, np.int32).reshape(cnt, width)

    bonds = section("!NBOND", 2)
    angles = section("!NTHETA", 3)
    # CHARMM parameters for TIP3 (par_all27_prot_lipid.inp): nonbonded eps, Rmin/2;
    # bond OT-HT 450 / 0.9572, angle HT-OT-HT 55 / 104.52 deg
    par = {}
    for l in open(os.path.join(REFTESTS, "data/par_all27_prot_lipid.inp"), errors="replace"):
        p = l.split("!")[0].split()
        if len(p) >= 4 and p[0] in ("OT", "HT") and len(p) in (4, 7):
            try:
                par[p[0]] = (abs(float(p[2])), float(p[3]))
            except ValueError:
                pass
    tnames = ["OT", "HT"]
    tidx = np.array([tnames.index(t) for t in types], np.int32)
    table = np.zeros((2, 2, 4))
    for a in range(2):
        for b in range(2):
            ea, ra = par[tnames[a]]
            eb, rb = par[tnames[b]]
            r = ra + rb
            e = math.sqrt(ea * eb)
            # CHARMM mixing (BuildTopology.cpp:664-672)
            A, B = r ** 12 * e, 2 * r ** 6 * e
            table[a, b] = (A, B, A, B)
    x0 = _pdb_xyz(os.path.join(REFTESTS, "data/water_423.pos.pdb"))
    v0 = _pdb_xyz(os.path.join(REFTESTS, "data/water_423.vel.pdb")) / 20.45482706  # PMConstants.cpp:71
    fr = _xyz_frames(os.path.join(REFTESTS, "expected/water_CHARMM_PERIODIC_L.forces"))
    pos = _xyz_frames(os.path.join(REFTESTS, "expected/water_CHARMM_PERIODIC_L.pos"))[0]
    vel = _xyz_frames(os.path.join(REFTESTS, "expected/water_CHARMM_PERIODIC_L.vel"))[0]
    en = np.array([[float(v) for v in l.split()] for l in
                   open(os.path.join(REFTESTS, "expected/water_CHARMM_PERIODIC_L.energy")) if l.strip()])
    # molecules = residues (BuildTopology assigns molecules from bonded connectivity)
    _, mol = np.unique(np.array(resid), return_inverse=True)
    np.savez_compressed(os.path.join(HERE, "water423.npz"), x0=x0, v0=v0, charge=np.array(q), mass=np.array(m),
                        type=tidx, lj_table=table, molecule=mol.astype(np.int32), bonds=bonds, angles=angles,
                        bond_k=np.full(len(bonds), 450.0), bond_r0=np.full(len(bonds), 0.9572),
                        angle_k=np.full(len(angles), 55.0), angle_t0=np.full(len(angles), math.radians(104.52)),
                        forces0=fr[0], forces200=fr[1], pos200=pos, vel200=vel, energy=en,
                        box=np.array([28.0, 28.0, 28.0]), origin=np.zeros(3), dt=0.5, steps=200, cutoff=6.5,
                        switchon=0.1, cellsize=6.5)


def copy_water_expected():
    """The reference's own expected output files of water_CHARMM_PERIODIC_L
    (data the reference's test holds), for the output-format checks."""
    import shutil
    dst = os.path.join(HERE, "water423_expected")
    os.makedirs(dst, exist_ok=True)
    for ext in ("forces", "pos", "vel", "energy", "energy.header", "dcd"):
        shutil.copyfile(os.path.join(REFTESTS, f"expected/water_CHARMM_PERIODIC_L.{ext}"),
                        os.path.join(dst, f"water_CHARMM_PERIODIC_L.{ext}"))


# velocity starts: (tag, randomtype, seed, temperature); the system's masses
# are those of the water fixture with every 7th atom made heavier
VELOCITY_CASES = (("t0", 0, 1234, 300.0), ("t1", 1, 4242, 310.0), ("t0_big_seed", 0, 3000000000, 250.0))


def make_many_types():
    """A 40-type LJ table (more types than the pair kernel's LDS copy holds):
    the united-atom chains with every site given one of 40 seeded types, LJ +
    Coulomb C2 cutoff pairs with MODIFIED 1-4 pairs."""
    c = S.chain_box(512, 52.0, seed=3, n_types=40)
    ta, tb, sa, sb = CUTOFF_CASES["c2c2"]
    r = refio.system_records(c)
    r.update(mode=np.int32(1), term_a=np.int32(ta), term_b=np.int32(tb), sw_a=np.array(sa, float),
             sw_b=np.array(sb, float), alg_cutoff=10.0, cellsize=5.0, virial=np.int32(1))
    o = refio.run(r)
    np.savez_compressed(os.path.join(HERE, "ref_lj40.npz"), x=c.x, f=o["f"].reshape(-1, 3), e=o["e"],
                        census=np.array(o["census"], dtype=np.uint64))


def make_velocity():
    """The reference's velocity start (MainModule's randomNumber(seed,
    randomtype) seeding, randomVelocity, kineticEnergy, temperature) and
    further uniform and Gaussian draws, for both random-number types."""
    w = water_fixture_system()
    mass = w.mass.copy()
    mass[::7] *= 3.5
    out = {"mass": mass}
    for tag, rtype, seed, temp in VELOCITY_CASES:
        r = {"n": np.int32(w.n), "x": np.zeros(3 * w.n), "mass": mass, "mode": np.int32(5),
             "seed": np.int32(np.uint32(seed).view(np.int32)), "randomtype": np.int32(rtype),
             "temperature": temp, "ndraw": np.int32(4000)}
        o = refio.run(r)
        out[f"{tag}_v"] = o["f"].reshape(-1, 3)
        out[f"{tag}_u0"] = o["u0"][0]
        out[f"{tag}_ke"], out[f"{tag}_temp"] = o["ke_t"][0], o["ke_t"][1]
        out[f"{tag}_u"], out[f"{tag}_g"] = o["u"], o["g"]
        out[f"{tag}_params"] = np.array([rtype, seed, temp], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "ref_velocity.npz"), **out)


# the reference's GROMACS/AMBER implicit-solvent regressions: conf -> (pdb,
# top, force-field directory, GB parameter file or None), all under test/tests/data
GB_LL_CASES = {
    "alanylalanine": ("alanylalanine_capped_processed.withH.pdb", "alanylalanine_capped_processed.top", "ffamber96",
                      None),
    "wwwfip35": ("www-fip35.pdb", "www-fip35.top", "ffamber03", "params.agb"),
}


def _path_rec(p):
    return np.frombuffer(p.encode(), dtype=np.uint8).astype(np.int32)


def make_gb_ll(name):
    """test/tests/{alanylalanine,www-fip35}_FFAMBER_GB_LL.conf at step 0: the
    reference reads the inputs with its own parsers and builds the topology
    (ref_driver mode 7); the fixture keeps that topology as arrays and each GB
    group force of it (GBForce after the Born-radii and partial-sum passes,
    GBACEForce, LennardJones, Coulomb; NonbondedSimpleFull, vacuum), and its
    harmonic bonds and angles with their forces."""
    pdb, top, ff, gb = GB_LL_CASES[name]
    data = os.path.join(REFTESTS, "data")
    r = {"mode": np.int32(7), "n": np.int32(0), "path_pdb": _path_rec(os.path.join(data, pdb)),
         "path_top": _path_rec(os.path.join(data, top)), "path_ff": _path_rec(os.path.join(data, ff)),
         "eps_in": 1.0, "eps_out": 78.3}
    if gb:
        r["path_gb"] = _path_rec(os.path.join(data, gb))
    # the parsers resolve the force-field directory relative to the run's directory
    cwd = os.getcwd()
    os.chdir(REFTESTS)
    try:
        o = refio.run(r)
    finally:
        os.chdir(cwd)
    n = o["charge"].size
    T = int(o["scalars"][2])
    ex = o["excl"].reshape(-1, 3).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, f"ref_gb_ll_{name}.npz"), x=o["x"].reshape(-1, 3), charge=o["charge"],
                        mass=o["mass"], type=o["type"].astype(np.int32), molecule=o["molecule"].astype(np.int32),
                        gb_radius=o["gb_radius"], gb_scaling=o["gb_scaling"], gb_offset=o["gb_offset"],
                        born=o["born"], lj_table=o["lj"].reshape(T, T, 4), exclusions=ex,
                        coulomb_scaling=o["scalars"][0], lj_scaling=o["scalars"][1],
                        f_gb=o["f_gb"].reshape(-1, 3), e_gb=o["e_gb"], f_ace=o["f_ace"].reshape(-1, 3),
                        e_ace=o["e_ace"], f_lj=o["f_lj"].reshape(-1, 3), e_lj=o["e_lj"],
                        f_coul=o["f_coul"].reshape(-1, 3), e_coul=o["e_coul"],
                        bonds=o["bonds"].reshape(-1, 4), angles=o["angles"].reshape(-1, 7),
                        f_bonded=o["f_bonded"].reshape(-1, 3), e_bonded=o["e_bonded"])
    print(name, n, "atoms", T, "types", len(ex), "exclusions")


def main():
    if sys.argv[1:] == ["gb_ll"]:
        for name in GB_LL_CASES:
            make_gb_ll(name)
        return
    w = water_fixture_system()
    if sys.argv[1:] == ["velocity"]:
        make_velocity()
        return
    if sys.argv[1:] == ["vacuum"]:
        d = droplet_fixture_system()
        for name, case in VACUUM_CASES.items():
            make_vacuum(name, d, case)
        return
    if sys.argv[1:] == ["lj40"]:
        make_many_types()
        return
    if sys.argv[1:] == ["cutoff_switch"]:
        for name in ("coul_cut", "cut_cut_two", "cut_c2_two", "c2_cut_two"):
            make_cutoff(name, w, CUTOFF_CASES[name])
        return
    if sys.argv[1:] == ["ewald_recip"]:
        make_ewald_reciprocal(w)
        return
    if sys.argv[1:] == ["expected"]:
        copy_water_expected()
        return
    if sys.argv[1:] == ["gb_ace"]:
        make_gb_ace(200, seed=13)
        return
    if sys.argv[1:] == ["molvirial"]:
        make_molvirial(w)
        return
    if sys.argv[1:] == ["cutoff_two"]:
        make_cutoff("c2c2_two", w, CUTOFF_CASES["c2c2_two"])
        make_cutoff("c2c1_two", w, CUTOFF_CASES["c2c1_two"])
        return
    if sys.argv[1:] == ["modified"]:
        make_modified()
        make_gb_modified()
        return
    for name, case in CUTOFF_CASES.items():
        make_cutoff(name, w, case)
    make_ewald(w)
    make_ewald_reciprocal(w)
    make_bonded(w)
    make_gb(300, seed=7)
    make_gb_ace(200, seed=13)
    make_molvirial(w)
    make_modified()
    make_gb_modified()
    make_water423()
    copy_water_expected()
    make_velocity()
    make_many_types()
    for name in GB_LL_CASES:
        make_gb_ll(name)
    d = droplet_fixture_system()
    for name, case in VACUUM_CASES.items():
        make_vacuum(name, d, case)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
