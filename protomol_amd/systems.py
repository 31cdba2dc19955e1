"""Synthetic systems of the benchmark configurations, generated from seeds.

A ``System`` holds what ProtoMol's ``GenericTopology`` holds for the nonbonded
path (protomol/topology/GenericTopology.h:65-210): per-atom scaled charge,
mass, LJ type or per-atom sigma/epsilon, molecule id, the exclusion table in
``ExclusionTable::add`` order, harmonic bonds/angles, GB parameters and an
orthogonal periodic box.  Units are ProtoMol's: Angstrom, kcal/mol, amu,
charges scaled by sqrt(Coulomb constant) (protomol/base/PMConstants.cpp:67,
protomol/topology/BuildTopology.cpp:92).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SQRT_COULOMB = 18.2226123264      # PMConstants.cpp:67
BOLTZMANN = 0.001987191           # PMConstants.cpp:70
INV_TIMEFACTOR = 0.02045482828087295  # PMConstants.cpp:54

EXCL_FULL = 1
EXCL_MODIFIED = 2

# three-site water: O-H 0.9572 A, H-O-H 104.52 deg, charges -0.834/+0.417,
# LJ on O only (sigma 3.1507 A, epsilon 0.1521 kcal/mol)
WATER_ROH = 0.9572
WATER_THETA = math.radians(104.52)
WATER_Q = (-0.834, 0.417, 0.417)
WATER_MASS = (15.9994, 1.008, 1.008)
WATER_SIGMA = 3.1507
WATER_EPS = 0.1521
# intramolecular model: the CHARMM TIP3 bond/angle constants of the
# reference's water test (par_all27_prot_lipid.inp: OT-HT 450/0.9572, HT-OT-HT 55/104.52)
WATER_KB = 450.0
WATER_KA = 55.0


@dataclasses.dataclass
class System:
    x: np.ndarray                  # (n, 3) float64
    charge: np.ndarray             # (n,) scaled charges
    mass: np.ndarray               # (n,)
    molecule: np.ndarray           # (n,) int32
    box: np.ndarray                # (3,) orthogonal edges (0 for vacuum)
    origin: np.ndarray             # (3,)
    periodic: bool = True
    type: np.ndarray | None = None           # (n,) int32, table mode
    lj_table: np.ndarray | None = None       # (T, T, 4) A, B, A14, B14
    lj_sigma: np.ndarray | None = None       # (n,) per-atom mode
    lj_epsilon: np.ndarray | None = None     # (n,)
    lj_scaling: float = 1.0
    exclusions: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    coulomb_scaling: float = 1.0
    bonds: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bond_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    bond_r0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    angles: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    angle_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    angle_t0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    gb_radius: np.ndarray | None = None
    gb_scaling: np.ndarray | None = None
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def lj_table_full(self) -> np.ndarray:
        """A T x T x 4 table; per-atom parameters become one type per atom
        with GROMACS mixing (BuildTopology.cpp:650-659)."""
        if self.lj_table is not None:
            return self.lj_table
        s, e = self.lj_sigma, self.lj_epsilon
        sij = 0.5 * (s[:, None] + s[None, :])
        eij = np.sqrt(e[:, None] * e[None, :])
        A = sij ** 12 * eij * 4.0
        B = sij ** 6 * eij * 4.0
        return np.stack([A, B, self.lj_scaling * A, self.lj_scaling * B], axis=-1)

    def lj_types(self) -> np.ndarray:
        if self.type is not None:
            return self.type
        return np.arange(self.n, dtype=np.int32)


def lj_gromacs(sigma: float, eps: float) -> tuple[float, float]:
    """A = 4 eps sigma^12, B = 4 eps sigma^6 (BuildTopology.cpp:650-659)."""
    return 4.0 * eps * sigma ** 12, 4.0 * eps * sigma ** 6


def _random_rotations(rng: np.random.Generator, m: int) -> np.ndarray:
    # uniform rotations from unit quaternions
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def _relieve_contacts(rng, centers, R, local, lo, box, min_contact, n_try=32, max_iter=200):
    """Re-orient molecules whose hydrogens come closer than ``min_contact`` to
    a hydrogen of another molecule (periodic box of edge ``box`` whose lower
    corner is ``lo``).  Each offender takes the best of ``n_try`` random
    orientations, judged by the distance from its hydrogens to the other
    molecules' hydrogens.  (O-H contacts are hydrogen bonds and stay.)"""
    from scipy.spatial import cKDTree
    xyz = centers[:, None, :] + np.einsum("mij,aj->mai", R, local)
    n_mol = len(centers)

    def wrap(p):
        q = np.mod(p - lo, box)
        q[q >= box] = 0.0
        return q

    for _ in range(max_iter):
        hyd = xyz[:, 1:3, :].reshape(-1, 3)          # H atoms, molecule = index // 2
        tree = cKDTree(wrap(hyd), boxsize=box)
        pairs = tree.query_pairs(min_contact, output_type="ndarray")
        mi, mj = pairs[:, 0] // 2, pairs[:, 1] // 2
        inter = mi != mj
        if not inter.any():
            return xyz
        # one molecule of each contact, chosen at random: a molecule hemmed in
        # by its neighbours' hydrogens may need a neighbour to turn instead
        pick = rng.random(int(inter.sum())) < 0.5
        redo = np.unique(np.where(pick, mi[inter], mj[inter]))
        cand = _random_rotations(rng, len(redo) * n_try).reshape(len(redo), n_try, 3, 3)
        pos = centers[redo, None, None, :] + np.einsum("mtij,aj->mtai", cand, local[1:3])
        # the two nearest hydrogens: one may be the molecule's own other H
        d, idx = tree.query(wrap(pos.reshape(-1, 3)), k=3)
        own = (idx // 2) == np.repeat(redo, n_try * 2)[:, None]
        d = np.where(own | (idx >= 2 * n_mol), np.inf, d)
        score = d.min(axis=1).reshape(len(redo), n_try, 2).min(axis=2)
        best = np.argmax(score, axis=1)
        R[redo] = cand[np.arange(len(redo)), best]
        xyz[redo] = centers[redo, None, :] + np.einsum("mij,aj->mai", R[redo], local)
    raise RuntimeError("water_box: could not remove close contacts")


def water_box(n_mol: int, box: float, seed: int, origin=(0.0, 0.0, 0.0), jitter: float = 0.0,
              atom_noise: float = 0.0, min_contact: float = 0.0) -> System:
    """``n_mol`` rigid-geometry 3-site waters on a cubic lattice in a cubic box.

    The lattice has m = ceil(n_mol^(1/3)) sites per edge (spacing box/m); when
    n_mol < m^3 a seeded subset of sites is used.  Each molecule gets a seeded
    random orientation.  Atoms are ordered O, H, H per molecule.  ``jitter``
    moves molecule centres uniformly by up to that many A, ``atom_noise`` adds
    Gaussian noise (A) to every atom, away from the ideal geometry.
    ``min_contact`` > 0 re-draws the orientation of molecules that have a
    hydrogen closer than that to a hydrogen of another molecule (periodic),
    until no such contact is left: random orientations on a ~3 A lattice put
    some H-H pairs at 1.2 A, whose repulsion would heat the start of a run.
    """
    rng = np.random.default_rng(seed)
    m = int(math.ceil(round(n_mol ** (1.0 / 3.0), 9)))
    while m ** 3 < n_mol:
        m += 1
    a = box / m
    sites = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    if n_mol < m ** 3:
        keep = np.sort(rng.choice(m ** 3, size=n_mol, replace=False))
        sites = sites[keep]
    org = np.asarray(origin, float)
    centers = (sites + 0.5) * a - 0.5 * box + org
    if jitter:
        centers = centers + rng.uniform(-jitter, jitter, size=centers.shape)
    half = 0.5 * WATER_THETA
    local = np.array([[0.0, 0.0, 0.0],
                      [WATER_ROH * math.sin(half), 0.0, WATER_ROH * math.cos(half)],
                      [-WATER_ROH * math.sin(half), 0.0, WATER_ROH * math.cos(half)]])
    R = _random_rotations(rng, n_mol)
    xyz = centers[:, None, :] + np.einsum("mij,aj->mai", R, local)
    if min_contact > 0.0:
        xyz = _relieve_contacts(rng, centers, R, local, org - 0.5 * box, box, min_contact)
    n = 3 * n_mol
    x = xyz.reshape(n, 3).astype(np.float64)
    if atom_noise:
        x = x + rng.normal(0.0, atom_noise, size=x.shape)
    charge = np.tile(np.array(WATER_Q) * SQRT_COULOMB, n_mol)
    mass = np.tile(np.array(WATER_MASS), n_mol)
    typ = np.tile(np.array([0, 1, 1], np.int32), n_mol)
    mol = np.repeat(np.arange(n_mol, dtype=np.int32), 3)
    A, B = lj_gromacs(WATER_SIGMA, WATER_EPS)
    table = np.zeros((2, 2, 4))
    table[0, 0] = (A, B, A, B)
    base = 3 * np.arange(n_mol, dtype=np.int32)
    # 1-2 (O-H1, O-H2) and 1-3 (H1-H2) exclusions in buildExclusionTable order:
    # bonds first, then angle ends (BuildTopology.cpp:1119-1131)
    ex = np.concatenate([
        np.stack([base, base + 1, np.full(n_mol, EXCL_FULL)], 1),
        np.stack([base, base + 2, np.full(n_mol, EXCL_FULL)], 1),
        np.stack([base + 1, base + 2, np.full(n_mol, EXCL_FULL)], 1),
    ]).astype(np.int32)
    order = np.argsort(np.concatenate([2 * np.arange(n_mol), 2 * np.arange(n_mol) + 1, 2 * n_mol + np.arange(n_mol)]),
                       kind="stable")
    ex = ex[order]
    bonds = np.stack([np.stack([base, base + 1], 1), np.stack([base, base + 2], 1)], 1).reshape(-1, 2).astype(np.int32)
    angles = np.stack([base + 1, base, base + 2], 1).astype(np.int32)
    return System(x=x, charge=charge, mass=mass, molecule=mol, box=np.full(3, float(box)), origin=org.copy(),
                  periodic=True, type=typ, lj_table=table, exclusions=ex,
                  bonds=bonds, bond_k=np.full(len(bonds), WATER_KB), bond_r0=np.full(len(bonds), WATER_ROH),
                  angles=angles, angle_k=np.full(n_mol, WATER_KA), angle_t0=np.full(n_mol, WATER_THETA),
                  name=f"water{n_mol}")


def vacuum_region(x: np.ndarray, margin: float = 10.0) -> tuple[np.ndarray, np.ndarray]:
    """(box, origin) of the region a vacuum system's atoms occupy: their
    bounding box grown by ``margin`` on every side.  The cutoff path keeps the
    atoms of a vacuum system in this region (origin +- box / 2)."""
    lo, hi = x.min(axis=0), x.max(axis=0)
    return (hi - lo) + 2.0 * margin, 0.5 * (lo + hi)


def water_droplet(radius: float, seed: int, margin: float = 10.0) -> System:
    """The whole molecules of a water box whose oxygen lies within ``radius``
    A of the box centre (its origin), in vacuum (VacuumBoundaryConditions), with the
    region of ``vacuum_region``."""
    edge = 2.0 * radius + 6.0
    n_mol = int(math.ceil(0.0334 * edge ** 3))
    w = water_box(n_mol, edge, seed=seed, jitter=0.3, atom_noise=0.03)
    o = w.x[0::3]
    keep = np.flatnonzero(np.linalg.norm(o - np.asarray(w.origin), axis=1) <= radius)
    atoms = (3 * keep[:, None] + np.arange(3)[None, :]).reshape(-1)
    remap = -np.ones(w.n, np.int64)
    remap[atoms] = np.arange(atoms.size)
    ex = w.exclusions[np.isin(w.exclusions[:, 0], atoms) & np.isin(w.exclusions[:, 1], atoms)]
    ex = np.column_stack([remap[ex[:, 0]], remap[ex[:, 1]], ex[:, 2]]).astype(np.int32)
    bsel = np.isin(w.bonds[:, 0], atoms)
    asel = np.isin(w.angles[:, 0], atoms)
    x = w.x[atoms].copy()
    box, org = vacuum_region(x, margin)
    mol = (np.arange(atoms.size) // 3).astype(np.int32)
    return dataclasses.replace(
        w, x=x, charge=w.charge[atoms], mass=w.mass[atoms], molecule=mol, box=box, origin=org, periodic=False,
        type=w.type[atoms], exclusions=ex, bonds=remap[w.bonds[bsel]].astype(np.int32), bond_k=w.bond_k[bsel],
        bond_r0=w.bond_r0[bsel], angles=remap[w.angles[asel]].astype(np.int32), angle_k=w.angle_k[asel],
        angle_t0=w.angle_t0[asel], name=f"water droplet r={radius:g} A ({keep.size} molecules)")


def water_box_density(n_mol: int, seed: int, density: float = 0.0334) -> System:
    """Water box of ``n_mol`` molecules at ``density`` molecules/A^3."""
    box = (n_mol / density) ** (1.0 / 3.0)
    return water_box(n_mol, box, seed)


# united-atom alkane chain (TraPPE-like CH3 / CH2 sites): bonds 1.54 A,
# angles 114 deg; CH3 sigma 3.75 A eps 0.195, CH2 sigma 3.95 A eps 0.0914
# kcal/mol; 1-4 pairs with their own LJ coefficients (sigma 0.9x, eps 0.5x)
# and the Coulomb scale 1/1.2
CHAIN_BOND = 1.54
CHAIN_THETA = math.radians(114.0)
CHAIN_SIGMA = (3.75, 3.95)
CHAIN_EPS = (0.195, 0.0914)


def chain_box(n_mol: int, box: float, seed: int, length: int = 4, per_atom: bool = False,
              atom_noise: float = 0.05, n_types: int = 2) -> System:
    """Periodic box of ``n_mol`` united-atom chains of ``length`` sites on a
    cubic lattice, random orientations, planar zig-zag geometry plus Gaussian
    ``atom_noise`` (A).  Exclusions: 1-2 and 1-3 pairs FULL, 1-4 pairs
    MODIFIED, so the pair path's scaled (1-4) branch is exercised: charges
    +-0.25 e alternate along the chain and ``coulomb_scaling`` is 1/1.2.
    ``per_atom`` gives per-atom sigma / epsilon with GROMACS mixing and
    ``lj_scaling`` 0.5 instead of the type table (whose 1-4 entries use
    sigma 0.9x and epsilon 0.5x).  ``n_types`` > 2 gives every site one of
    that many types (seeded), each with its own sigma U[2.8, 4.0] A and
    epsilon U[0.05, 0.2] kcal/mol, mixed and scaled the same way."""
    rng = np.random.default_rng(seed)
    m = 1
    while m ** 3 < n_mol:
        m += 1
    a = box / m
    sites = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    if n_mol < m ** 3:
        sites = sites[np.sort(rng.choice(m ** 3, size=n_mol, replace=False))]
    centers = (sites + 0.5) * a - 0.5 * box
    half = 0.5 * (math.pi - CHAIN_THETA)
    k = np.arange(length)
    local = np.stack([k * CHAIN_BOND * math.cos(half), (k % 2) * CHAIN_BOND * math.sin(half), np.zeros(length)], 1)
    local -= local.mean(axis=0)
    R = _random_rotations(rng, n_mol)
    x = (centers[:, None, :] + np.einsum("mij,aj->mai", R, local)).reshape(-1, 3)
    x = x + rng.normal(0.0, atom_noise, size=x.shape)
    n = n_mol * length
    end = (k == 0) | (k == length - 1)
    typ = np.tile(np.where(end, 0, 1).astype(np.int32), n_mol)
    charge = np.tile(np.where(k % 2 == 0, 0.25, -0.25) * SQRT_COULOMB, n_mol)
    if length % 2:
        charge[length - 1::length] = 0.0
    mass = np.tile(np.where(end, 15.035, 14.027), n_mol)
    mol = np.repeat(np.arange(n_mol, dtype=np.int32), length)
    base = length * np.arange(n_mol, dtype=np.int32)
    ex = []
    for d, cls in ((1, EXCL_FULL), (2, EXCL_FULL), (3, EXCL_MODIFIED)):
        for i in range(length - d):
            ex.append(np.stack([base + i, base + i + d, np.full(n_mol, cls)], 1))
    ex = np.concatenate(ex).astype(np.int32)
    bonds = np.concatenate([np.stack([base + i, base + i + 1], 1) for i in range(length - 1)]).astype(np.int32)
    kw = dict(x=x, charge=charge, mass=mass, molecule=mol, box=np.full(3, float(box)), origin=np.zeros(3),
              periodic=True, exclusions=ex, coulomb_scaling=1.0 / 1.2, bonds=bonds,
              name=f"chain{length}x{n_mol}")
    if per_atom:
        sig = np.array(CHAIN_SIGMA)[typ]
        eps = np.array(CHAIN_EPS)[typ]
        return System(lj_sigma=sig, lj_epsilon=eps, lj_scaling=0.5, **kw)
    sigmas, epss = list(CHAIN_SIGMA), list(CHAIN_EPS)
    if n_types > 2:
        trng = np.random.default_rng(seed + 7919)
        sigmas = trng.uniform(2.8, 4.0, size=n_types).tolist()
        epss = trng.uniform(0.05, 0.2, size=n_types).tolist()
        typ = trng.integers(0, n_types, size=n).astype(np.int32)
    nt = len(sigmas)
    table = np.zeros((nt, nt, 4))
    for ti in range(nt):
        for tj in range(nt):
            sg = 0.5 * (sigmas[ti] + sigmas[tj])
            ep = math.sqrt(epss[ti] * epss[tj])
            A, B = 4 * ep * sg ** 12, 4 * ep * sg ** 6
            A14, B14 = 4 * 0.5 * ep * (0.9 * sg) ** 12, 4 * 0.5 * ep * (0.9 * sg) ** 6
            table[ti, tj] = (A, B, A14, B14)
    return System(type=typ, lj_table=table, **kw)


def rock_salt(n_cells: int, spacing: float = 2.82, origin=(0.0, 0.0, 0.0), jitter: float = 0.0,
              seed: int = 0) -> System:
    """NaCl lattice of n_cells^3 conventional cells (8 ions each) with
    nearest-neighbour distance `spacing`: charges +-1 e (scaled), LJ off, each
    ion its own molecule.  Unjittered, its Coulomb energy is the Madelung
    energy -N M q^2 / (2 spacing), M = 1.747564594633..."""
    m = 2 * n_cells
    idx = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    sign = np.where(idx.sum(1) % 2 == 0, 1.0, -1.0)
    box = np.full(3, m * spacing)
    org = np.asarray(origin, np.float64)
    x = (idx + 0.5) * spacing - box / 2 + org
    if jitter:
        x = x + np.random.default_rng(seed).uniform(-jitter, jitter, x.shape)
    n = len(x)
    return System(x=x, charge=sign * SQRT_COULOMB, mass=np.where(sign > 0, 22.98977, 35.453),
                  molecule=np.arange(n, dtype=np.int32), box=box, origin=org, periodic=True,
                  type=np.zeros(n, np.int32), lj_table=np.zeros((1, 1, 4)), name=f"nacl{n}")


def ensemble_sizes(count: int, lo: int, hi: int, seed: int) -> list[int]:
    """Molecule counts drawn log-uniform in [lo, hi] by seed."""
    rng = np.random.default_rng(seed)
    return [int(math.floor(math.exp(v))) for v in rng.uniform(math.log(lo), math.log(hi), size=count)]


_ELEMENTS = ("C", "N", "O", "H", "S", "P")
_ELEMENT_MASS = {"C": 12.011, "N": 14.007, "O": 15.999, "H": 1.008, "S": 32.06, "P": 30.974}
# GB scaling factor by the first letter of the atom type (BuildTopology.cpp:843-858)
_ELEMENT_GBS = {"H": 0.85, "C": 0.72, "N": 0.79, "O": 0.85, "P": 0.86, "S": 0.96}


CHAIN_KB = 310.0  # kcal/mol/A^2, the CHARMM aliphatic C-C bond constant
CHAIN_KA = 50.0   # kcal/mol/rad^2


def gb_chain(n_atoms: int, seed: int, bond: float = 1.53, min_dist: float = 2.5) -> System:
    """Bead polymer in vacuum: seeded self-avoiding random walk of ``bond``-A
    steps (non-bonded beads kept at least ``min_dist`` apart), charges
    N(0, 0.3) e shifted to net zero, LJ sigma U[2.5, 4.0] A, epsilon
    U[0.05, 0.2] kcal/mol (per-atom GROMACS mixing), intrinsic Born radii
    U[1.2, 2.0] A, 1-2 and 1-3 exclusions along the chain.  The chain's
    covalent geometry holds it together in MD: harmonic bonds (310 kcal/mol/A^2,
    rest length ``bond``) and angles (50 kcal/mol/rad^2, each at its start
    value, so the start is unstrained); without them the excluded bonded
    neighbours, which feel no Lennard-Jones repulsion, drift onto each other."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n_atoms, 3))
    cell = min_dist
    grid: dict[tuple[int, int, int], list[int]] = {}

    def key(p):
        return (int(math.floor(p[0] / cell)), int(math.floor(p[1] / cell)), int(math.floor(p[2] / cell)))

    def ok(p, upto):
        k = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if j < upto - 1 and np.sum((pos[j] - p) ** 2) < min_dist * min_dist:
                            return False
        return True

    grid.setdefault(key(pos[0]), []).append(0)
    i = 1
    stuck = 0
    while i < n_atoms:
        for _ in range(100):
            v = rng.normal(size=3)
            p = pos[i - 1] + bond * v / np.linalg.norm(v)
            if ok(p, i):
                break
        else:
            # dead end: drop the last few beads and grow again
            stuck += 1
            if stuck > 100 * n_atoms:
                raise RuntimeError("self-avoiding walk got stuck")
            back = min(i - 1, 5)
            for j in range(i - back, i):
                grid[key(pos[j])].remove(j)
            i -= back
            continue
        pos[i] = p
        grid.setdefault(key(p), []).append(i)
        i += 1
    pos -= pos.mean(axis=0)
    q = rng.normal(0.0, 0.3, size=n_atoms)
    q -= q.mean()
    elem = rng.choice(len(_ELEMENTS), size=n_atoms, p=[0.35, 0.15, 0.15, 0.25, 0.05, 0.05])
    mass = np.array([_ELEMENT_MASS[_ELEMENTS[e]] for e in elem])
    gbs = np.array([_ELEMENT_GBS[_ELEMENTS[e]] for e in elem])
    sigma = rng.uniform(2.5, 4.0, size=n_atoms)
    eps = rng.uniform(0.05, 0.2, size=n_atoms)
    rho = rng.uniform(1.2, 2.0, size=n_atoms)
    idx = np.arange(n_atoms, dtype=np.int32)
    bonds = np.stack([idx[:-1], idx[1:]], 1)
    angles = np.stack([idx[:-2], idx[1:-1], idx[2:]], 1)
    ex = np.concatenate([np.concatenate([bonds, np.full((len(bonds), 1), EXCL_FULL)], 1),
                         np.concatenate([angles[:, [0, 2]], np.full((len(angles), 1), EXCL_FULL)], 1)]).astype(np.int32)
    u = pos[angles[:, 0]] - pos[angles[:, 1]]
    w = pos[angles[:, 2]] - pos[angles[:, 1]]
    theta0 = np.arctan2(np.linalg.norm(np.cross(u, w), axis=1), np.sum(u * w, axis=1))
    return System(x=pos, charge=q * SQRT_COULOMB, mass=mass, molecule=np.zeros(n_atoms, np.int32),
                  box=np.zeros(3), origin=np.zeros(3), periodic=False, lj_sigma=sigma, lj_epsilon=eps,
                  lj_scaling=1.0, exclusions=ex, bonds=bonds.astype(np.int32),
                  bond_k=np.full(len(bonds), CHAIN_KB), bond_r0=np.full(len(bonds), bond),
                  angles=angles.astype(np.int32), angle_k=np.full(len(angles), CHAIN_KA), angle_t0=theta0,
                  gb_radius=rho, gb_scaling=gbs, name=f"gbchain{n_atoms}")


def maxwell_velocities(sys: System, temperature: float, seed: int, randomtype: int = 0) -> np.ndarray:
    """Seeded Maxwell velocities, sigma = sqrt(kB T / m) per component, in
    ProtoMol velocity units: the velocity start of a ProtoMol run with this
    seed and randomtype (MainModule.cpp:88-102, 143-147;
    TopologyUtilities.cpp:21-48), drawn by the library's pm_random_velocity."""
    from protomol_amd import _lib
    return _lib.random_velocity(sys.mass, temperature, seed, randomtype)


def kinetic_energy(sys: System, v: np.ndarray) -> float:
    """TopologyUtilities.cpp:163-171"""
    return 0.5 * float(np.sum(sys.mass * np.sum(v * v, axis=1)))


def temperature(sys: System, v: np.ndarray, dof: int | None = None) -> float:
    """T = 2 KE / (kB * dof), dof = 3N - 3 (BuildTopology.cpp:121, TopologyUtilities.cpp:109-111)."""
    if dof is None:
        dof = 3 * sys.n - 3
    return 2.0 * kinetic_energy(sys, v) / (BOLTZMANN * dof)
