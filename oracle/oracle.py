"""ctypes wrapper of liboracle.so, the FP64 C restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker / CPU baseline, never as
the product path.  Build with ``make -C oracle lib``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SW_UNIVERSAL, SW_CUTOFF, SW_C1, SW_C2, SW_CN, SW_CMPCNCN = range(6)
TERM_NONE, TERM_LJ, TERM_COULOMB, TERM_EWALD_REAL = range(4)
COULOMB, LENNARDJONES, BOND, ANGLE = 0, 1, 2, 3
VIRIALXX = 7
NSLOTS = 36

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSwitch(C.Structure):
    _fields_ = [("kind", C.c_int), ("use", C.c_int), ("modify", C.c_int),
                ("cutoff", C.c_double), ("cutoff2", C.c_double),
                ("switchon", C.c_double), ("switchon2", C.c_double),
                ("switchoff", C.c_double), ("switchoff2", C.c_double),
                ("s1", C.c_double), ("s2", C.c_double), ("s3", C.c_double),
                ("c1_15rc1", C.c_double), ("c1_05rc3", C.c_double), ("c1_15rc3", C.c_double),
                ("ord_int", C.c_int), ("ord_idx", C.c_int),
                ("irange", C.c_double * 8), ("itrange", C.c_double * 8)]


class OrcSystem(C.Structure):
    _fields_ = [("n", C.c_int), ("charge", _dp), ("type", _ip), ("molecule", _ip), ("mass", _dp),
                ("ntypes", C.c_int), ("lj", _dp), ("nexcl", C.c_int), ("excl", _ip),
                ("box", C.c_double * 3), ("origin", C.c_double * 3),
                ("coulomb_scaling", C.c_double), ("periodic", C.c_int)]


class OrcNonbonded(C.Structure):
    _fields_ = [("term_a", C.c_int), ("term_b", C.c_int), ("sw_a", OrcSwitch), ("sw_b", OrcSwitch),
                ("algorithm_cutoff", C.c_double), ("cellsize", C.c_double),
                ("virial", C.c_int), ("molvirial", C.c_int)]


class OrcCensus(C.Structure):
    _fields_ = [("count", C.c_uint64), ("hash", C.c_uint64)]


class OrcEwald(C.Structure):
    _fields_ = [("alpha", C.c_double), ("accuracy", C.c_double), ("rc", C.c_double), ("rc2", C.c_double),
                ("two_alpha_pi", C.c_double), ("cellsize", C.c_double), ("correction", C.c_int),
                ("virial", C.c_int), ("molvirial", C.c_int)]


class OrcGB(C.Structure):
    _fields_ = [("vdw_radius", _dp), ("scaling", _dp), ("offset", C.c_double),
                ("alpha_obc", C.c_double), ("beta_obc", C.c_double), ("gamma_obc", C.c_double),
                ("solute_dielec", C.c_double), ("solvent_dielec", C.c_double), ("block_size", C.c_int),
                ("ace", C.c_int), ("ace_sigma", C.c_double), ("ace_rho_s", C.c_double),
                ("virial", C.c_int), ("molvirial", C.c_int)]


class OrcRng(C.Structure):
    _fields_ = [("type", C.c_int), ("x48", C.c_ushort * 3), ("iv", C.c_int32 * 32), ("iy", C.c_int32),
                ("idum", C.c_int32), ("have_spare", C.c_int), ("spare", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle lib`")
        L = C.CDLL(LIB_PATH)
        L.orc_switch_init.argtypes = [C.POINTER(OrcSwitch), C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_switch_eval.argtypes = [C.POINTER(OrcSwitch), C.c_double, _dp, _dp]
        L.orc_min_diff.argtypes = [C.POINTER(OrcSystem), _dp, _dp, _dp]
        L.orc_min_diff.restype = C.c_double
        L.orc_nonbonded_cutoff.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcNonbonded), _dp, _dp, _dp,
                                           C.POINTER(OrcCensus), C.c_int]
        L.orc_census_bruteforce.argtypes = [C.POINTER(OrcSystem), _dp, C.c_double, C.POINTER(OrcCensus)]
        L.orc_ewald_init.argtypes = [C.POINTER(OrcEwald), C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_ewald_real.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcEwald), _dp, _dp, _dp,
                                     C.POINTER(OrcCensus), C.c_int]
        L.orc_ewald_kvectors.argtypes = [_dp, C.c_double, C.c_double, C.POINTER(_ip), _ip]
        L.orc_ewald_reciprocal.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcEwald), _dp, _dp, _dp, C.c_int]
        L.orc_poly5.argtypes = [C.c_double]
        L.orc_poly5.restype = C.c_double
        L.orc_rng_seed.argtypes = [C.POINTER(OrcRng), C.c_uint32, C.c_int]
        L.orc_rng_uniform.argtypes = [C.POINTER(OrcRng)]
        L.orc_rng_uniform.restype = C.c_double
        L.orc_rng_gaussian.argtypes = [C.POINTER(OrcRng)]
        L.orc_rng_gaussian.restype = C.c_double
        L.orc_random_velocity.argtypes = [C.POINTER(OrcSystem), C.c_double, C.POINTER(OrcRng), _dp]
        L.orc_kinetic_energy.argtypes = [C.POINTER(OrcSystem), _dp]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_gb_evaluate.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcGB), _dp, _dp, _dp, _dp]
        L.orc_bonded.argtypes = [C.POINTER(OrcSystem), C.c_int, _ip, _dp, _dp, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp,
                                 C.c_int]
        L.orc_n_molecules.argtypes = [C.POINTER(OrcSystem)]
        L.orc_molecule_com.argtypes = [C.POINTER(OrcSystem), _dp, _dp]
        L.orc_pair_hash.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_pair_hash.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


class OracleSystem:
    """Keeps the numpy buffers alive behind an OrcSystem."""

    def __init__(self, sys):
        self.sys = sys
        self.charge = np.ascontiguousarray(sys.charge, np.float64)
        self.mass = np.ascontiguousarray(sys.mass, np.float64)
        self.molecule = np.ascontiguousarray(sys.molecule, np.int32)
        self.type = np.ascontiguousarray(sys.lj_types(), np.int32)
        self.lj = np.ascontiguousarray(sys.lj_table_full(), np.float64)
        self.excl = np.ascontiguousarray(sys.exclusions, np.int32)
        s = OrcSystem()
        s.n = sys.n
        s.charge = _p(self.charge)
        s.type = _p(self.type, _ip)
        s.molecule = _p(self.molecule, _ip)
        s.mass = _p(self.mass)
        s.ntypes = self.lj.shape[0]
        s.lj = _p(self.lj)
        s.nexcl = len(self.excl)
        s.excl = _p(self.excl, _ip)
        for k in range(3):
            s.box[k] = float(sys.box[k])
            s.origin[k] = float(sys.origin[k])
        s.coulomb_scaling = float(sys.coulomb_scaling)
        s.periodic = 1 if sys.periodic else 0
        self.c = s


def make_switch(kind, switchon=0.0, cutoff=0.0, switchoff=0.0, order=0.0):
    sw = OrcSwitch()
    lib().orc_switch_init(C.byref(sw), kind, switchon, cutoff, switchoff, order)
    return sw


def switch_eval(sw, r2):
    v = C.c_double()
    d = C.c_double()
    lib().orc_switch_eval(C.byref(sw), r2, C.byref(v), C.byref(d))
    return v.value, d.value


def nonbonded_cutoff(sys, term_a, sw_a, term_b=TERM_NONE, sw_b=None, algorithm_cutoff=None, cellsize=5.0,
                     virial=False, nthreads=1, x=None, census=True, molvirial=False):
    """NonbondedCutoffSystemForce::evaluate on fresh force/energy buffers.
    Returns (forces (n,3), energies[36], (count, hash))."""
    os_ = OracleSystem(sys)
    nb = OrcNonbonded()
    nb.term_a, nb.term_b = term_a, term_b
    nb.sw_a = sw_a
    if sw_b is not None:
        nb.sw_b = sw_b
    cut = max(sw_a.cutoff, sw_b.cutoff if sw_b is not None and term_b else 0.0)
    nb.algorithm_cutoff = algorithm_cutoff if algorithm_cutoff else cut
    nb.cellsize = cellsize
    nb.virial = 1 if virial else 0
    nb.molvirial = 1 if molvirial else 0
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    f = np.zeros(3 * sys.n)
    e = np.zeros(NSLOTS)
    cen = OrcCensus()
    rc = lib().orc_nonbonded_cutoff(C.byref(os_.c), C.byref(nb), _p(xx), _p(f), _p(e),
                                    C.byref(cen) if census else None, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_nonbonded_cutoff failed: {rc}")
    return f.reshape(-1, 3), e, (cen.count, cen.hash)


def ewald_real(sys, alpha, accuracy, cellsize=5.0, correction=True, virial=False, nthreads=1, x=None,
               molvirial=False, census=True):
    os_ = OracleSystem(sys)
    ew = OrcEwald()
    lib().orc_ewald_init(C.byref(ew), alpha, accuracy, cellsize, 1 if correction else 0, 1 if virial else 0)
    ew.molvirial = 1 if molvirial else 0
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    f = np.zeros(3 * sys.n)
    e = np.zeros(NSLOTS)
    cen = OrcCensus()
    rc = lib().orc_ewald_real(C.byref(os_.c), C.byref(ew), _p(xx), _p(f), _p(e), C.byref(cen) if census else None,
                              nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_ewald_real failed: {rc}")
    return f.reshape(-1, 3), e, (cen.count, cen.hash), ew.rc


def ewald_kvectors(box, alpha, accuracy):
    """(h, k, l) of every reciprocal-space vector, in the reference's order,
    and (hmax, kmax, lmax, hklmax)."""
    out = _ip()
    hkl = np.zeros(4, np.int32)
    boxa = np.ascontiguousarray(box, np.float64)
    nk = lib().orc_ewald_kvectors(_p(boxa), alpha, accuracy, C.byref(out), _p(hkl, _ip))
    if nk < 0:
        raise RuntimeError("orc_ewald_kvectors failed")
    kv = np.ctypeslib.as_array(out, shape=(max(nk, 1) * 4,))[: 4 * nk].reshape(-1, 4)[:, :3].copy()
    C.CDLL(None).free(C.cast(out, C.c_void_p))
    return kv, tuple(int(v) for v in hkl)


def ewald_reciprocal(sys, alpha, accuracy, virial=False, nthreads=1, x=None, molvirial=False):
    """Reciprocal-space Ewald term alone: forces (n, 3), 36 energy slots."""
    os_ = OracleSystem(sys)
    ew = OrcEwald()
    lib().orc_ewald_init(C.byref(ew), alpha, accuracy, 5.0, 0, 1 if virial else 0)
    ew.molvirial = 1 if molvirial else 0
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    f = np.zeros(3 * sys.n)
    e = np.zeros(NSLOTS)
    rc = lib().orc_ewald_reciprocal(C.byref(os_.c), C.byref(ew), _p(xx), _p(f), _p(e), nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_ewald_reciprocal failed: {rc}")
    return f.reshape(-1, 3), e


def census_bruteforce(sys, cut2, x=None):
    os_ = OracleSystem(sys)
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    cen = OrcCensus()
    lib().orc_census_bruteforce(C.byref(os_.c), _p(xx), cut2, C.byref(cen))
    return cen.count, cen.hash


ACE_SIGMA = 2.26 / 418.4   # GBACEForce -solvationparam default (GBACEForce.cpp:218-219)
ACE_RHO_S = 1.4            # -watersphereradius default


def gb_evaluate(sys, offset=0.09, a=1.0, b=0.8, g=4.85, eps_in=1.0, eps_out=78.3, x=None, ace=False,
                ace_sigma=ACE_SIGMA, ace_rho_s=ACE_RHO_S, virial=False, molvirial=False):
    os_ = OracleSystem(sys)
    rho = np.ascontiguousarray(sys.gb_radius, np.float64)
    S = np.ascontiguousarray(sys.gb_scaling, np.float64)
    gb = OrcGB()
    gb.vdw_radius = _p(rho)
    gb.scaling = _p(S)
    gb.offset, gb.alpha_obc, gb.beta_obc, gb.gamma_obc = offset, a, b, g
    gb.solute_dielec, gb.solvent_dielec = eps_in, eps_out
    gb.block_size = 64
    gb.ace = 1 if ace else 0
    gb.ace_sigma, gb.ace_rho_s = ace_sigma, ace_rho_s
    gb.virial, gb.molvirial = (1 if virial else 0), (1 if molvirial else 0)
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    f = np.zeros(3 * sys.n)
    e = np.zeros(NSLOTS)
    born = np.zeros(sys.n)
    rc = lib().orc_gb_evaluate(C.byref(os_.c), C.byref(gb), _p(xx), _p(f), _p(e), _p(born))
    if rc != 0:
        raise RuntimeError("orc_gb_evaluate failed")
    return f.reshape(-1, 3), e, born


def bonded(sys, x=None, virial=False):
    os_ = OracleSystem(sys)
    bonds = np.ascontiguousarray(sys.bonds, np.int32)
    angles = np.ascontiguousarray(sys.angles, np.int32)
    bk, br = np.ascontiguousarray(sys.bond_k, np.float64), np.ascontiguousarray(sys.bond_r0, np.float64)
    ak, at = np.ascontiguousarray(sys.angle_k, np.float64), np.ascontiguousarray(sys.angle_t0, np.float64)
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    f = np.zeros(3 * sys.n)
    e = np.zeros(NSLOTS)
    lib().orc_bonded(C.byref(os_.c), len(bonds), _p(bonds, _ip), _p(bk), _p(br), len(angles), _p(angles, _ip),
                     _p(ak), _p(at), _p(xx), _p(f), _p(e), 1 if virial else 0)
    return f.reshape(-1, 3), e


def molecule_com(sys, x=None):
    """Centres of mass of the molecules, (n_molecules, 3)."""
    os_ = OracleSystem(sys)
    xx = np.ascontiguousarray(sys.x if x is None else x, np.float64).reshape(-1)
    nm = lib().orc_n_molecules(C.byref(os_.c))
    com = np.zeros(3 * max(nm, 1))
    lib().orc_molecule_com(C.byref(os_.c), _p(xx), _p(com))
    return com[:3 * nm].reshape(-1, 3)


def velocity_start(sys, temperature, seed, randomtype=0, ndraw=0):
    """The velocity start of a ProtoMol run: randomNumber(seed, randomtype)
    seeds the generator and returns one number (MainModule.cpp:88-102), then
    randomVelocity(temperature, ...) (TopologyUtilities.cpp:21-48); then
    `ndraw` further uniform and then Gaussian numbers of the same streams.
    Returns (first number, velocities (n, 3), uniforms, gaussians)."""
    os_ = OracleSystem(sys)
    r = OrcRng()
    L = lib()
    L.orc_rng_seed(C.byref(r), seed, randomtype)
    u0 = L.orc_rng_uniform(C.byref(r))
    v = np.zeros(3 * sys.n)
    L.orc_random_velocity(C.byref(os_.c), temperature, C.byref(r), _p(v))
    u = np.array([L.orc_rng_uniform(C.byref(r)) for _ in range(ndraw)])
    g = np.array([L.orc_rng_gaussian(C.byref(r)) for _ in range(ndraw)])
    return u0, v.reshape(-1, 3), u, g


def kinetic_energy(sys, v):
    """kineticEnergy (TopologyUtilities.cpp:163-171)"""
    os_ = OracleSystem(sys)
    vv = np.ascontiguousarray(v, np.float64).reshape(-1)
    return lib().orc_kinetic_energy(C.byref(os_.c), _p(vv))


def vacuum_lj_coulomb(c, x=None, block=512):
    """All-pairs LennardJonesForce + CoulombForce in vacuum as
    OneAtomPair<VacuumBoundaryConditions, Universal, ...> sums them
    (OneAtomPair.h:47-107, NonbondedSimpleFullSystemForce.h:125-150): the
    type table's A, B (A14, B14 for a MODIFIED pair, LennardJonesForce.h:44-50)
    or per-atom GROMACS mixing (BuildTopology.cpp:650-659, MODIFIED pairs
    scaled by lj_scaling), exclusions only within a molecule, MODIFIED Coulomb
    scaled by coulomb_scaling (CoulombForce.h:29-40).  Every pair in row
    blocks with the excluded ones skipped, then the MODIFIED pairs in their
    own form; FP64 numpy.  Returns forces (n, 3) and 36 slots (Coulomb,
    LJ; no virial)."""
    x = np.asarray(c.x if x is None else x, np.float64)
    n = c.n
    q = np.asarray(c.charge, np.float64)
    table = c.lj_table is not None
    if table:
        tab = np.asarray(c.lj_table, np.float64)
        ty = np.asarray(c.type)
    else:
        sh, se = 0.5 * np.asarray(c.lj_sigma), np.sqrt(np.asarray(c.lj_epsilon))
    f = np.zeros((n, 3))
    elj = ecou = 0.0

    def coef(i, j, mod):
        if table:
            t = tab[ty[i], ty[j]]
            return (t[..., 2], t[..., 3]) if mod else (t[..., 0], t[..., 1])
        sg = sh[i] + sh[j]
        ep = se[i] * se[j]
        s6 = sg ** 6
        B = 4.0 * ep * s6
        A = B * s6
        return (A * c.lj_scaling, B * c.lj_scaling) if mod else (A, B)

    def terms(i, j, d, mod=False):
        rr = 1.0 / np.sum(d * d, axis=-1)
        A, B = coef(i, j, mod)
        r6 = rr ** 3
        el = A * r6 * r6 - B * r6
        fl = (12 * A * r6 * r6 - 6 * B * r6) * rr
        ec = q[i] * q[j] * np.sqrt(rr)
        return el, fl, ec, ec * rr

    # the excluded pairs (last add wins, one molecule), skipped inside the
    # blocks as OneAtomPair skips them; MODIFIED pairs evaluated on their own
    mol = np.asarray(c.molecule)
    last = {}
    for a, b, cls in c.exclusions:
        if mol[a] == mol[b] and a != b:
            last[(min(a, b), max(a, b))] = cls
    pairs = np.array([(a, b, k) for (a, b), k in last.items()], dtype=np.int64).reshape(-1, 3)
    for b0 in range(0, n, block):
        b1 = min(n, b0 + block)
        i = np.arange(b0, b1)[:, None]
        j = np.arange(n)[None, :]
        d = x[None, :, :] - x[i[:, 0], None, :]
        mask = j > i
        sel = (pairs[:, 0] >= b0) & (pairs[:, 0] < b1)
        mask[pairs[sel, 0] - b0, pairs[sel, 1]] = False
        el, fl, ec, fc = terms(i, j, np.where(mask[..., None], d, 1.0))
        el, ec = np.where(mask, el, 0.0), np.where(mask, ec, 0.0)
        fs = np.where(mask, fl + fc, 0.0)
        fij = d * fs[..., None]
        f[i[:, 0]] -= fij.sum(axis=1)
        f += fij.sum(axis=0)
        elj += el.sum()
        ecou += ec.sum()
    mod = pairs[pairs[:, 2] == 2]
    if len(mod):
        # MODIFIED: LJ with its 1-4 coefficients, Coulomb scaled by coulomb_scaling
        i, j = mod[:, 0], mod[:, 1]
        d = x[j] - x[i]
        el, fl, ec, fc = terms(i, j, d, mod=True)
        ec, fc = ec * c.coulomb_scaling, fc * c.coulomb_scaling
        fij = d * (fl + fc)[:, None]
        np.add.at(f, i, -fij)
        np.add.at(f, j, fij)
        elj += float(np.sum(el))
        ecou += float(np.sum(ec))
    e = np.zeros(36)
    e[0], e[1] = ecou, elj
    return f, e


def pair_hash(lo, hi):
    return lib().orc_pair_hash(lo, hi)
