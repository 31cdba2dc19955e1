"""The reference's GROMACS/AMBER implicit-solvent regressions
(test/tests/alanylalanine_FFAMBER_GB_LL.conf, www-fip35_FFAMBER_GB_LL.conf)
at their start positions: the topology the reference's parsers build from
the PDB, the .top with its ffamber parameters and the GB parameter file
(charges, masses, 7 / 28 LJ types with 1-4 coefficients, FULL and MODIFIED
exclusions, GB radii and scales), and the forces and energies of each term
of the GB group (GBBornRadii -> GBPartialSum -> GBForce, GBACEForce,
LennardJones, Coulomb; NonbondedSimpleFull, vacuum) and of the harmonic
bonds and angles, written by the reference's own classes
(tests/golden/ref_gb_ll_*.npz, ref_driver mode 7).  The dihedral terms of
those runs are not on this path.  On the device: one evaluation against the
reference, and 20 Langevin-leapfrog steps at T = 0 (the runs' integrator,
1 fs, their gamma) against the FP64 oracle's integration of the same
forces."""
import math

import numpy as np
import pytest

import oracle as O
import pmtest as T
from protomol_amd import _lib as L
from protomol_amd import systems as S

CASES = ("alanylalanine", "wwwfip35")
GB_PARAMS = dict(offset=0.09, a=1.0, b=0.8, g=4.85, eps_in=1.0, eps_out=78.3)


GAMMA = {"alanylalanine": 91.0, "wwwfip35": 80.0}   # the runs' Langevin gamma (1/ps)


def gb_ll_system(name):
    g = T.golden(f"ref_gb_ll_{name}.npz")
    n = g["charge"].size
    assert np.all(g["gb_offset"] == 0.09)
    assert np.all(g["angles"][:, 5:] == 0.0)   # no Urey-Bradley terms
    b, a = g["bonds"], g["angles"]
    return S.System(x=g["x"], charge=g["charge"], mass=g["mass"], molecule=g["molecule"], box=np.zeros(3),
                    origin=np.zeros(3), periodic=False, type=g["type"], lj_table=g["lj_table"],
                    exclusions=g["exclusions"], coulomb_scaling=float(g["coulomb_scaling"]),
                    lj_scaling=float(g["lj_scaling"]), gb_radius=g["gb_radius"], gb_scaling=g["gb_scaling"],
                    bonds=b[:, :2].astype(np.int32), bond_k=b[:, 2], bond_r0=b[:, 3],
                    angles=a[:, :3].astype(np.int32), angle_k=a[:, 3], angle_t0=a[:, 4],
                    name=f"{name} ({n} atoms)"), g


def ref_total(g):
    f = g["f_gb"] + g["f_ace"] + g["f_lj"] + g["f_coul"] + g["f_bonded"]
    e = g["e_gb"] + g["e_ace"] + g["e_lj"] + g["e_coul"] + g["e_bonded"]
    return f, e


def oracle_forces(s, x):
    f, e, _ = O.gb_evaluate(s, ace=True, x=x, **GB_PARAMS)
    fl, el = O.vacuum_lj_coulomb(s, x=x)
    fb, eb = O.bonded(s, x=x)
    return f + fl + fb, e + el + eb


def gb_desc(bonded=True):
    d = L.ForceDesc()
    d.gb_offset, d.gb_alpha, d.gb_beta, d.gb_gamma = 0.09, 1.0, 0.8, 4.85
    d.gb_solute_dielec, d.gb_solvent_dielec = 1.0, 78.3
    d.flags = L.FLAG_GB | L.FLAG_GB_ACE | (L.FLAG_BONDED if bonded else 0)
    return d


def close(a, b, rel):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b))) <= rel * max(1.0, float(np.max(np.abs(b))))


@pytest.mark.parametrize("name", CASES)
def test_oracle_gb_ll_vs_reference(name):
    s, g = gb_ll_system(name)
    f, e, born = O.gb_evaluate(s, ace=True, virial=True, **GB_PARAMS)
    assert close(born, g["born"], 1e-12)
    assert close(f, g["f_gb"] + g["f_ace"], 1e-11)
    assert close(e, g["e_gb"] + g["e_ace"], 1e-11)
    fl, el = O.vacuum_lj_coulomb(s)
    assert close(fl, g["f_lj"] + g["f_coul"], 1e-11)
    assert abs(el[0] - g["e_coul"][0]) <= 1e-11 * abs(g["e_coul"][0])
    assert abs(el[1] - g["e_lj"][1]) <= 1e-11 * abs(g["e_lj"][1])
    fb, eb = O.bonded(s, virial=True)
    assert close(fb, g["f_bonded"], 1e-11)
    assert close(eb, g["e_bonded"], 1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gb_ll_evaluate_vs_reference(name):
    """The whole force set of the run but its dihedrals, one evaluation on the
    device (FP32 pair arithmetic, FP64 sums) against the reference."""
    print(f"case gb_ll evaluate {name}", flush=True)
    s, g = gb_ll_system(name)
    ref_f, ref_e = ref_total(g)
    with L.Context(gb_desc(), [s]) as ctx:
        f = np.zeros_like(s.x)
        e = np.zeros(L.NSLOTS)
        ctx.evaluate(0, s.x, f, e)
    assert T.rms_rel(f, ref_f) <= 1e-5
    # per atom, against the force scale of the system
    assert np.max(np.linalg.norm(f - ref_f, axis=1)) <= 1e-4 * np.max(np.linalg.norm(ref_f, axis=1))
    npairs = s.n * (s.n - 1) // 2
    for slot in (0, 1, 2, 3):   # Coulomb + GB + ACE, LJ, bonds, angles
        tol = 1e-5 * max(1.0, abs(ref_e[slot])) + 1e-7 * npairs
        assert abs(e[slot] - ref_e[slot]) <= tol, (slot, e[slot], ref_e[slot], tol)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gb_ll_langevin_zero_temperature_vs_oracle(name):
    """20 LangevinLeapfrog steps at 1 fs with the run's gamma and T = 0 (the
    noise term off, the damping and force coupling on) from the start
    positions at rest: the device run against the FP64 oracle's forces and
    the numpy LangevinLeapfrogIntegrator (LangevinLeapfrogIntegrator.cpp:46-70)."""
    print(f"case gb_ll langevin {name}", flush=True)
    s, g = gb_ll_system(name)
    v0 = np.zeros_like(s.x)
    steps, dt, gamma = 20, 1.0, GAMMA[name]
    xo, vo, Fo, _ = T.langevin_oracle(lambda x: oracle_forces(s, x), s.x.copy(), v0.copy(), s.mass, dt, steps, 0.0,
                                      gamma, seed=1234)
    with L.Context(gb_desc(), [s]) as ctx:
        ctx.set_langevin(0.0, gamma, 1234)
        ctx.md_set_state(s.x, v0)
        ctx.md_run(steps, dt, energies=False)
        x, v, F, _ = ctx.md_get_state()
    assert np.max(np.abs(x - xo)) < 1e-5
    assert T.rms_rel(v, vo) < 1e-4
    assert T.rms_rel(F, Fo) < 1e-5
