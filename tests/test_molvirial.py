"""The molecular virial (ScalarStructure::molecularVirial) and the atomic
virial of the bonded and GB terms: the oracle against the reference's own
classes, and the identities that fix the molecular virial without them."""
import math

import numpy as np

import oracle as O
import pmtest as T
from protomol_amd import systems as S

ACC = math.exp(-10.24)
VIR = slice(7, 16)
MOLVIR = slice(16, 25)


def close(a, b, rel=1e-11):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rel * max(1.0, np.max(np.abs(b)))


def gb_multi_chain(n=200, seed=17, per=50):
    g = S.gb_chain(n, seed=seed)
    g.molecule = (np.arange(n) // per).astype(np.int32)
    return g


def test_molecule_com():
    # molecularCenterOfMass: mass-weighted mean of a molecule's atoms
    w = T.fixture_water()
    com = O.molecule_com(w)
    m = w.mass.reshape(-1, 3)
    ref = np.einsum("ma,mak->mk", m, w.x.reshape(-1, 3, 3)) / m.sum(1)[:, None]
    assert com.shape == (w.n // 3, 3)
    assert np.max(np.abs(com - ref)) < 1e-12


def test_molvirial_cutoff_matches_reference():
    g = T.golden("ref_molvirial.npz")
    w = T.fixture_water()
    assert np.array_equal(w.x, g["x"])
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    for vir in (1, 0):
        f, e, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw, algorithm_cutoff=10.0, cellsize=5.0,
                                     virial=bool(vir), molvirial=True)
        assert close(f, g[f"f_cut_v{vir}"])
        assert close(e, g[f"e_cut_v{vir}"])
        assert np.any(e[MOLVIR] != 0)


def test_molvirial_ewald_matches_reference():
    g = T.golden("ref_molvirial.npz")
    w = T.fixture_water()
    for vir in (1, 0):
        fr, er, _, _ = O.ewald_real(w, 0.32, ACC, virial=bool(vir), molvirial=True)
        fk, ek = O.ewald_reciprocal(w, 0.32, ACC, virial=bool(vir), molvirial=True)
        assert close(fk, g[f"f_rec_v{vir}"])
        assert close(ek, g[f"e_rec_v{vir}"])
        assert close(fr + fk, g[f"f_ew_v{vir}"])
        assert close(er + ek, g[f"e_ew_v{vir}"])


def test_bonded_virial_matches_reference():
    g = T.golden("ref_molvirial.npz")
    w = T.fixture_water()
    _, e = O.bonded(w, virial=True)
    assert close(e, g["e_bonded_v1"], rel=1e-13)


def test_gb_virials_match_reference():
    g = T.golden("ref_molvirial.npz")
    c = gb_multi_chain(200, int(g["gb_seed"]), int(g["gb_per"]))
    assert np.array_equal(c.x, g["gb_x"])
    p = g["gb_params"]
    for vir in (1, 0):
        f, e, _ = O.gb_evaluate(c, offset=p[0], a=p[1], b=p[2], g=p[3], eps_in=p[4], eps_out=p[5],
                                virial=bool(vir), molvirial=True)
        fl, el = T.vacuum_lj_coulomb(c, virial=bool(vir), molvirial=True)
        assert close(f + fl, g[f"f_gb_v{vir}"], rel=1e-9)
        assert close(e + el, g[f"e_gb_v{vir}"], rel=1e-9)


def test_molvirial_single_atom_molecules():
    # every atom its own molecule: the centre is the atom, so the molecular
    # virial is the atomic one
    w = T.fixture_water()
    w.molecule = np.arange(w.n, dtype=np.int32)
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    _, e, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw, virial=True, molvirial=True)
    assert close(e[MOLVIR], e[VIR], rel=1e-12)
    _, e, _, _ = O.ewald_real(w, 0.32, ACC, correction=False, virial=True, molvirial=True)
    assert close(e[MOLVIR], e[VIR], rel=1e-12)


UPPER = [0, 1, 2, 4, 5, 8]  # xx xy xz yy yz zz of a row-major 3x3


def _intermolecular_identity(w, f, e):
    # W_mol = W_atomic - sum_a F_a (x) u_a, u_a = x_a - R_M(a) (minimal image),
    # when every evaluated pair joins two molecules; the reference keeps the
    # upper triangle (addVirial copies it into the lower one)
    com = O.molecule_com(w)
    u = w.x - com[w.molecule]
    u -= w.box * np.rint(u / w.box)
    m = e[MOLVIR]
    assert m[3] == m[1] and m[6] == m[2] and m[7] == m[5]
    return (e[VIR] - np.einsum("ai,aj->ij", f, u).reshape(-1))[UPPER]


def test_molvirial_identity_cutoff():
    w = T.fixture_water()
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    f, e, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw, virial=True, molvirial=True)
    assert close(e[MOLVIR][UPPER], _intermolecular_identity(w, f, e), rel=1e-10)


def test_molvirial_identity_ewald_real():
    # the real-space sum without the correction term (whose pairs lie inside
    # molecules and add to the atomic virial only)
    w = T.fixture_water()
    f, e, _, _ = O.ewald_real(w, 0.32, ACC, correction=False, virial=True, molvirial=True)
    assert close(e[MOLVIR][UPPER], _intermolecular_identity(w, f, e), rel=1e-10)


def test_molvirial_off_leaves_slots():
    w = T.fixture_water()
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    _, e, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw, virial=True)
    assert np.all(e[MOLVIR] == 0)
    _, e, _ = O.nonbonded_cutoff(w, O.TERM_LJ, sw, O.TERM_COULOMB, sw)
    assert np.all(e[7:] == 0)
