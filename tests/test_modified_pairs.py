"""Modified (1-4) pairs: ExclusionTable's EXCLUSION_MODIFIED class, whose pairs
take the 1-4 LJ coefficients (LennardJonesParameters::A14 / B14, or
lj_scaling with per-atom parameters) and coulombScalingFactor
(OneAtomPair.h:60-80, LennardJonesForce.h:45-58, CoulombForce.h:29-40,
NonbondedFullEwaldSystemForce.h:818-834 and :1097-1131).

The fixture is a box of united-atom chains whose 1-4 pairs are MODIFIED; its
reference outputs come from the reference's own classes (tests/golden/
make_golden.py, "modified").  The CPU tests pin the oracle to them and to an
identity that does not need them; the GPU tests hold the HIP path to them.
"""
import math

import numpy as np
import pytest

import oracle as O
import pmtest as T
from protomol_amd import _lib as L
from protomol_amd import systems as S

ACC = math.exp(-10.24)
SW = (3, 8.0, 10.0, 0.0, 0.0)  # C2 8-10 A
RTOL_F = 1e-5
RTOL_E = 1e-5


def chains(tag):
    return S.chain_box(512, 52.0, seed=3, per_atom=(tag == "atom"))


def close(a, b, rel=1e-11):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rel * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("tag", ["table", "atom"])
def test_oracle_modified_cutoff_matches_reference(tag):
    g = T.golden("ref_modified.npz")
    c = chains(tag)
    assert np.array_equal(c.x, g[f"x_{tag}"])
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    f, e, cen = O.nonbonded_cutoff(c, O.TERM_LJ, sw, O.TERM_COULOMB, sw, algorithm_cutoff=10.0, cellsize=5.0,
                                   virial=True)
    assert cen == (int(g[f"census_{tag}"][0]), int(g[f"census_{tag}"][1]))
    assert close(f, g[f"f_cut_{tag}"])
    assert close(e, g[f"e_cut_{tag}"])


@pytest.mark.parametrize("tag", ["table", "atom"])
def test_oracle_modified_ewald_matches_reference(tag):
    g = T.golden("ref_modified.npz")
    c = chains(tag)
    f, e, _, _ = O.ewald_real(c, 0.32, ACC, cellsize=5.0, virial=True)
    assert close(f, g[f"f_ew_{tag}"])
    assert close(e, g[f"e_ew_{tag}"])


def test_oracle_modified_with_unit_scales_is_plain_pair():
    # with A14 = A, B14 = B and coulombScalingFactor 1 a MODIFIED pair is an
    # ordinary pair: dropping those exclusions changes nothing (to rounding)
    c = chains("table")
    tab = c.lj_table.copy()
    tab[..., 2:] = tab[..., :2]
    c.lj_table, c.coulomb_scaling = tab, 1.0
    plain = S.System(**{**c.__dict__, "exclusions": c.exclusions[c.exclusions[:, 2] != S.EXCL_MODIFIED]})
    sw = O.make_switch(O.SW_C2, 8.0, 10.0)
    f1, e1, c1 = O.nonbonded_cutoff(c, O.TERM_LJ, sw, O.TERM_COULOMB, sw, virial=True)
    f2, e2, c2 = O.nonbonded_cutoff(plain, O.TERM_LJ, sw, O.TERM_COULOMB, sw, virial=True)
    assert c1 == c2
    assert close(f1, f2, 1e-12) and close(e1, e2, 1e-12)
    # and the scales do act: the default chains differ
    f3, _, _ = O.nonbonded_cutoff(chains("table"), O.TERM_LJ, sw, O.TERM_COULOMB, sw)
    assert not close(f3, f1, 1e-6)


def _desc(ewald):
    d = L.ForceDesc()
    sw = L.switch(*SW)
    if ewald:
        d.term_a = L.TERM_NONE
        d.term_b = L.TERM_EWALD_REAL
        d.ewald_alpha = 0.32
        d.ewald_accuracy = ACC
        d.ewald_correction = 1
    else:
        d.term_a, d.switch_a = L.TERM_LJ, sw
        d.term_b, d.switch_b = L.TERM_COULOMB, sw
    d.cutoff = 10.0
    d.skin = 2.0
    d.rebuild_interval = 10
    d.flags = L.FLAG_VIRIAL
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["table", "atom"])
@pytest.mark.parametrize("model", ["cutoff", "ewald"])
def test_modified_pairs_vs_reference_golden(tag, model):
    print(f"case modified/{tag}/{model}", flush=True)
    g = T.golden("ref_modified.npz")
    c = chains(tag)
    ref_f, ref_e = g[f"f_{'cut' if model == 'cutoff' else 'ew'}_{tag}"], g[f"e_{'cut' if model == 'cutoff' else 'ew'}_{tag}"]
    with L.Context(_desc(model == "ewald"), [c]) as ctx:
        f = np.zeros_like(c.x)
        e = np.zeros(L.NSLOTS)
        ctx.evaluate(0, c.x, f, e)
        cen = ctx.census(0, c.x)
    if model == "cutoff":
        assert cen == (int(g[f"census_{tag}"][0]), int(g[f"census_{tag}"][1]))
    assert T.rms_rel(f, ref_f) <= RTOL_F
    for s in (0, 1):
        tol = RTOL_E * max(1.0, abs(ref_e[s])) + 1e-7 * cen[0]
        assert abs(e[s] - ref_e[s]) <= tol, (s, e[s], ref_e[s], tol)
    vtol = RTOL_E * max(1.0, np.max(np.abs(ref_e[7:16]))) + 1e-7 * cen[0]
    assert np.max(np.abs(e[7:16] - ref_e[7:16])) <= vtol


GB_PARAMS = dict(offset=0.09, a=1.0, b=0.8, g=4.85, eps_in=1.0, eps_out=78.3)


def gb_chain_modified(n, seed):
    # the chain of make_golden.gb_chain_modified: 1-4 pairs MODIFIED,
    # lj_scaling 0.5, coulomb_scaling 1/1.2
    g = S.gb_chain(n, seed=seed)
    ex14 = np.stack([np.arange(n - 3), np.arange(3, n), np.full(n - 3, S.EXCL_MODIFIED)], 1)
    g.exclusions = np.concatenate([g.exclusions, ex14]).astype(np.int32)
    g.lj_scaling, g.coulomb_scaling = 0.5, 1.0 / 1.2
    return g


def test_oracle_gb_modified_matches_reference():
    # GB (GBBornRadii, GBPartialSum, GBForce) plus the scaled LJ / Coulomb of
    # the MODIFIED pairs against the reference's classes
    g = T.golden("ref_gb_modified.npz")
    c = gb_chain_modified(int(g["n"]), int(g["seed"]))
    assert np.array_equal(c.x, g["x"])
    f, e, _ = O.gb_evaluate(c, **GB_PARAMS)
    fl, el = T.vacuum_lj_coulomb(c)
    assert close(f + fl, g["f"], rel=1e-11)
    assert close(e + el, g["e"], rel=1e-11)
    # the blocked all-pairs sum (MODIFIED pairs rescaled after the fact)
    fb, eb = O.vacuum_lj_coulomb(c, block=29)
    assert close(f + fb, g["f"], rel=1e-11)
    assert close(e[:2] + eb[:2], g["e"][:2], rel=1e-11)


def gb_chain_table(n, seed):
    # the chain of make_golden.gb_chain_table: a 4-type LJ table with its own
    # A14 / B14 in place of per-atom LJ
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


@pytest.mark.gpu
@pytest.mark.parametrize("lj", ["atom", "table"])
def test_gb_modified_pairs_vs_reference_golden(lj):
    print(f"case modified/gb/{lj}", flush=True)
    g = T.golden("ref_gb_modified.npz")
    make = gb_chain_modified if lj == "atom" else gb_chain_table
    c = make(int(g["n"]), int(g["seed"]))
    ref_f, ref_e = (g["f"], g["e"]) if lj == "atom" else (g["f_table"], g["e_table"])
    d = L.ForceDesc()
    d.gb_offset, d.gb_alpha, d.gb_beta, d.gb_gamma = 0.09, 1.0, 0.8, 4.85
    d.gb_solute_dielec, d.gb_solvent_dielec = 1.0, 78.3
    d.flags = L.FLAG_GB
    with L.Context(d, [c]) as ctx:
        f = np.zeros_like(c.x)
        e = np.zeros(L.NSLOTS)
        ctx.evaluate(0, c.x, f, e)
    assert T.rms_rel(f, ref_f) <= RTOL_F
    npairs = 3 * c.n * (c.n - 1) // 2
    for s in (0, 1):
        tol = RTOL_E * max(1.0, abs(ref_e[s])) + 1e-7 * npairs
        assert abs(e[s] - ref_e[s]) <= tol, (s, e[s], ref_e[s], tol)


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["table", "atom"])
def test_modified_pairs_md_forces_vs_reference_golden(tag):
    # the MD path's forces-only pair kernel (F(x0) of pm_md_set_state) on the
    # same MODIFIED-pair chains
    print(f"case modified/md/{tag}", flush=True)
    g = T.golden("ref_modified.npz")
    c = chains(tag)
    with L.Context(_desc(False), [c]) as ctx:
        ctx.md_set_state(c.x, np.zeros_like(c.x))
        _, _, f, _ = ctx.md_get_state()
    assert T.rms_rel(f, g[f"f_cut_{tag}"]) <= RTOL_F
