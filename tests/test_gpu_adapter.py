"""The SystemForce adapters on the device, beside the reference's own forces
in one program (oracle/_ref/adapter_check, built from the reference sources
and linked to libprotomol_amd.so): each configuration line made through
ProtoMol's ForceFactory twice -- the reference's algorithm and the adapter's
-- on one topology built with ProtoMol's classes, evaluated as a force group
does.  Forces to 1e-5 RMS relative, energies to the pair tolerance."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pmtest as T
from protomol_amd import systems as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
NAMES = ["LJ C2 + Coulomb C1", "LJ C2 + Coulomb C2", "LJ C2 6-8 A", "Coulomb C1", "LJ Cn + Coulomb Cn",
         "Ewald real + correction", "full Ewald (CoulombEwald)"]


def _run(records):
    import refio
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.rec"), os.path.join(d, "out.rec")
        refio.write_records(fin, records)
        p = subprocess.run([CHECK, fin, fout], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr[-2000:]
        return refio.read_records(fout)


def _check(out, k, name, npairs):
    fr, fh = out[f"f_ref{k}"].reshape(-1, 3), out[f"f_hip{k}"].reshape(-1, 3)
    er, eh = out[f"e_ref{k}"], out[f"e_hip{k}"]
    rms = T.rms_rel(fh, fr)
    print(f"  {name}: force RMS relative {rms:.2e}, energies {eh[0]:.6f} {eh[1]:.6f} {eh[6]:.6f} vs {er[0]:.6f} {er[1]:.6f} {er[6]:.6f}",
          flush=True)
    assert rms <= 1e-5, (name, rms)
    for s in range(7):  # COULOMB, LENNARDJONES, ..., OTHER (the GB energy)
        assert abs(eh[s] - er[s]) <= 1e-5 * max(1.0, abs(er[s])) + 1e-7 * npairs, (name, s, eh[s], er[s])
    tol = 1e-5 * max(1.0, np.max(np.abs(er[7:16]))) + 1e-7 * npairs
    assert np.max(np.abs(eh[7:16] - er[7:16])) <= tol, (name, eh[7:16], er[7:16])


@pytest.mark.skipif(not os.path.exists(CHECK), reason="oracle/_ref/adapter_check is built where the reference is")
def test_periodic_adapters_vs_reference_forces():
    import refio
    print("case adapters periodic (water 1,029 atoms)", flush=True)
    w = T.fixture_water()
    r = refio.system_records(w)
    r.update(mode=np.int32(1), cellsize=5.0)
    out = _run(r)
    assert int(out["cases"][0]) == len(NAMES)
    for k, name in enumerate(NAMES):
        _check(out, k, name, 200000)


@pytest.mark.skipif(not os.path.exists(CHECK), reason="oracle/_ref/adapter_check is built where the reference is")
def test_gb_force_group_adapters_vs_reference_forces():
    import refio
    print("case adapters GB force group (chain 300 atoms)", flush=True)
    c = S.gb_chain(300, seed=7)
    r = refio.system_records(c)
    r.update(mode=np.int32(3), gb_rho=c.gb_radius, gb_S=c.gb_scaling,
             gb_params=np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3]), cellsize=np.int32(5))
    out = _run(r)
    _check(out, 0, "GB + ACE + LJ + Coulomb, all pairs", 3 * 300 * 299 // 2)


@pytest.mark.skipif(not os.path.exists(CHECK), reason="oracle/_ref/adapter_check is built where the reference is")
def test_vacuum_cutoff_adapters_vs_reference_forces():
    """OneAtomPair[Two]<VacuumBoundaryConditions, ...> on NonbondedCutoffHIP:
    a water droplet, then the same droplet with one molecule moved out of the
    context's region (the adapter makes its context again)."""
    import refio
    print("case adapters vacuum cutoff (water droplet 1,164 atoms)", flush=True)
    d = S.water_droplet(14.0, seed=3)
    r = refio.system_records(d)
    r.update(mode=np.int32(6), cellsize=5.0, exit=15.0)
    out = _run(r)
    assert int(out["cases"][0]) == 4
    for k, name in enumerate(["LJ C2 6-8 A + Coulomb C2 8-10 A", "the same, one molecule out of the region",
                              "Coulomb C1", "Coulomb C1, one molecule out of the region"]):
        _check(out, k, name, 150000)
