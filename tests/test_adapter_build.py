"""The SystemForce adapters (include/protomol_amd_systemforce.hpp) inside
ProtoMol: the header compiles as C++98 against ProtoMol's own headers; made
through ProtoMol's ForceFactory from configuration strings (the reference's
own parameter parsing and make()), they hand the C ABI the force and system
records those strings and the topology describe.  oracle/_ref/adapter_check_stub
links a recording stand-in of the C ABI, so this runs on a CPU; the GPU
build is held to the reference's forces in test_gpu_adapter.py.

Needs the reference sources and oracle/_ref (`make -C oracle ref adapter`),
so it runs only in a container that has them."""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pmtest as T
from protomol_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
ARCHIVE = os.path.join(ROOT, "oracle", "_ref", "libprotomol_ref.a")
STUB = os.path.join(ROOT, "oracle", "_ref", "adapter_check_stub")
LIB = os.path.join(ROOT, "protomol_amd", "libprotomol_amd.so")
HAVE_REF = os.path.isdir(os.path.join(REF, "protomol")) and os.path.exists(ARCHIVE)

SRC = r'''
#include "protomol_amd_systemforce.hpp"
#include <protomol/topology/Topology.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/switch/C2SwitchingFunction.h>
#include <protomol/switch/C1SwitchingFunction.h>
using namespace ProtoMol;
typedef OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction, LennardJonesForce,
                       C1SwitchingFunction, CoulombForce> P;
int main() {
  NonbondedCutoffSystemForceHIP<P> f;
  NonbondedFullEwaldSystemForceHIP<true, false, true> g;
  NonbondedSimpleFullSystemForceHIP<GBForce> h;
  SystemForce *p = &f, *q = &g, *r = &h;
  return (p->getKeyword() == "NonbondedCutoffHIP" && q->getKeyword() == "FullEwaldHIP" &&
          r->getKeyword() == "NonbondedSimpleFullHIP" && q->numberOfBlocks(0, 0) == 1) ? 0 : 1;
}
'''


@pytest.mark.skipif(not (HAVE_REF and os.path.exists(LIB)), reason="needs the reference sources and oracle/_ref")
def test_adapter_compiles_and_links_against_reference():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.cpp")
        open(src, "w").write(SRC)
        out = os.path.join(d, "probe")
        cmd = ["g++", "-std=c++98", "-w", f"-I{REF}", f"-I{ROOT}/include", src, ARCHIVE,
               f"-L{ROOT}/protomol_amd", "-lprotomol_amd", "-o", out]
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-2000:]


def _run_stub(records):
    import refio
    with tempfile.TemporaryDirectory() as d:
        log = os.path.join(d, "stub.log")
        fin, fout = os.path.join(d, "in.rec"), os.path.join(d, "out.rec")
        refio.write_records(fin, records)
        p = subprocess.run([STUB, fin, fout], env=dict(os.environ, PM_STUB_LOG=log), capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-2000:]
        creates = []
        for ln in open(log):
            kv = dict(t.split("=", 1) for t in ln.split()[1:])
            creates.append(kv)
        return creates, refio.read_records(fout)


@pytest.mark.skipif(not (HAVE_REF and os.path.exists(STUB)), reason="needs oracle/_ref/adapter_check_stub")
def test_factory_strings_become_the_c_abi_records():
    """ForceFactory::make of each configuration line (oracle/adapter_check.cpp)
    hands pm_ctx_create the terms, switches and cutoffs of that line and the
    topology's atoms, LJ table, exclusions and box; evaluate adds into the
    caller's forces and energy slots."""
    import refio
    w = T.fixture_water()
    r = refio.system_records(w)
    r.update(mode=np.int32(1), cellsize=5.0)
    creates, out = _run_stub(r)
    assert int(out["cases"][0]) == 7 and len(creates) == 7
    f = lambda kv, k: float(kv[k])
    c2c1, c2c2, lj, coul, cncn, ewr, ewf = creates
    # OneAtomPairTwo<PBC, C2, LJ, C1, Coulomb>: -switchon 0.1 -cutoff 9 -cutoff 9, algorithm -cutoff 9
    assert (int(c2c1["term_a"]), int(c2c1["term_b"])) == (1, 2)
    assert (int(c2c1["sw_a.kind"]), f(c2c1, "sw_a.switchon"), f(c2c1, "sw_a.cutoff")) == (3, 0.1, 9.0)
    assert (int(c2c1["sw_b.kind"]), f(c2c1, "sw_b.cutoff"), f(c2c1, "cutoff")) == (2, 9.0, 9.0)
    assert (int(c2c2["sw_a.kind"]), int(c2c2["sw_b.kind"]), f(c2c2, "sw_b.switchon")) == (3, 3, 8.0)
    assert (int(lj["term_a"]), int(lj["term_b"]), f(lj, "sw_a.switchon"), f(lj, "sw_a.cutoff")) == (1, 0, 6.0, 8.0)
    assert (int(coul["term_a"]), int(coul["sw_a.kind"]), f(coul, "sw_a.cutoff")) == (2, 2, 10.0)
    assert (int(cncn["sw_a.kind"]), f(cncn, "sw_a.order"), f(cncn, "sw_a.switchon")) == (4, 6.0, 8.0)
    # Ewald: -alpha 0.32 -accuracy 3.5713e-05, correction on, reciprocal only in the second
    for kv, recip in ((ewr, 0), (ewf, 16)):
        assert (int(kv["term_a"]), int(kv["term_b"])) == (0, 3)
        assert f(kv, "alpha") == 0.32 and f(kv, "accuracy") == 3.5713e-05 and int(kv["correction"]) == 1
        assert int(kv["flags"]) & 16 == recip
    # the ScalarStructure's virial() switch becomes PM_FLAG_VIRIAL
    assert all(int(kv["flags"]) & 1 for kv in creates)
    # the topology: atoms, types, LJ table, exclusions (add order), box
    for kv in creates:
        assert int(kv["n_atoms"]) == w.n and int(kv["n_types"]) == 2 and int(kv["periodic"]) == 1
        assert kv["box"] == "22.5,22.5,22.5" and int(kv["n_excl"]) == len(w.exclusions)
        assert f(kv, "charge_w") == pytest.approx(float(np.sum(w.charge * np.arange(1, w.n + 1))), rel=1e-12)
        assert f(kv, "lj_sum") == pytest.approx(float(np.sum(w.lj_table_full())), rel=1e-12)
        ex = w.exclusions
        assert int(kv["excl_w"]) == int(np.sum(ex[:, 0].astype(np.int64) * 7 + ex[:, 1] * 3 + ex[:, 2]))
    # evaluate adds: the stand-in adds 1 to every component, 0.5 to the slots
    assert np.all(out["f_hip0"] == 1.0) and out["e_hip0"][0] == 0.5 and out["e_hip0"][1] == 0.5


@pytest.mark.skipif(not (HAVE_REF and os.path.exists(STUB)), reason="needs oracle/_ref/adapter_check_stub")
def test_full_group_is_one_device_evaluation():
    """The all-pairs lines of a GB force group (LennardJones, Coulomb,
    GBBornRadii, GBPartialSum, GBForce -soluteDielec 1 -solventDielec 78.3,
    GBACEForce) become one context: all-pairs, GB, ACE, LJ and Coulomb, the
    topology's OBC constants and the atoms' GBSA parameters; the group's
    forces are added once."""
    import refio
    c = S.gb_chain(120, seed=4)
    params = np.array([0.09, 1.0, 0.8, 4.85, 1.0, 78.3])
    r = refio.system_records(c)
    r.update(mode=np.int32(3), gb_rho=c.gb_radius, gb_S=c.gb_scaling, gb_params=params, cellsize=np.int32(5))
    creates, out = _run_stub(r)
    assert len(creates) == 1
    kv = creates[0]
    assert int(kv["flags"]) & (128 | 2 | 32) == 128 | 2 | 32
    assert (int(kv["term_a"]), int(kv["term_b"])) == (1, 2)
    assert float(kv["solute"]) == 1.0 and float(kv["solvent"]) == 78.3
    assert float(kv["ace_sigma"]) == pytest.approx(2.26 / 418.4) and float(kv["ace_rho_s"]) == 1.4
    assert (float(kv["gb_offset"]), float(kv["gb_alpha"]), float(kv["gb_beta"]), float(kv["gb_gamma"])) == \
        (0.09, 1.0, 0.8, 4.85)
    assert int(kv["periodic"]) == 0 and int(kv["n_atoms"]) == c.n
    assert float(kv["gb_r"]) == pytest.approx(float(np.sum(c.gb_radius)), rel=1e-12)
    assert float(kv["gb_s"]) == pytest.approx(float(np.sum(c.gb_scaling)), rel=1e-12)
    assert np.all(out["f_hip0"] == 1.0)


@pytest.mark.skipif(not (HAVE_REF and os.path.exists(STUB)), reason="needs oracle/_ref/adapter_check_stub")
def test_vacuum_cutoff_region_and_rebuild():
    """A vacuum topology on NonbondedCutoffHIP: the context's region is the
    positions' bounding box grown by 10 A; when an atom leaves it (the C ABI
    answers PM_ECAPACITY and adds nothing) the adapter makes the context
    again around the positions of that evaluation and adds their forces once."""
    import refio
    d = S.water_droplet(14.0, seed=3)
    r = refio.system_records(d)
    exit_ = 15.0
    r.update(mode=np.int32(6), cellsize=5.0, exit=exit_)
    creates, out = _run_stub(r)
    assert int(out["cases"][0]) == 4 and len(creates) == 4  # 2 lines x (first region, rebuilt region)
    far = int(np.argmax(d.x[:, 0]))
    moved = d.x.copy()
    moved[d.molecule == d.molecule[far], 0] += exit_
    for kv, x in zip(creates, (d.x, moved, d.x, moved)):
        lo, hi = x.min(axis=0), x.max(axis=0)
        box = np.array([float(v) for v in kv["box"].split(",")])
        org = np.array([float(v) for v in kv["origin"].split(",")])
        assert int(kv["periodic"]) == 0
        assert np.allclose(box, hi - lo + 20.0, rtol=0, atol=1e-12)
        assert np.allclose(org, 0.5 * (lo + hi), rtol=0, atol=1e-12)
    assert (int(creates[0]["term_a"]), int(creates[0]["term_b"]), float(creates[0]["sw_a.cutoff"])) == (1, 2, 8.0)
    assert (int(creates[2]["term_a"]), int(creates[2]["term_b"]), float(creates[2]["cutoff"])) == (2, 0, 10.0)
    for k in range(4):
        assert np.all(out[f"f_hip{k}"] == 1.0) and out[f"e_hip{k}"][0] == 0.5
