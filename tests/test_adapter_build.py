"""The SystemForce adapter (include/protomol_amd_systemforce.hpp) compiles as
C++98 against ProtoMol's own headers and links against the reference's
objects and libprotomol_amd.so: the drop-in boundary type-checks.  Needs the
reference sources and oracle/_ref (built by `make -C oracle ref`), so it runs
only in a container that has them."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
ARCHIVE = os.path.join(ROOT, "oracle", "_ref", "libprotomol_ref.a")
LIB = os.path.join(ROOT, "protomol_amd", "libprotomol_amd.so")

SRC = r'''
#include "protomol_amd_systemforce.hpp"
#include <protomol/topology/Topology.h>
#include <protomol/topology/CubicCellManager.h>
using namespace ProtoMol;
int main() {
  pm_force_desc d = pm_force_desc();
  d.term_a = PM_TERM_LJ;
  NonbondedCutoffSystemForceHIP f(d);
  NonbondedEwaldRealSystemForceHIP g(0.32, 3.5713e-5, true);
  SystemForce *p = &f;
  SystemForce *q = &g;
  return (p->getKeyword() == "NonbondedCutoffHIP" && q->numberOfBlocks(0, 0) == 1) ? 0 : 1;
}
'''


@pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "protomol")) and os.path.exists(ARCHIVE)
                         and os.path.exists(LIB)), reason="needs the reference sources and oracle/_ref")
def test_adapter_compiles_and_links_against_reference():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.cpp")
        open(src, "w").write(SRC)
        out = os.path.join(d, "probe")
        cmd = ["g++", "-std=c++98", "-w", f"-I{REF}", f"-I{ROOT}/include", src, ARCHIVE,
               f"-L{ROOT}/protomol_amd", "-lprotomol_amd", "-o", out]
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-2000:]
