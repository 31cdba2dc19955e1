"""Output formats: the reference's own expected files of its water
regression test (water_CHARMM_PERIODIC_L: XYZ forces trajectory, final XYZ
positions and velocities, allEnergiesFile and its header, DCD) are parsed and
written again through protomol_amd.io; the bytes must come out the same
(for the DCD, all but the title text, which names the file and the build)."""
import os

import numpy as np

from protomol_amd import io as pio

EXP = os.path.join(os.path.dirname(__file__), "golden", "water423_expected")
BASE = "water_CHARMM_PERIODIC_L"


def _path(ext):
    return os.path.join(EXP, f"{BASE}.{ext}")


def test_xyz_final_positions_and_velocities_roundtrip(tmp_path):
    for ext in ("pos", "vel"):
        (names, x), = pio.read_xyz_frames(_path(ext))
        out = tmp_path / f"out.{ext}"
        pio.write_xyz(str(out), x, names)
        assert out.read_bytes() == open(_path(ext), "rb").read()


def test_xyz_force_trajectory_roundtrip(tmp_path):
    frames = pio.read_xyz_frames(_path("forces"))
    assert len(frames) == 2 and len(frames[0][0]) == 423
    out = tmp_path / "out.forces"
    w = pio.XYZTrajectoryWriter(str(out))
    for names, f in frames:
        w.write(f, names)
    assert out.read_bytes() == open(_path("forces"), "rb").read()


def test_energies_roundtrip(tmp_path):
    rows = [[float(v) for v in ln.split()] for ln in open(_path("energy")) if ln.strip()]
    out = tmp_path / "out.energy"
    w = pio.EnergiesWriter(str(out))
    for r in rows:
        w.write(*r[:13], shadow=r[13])
    assert out.read_bytes() == open(_path("energy"), "rb").read()
    assert (tmp_path / "out.energy.header").read_bytes() == open(_path("energy.header"), "rb").read()


def test_dcd_roundtrip(tmp_path):
    hdr, frames = pio.read_dcd(_path("dcd"))
    assert hdr["sets"] == 2 and hdr["atoms"] == 423
    out = tmp_path / "out.dcd"
    w = pio.DCDWriter(str(out), timestep_fs=1.0, first_step=hdr["first_step"])
    for f in frames:
        w.write(f.astype(np.float64))
    a, b = out.read_bytes(), open(_path("dcd"), "rb").read()
    assert len(a) == len(b)
    # title record text (bytes 100 .. 260) names the file and the build date
    assert a[:100] == b[:100] and a[260:] == b[260:]
