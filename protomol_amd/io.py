"""Output in ProtoMol's file formats, for positions, velocities and forces
read back from the device and for the energies of a run.

  write_xyz            XYZWriter::write (io/XYZWriter.cpp:35-58): a count line,
                       then one line per atom: name, TAB, three fields of
                       width 24 with 15 significant digits (finXYZPosFile,
                       finXYZVelFile)
  XYZTrajectoryWriter  XYZTrajectoryWriter (io/XYZTrajectoryWriter.cpp:52-135):
                       the frame count on the first line (padded to 19
                       characters, rewritten on every frame), then frames as
                       above (XYZForceFile, XYZPosFile, XYZVelFile)
  EnergiesWriter       OutputEnergies (output/OutputEnergies.cpp:41-106): the
                       column header file and one line per output step
  DCDWriter            DCDTrajectoryWriter (io/DCDTrajectoryWriter.cpp:74-345):
                       CHARMM-style little-endian DCD, FP32 frames

The numbers are formatted as the reference's C++ streams format them: a
stream with precision p and no fixed/scientific flag prints like printf's
"%.pg", so a width-w field is "%w.pg".
"""
from __future__ import annotations

import os
import struct

import numpy as np

INV_TIMEFACTOR = 0.02045482828087295  # PMConstants.cpp:54


def _xyz_lines(coords, names):
    c = np.asarray(coords, np.float64).reshape(-1, 3)
    if len(names) != len(c):
        raise ValueError("coordinate and atom name counts differ")
    out = [f"{len(c)}\n"]
    for nm, (x, y, z) in zip(names, c):
        out.append(f"{nm}\t{x:24.15g}{y:24.15g}{z:24.15g}\n")
    return "".join(out)


def write_xyz(path, coords, names):
    """One frame, as XYZWriter::write."""
    with open(path, "w") as fh:
        fh.write(_xyz_lines(coords, names))


class XYZTrajectoryWriter:
    """Frames appended to one file; the first line holds the frame count."""

    def __init__(self, path):
        self.path = path
        self.frames = 0

    def write(self, coords, names):
        self.frames += 1
        head = (str(self.frames) + " " * 23)[:19] + "\n"
        if self.frames == 1:
            with open(self.path, "w") as fh:
                fh.write(head)
        else:
            with open(self.path, "r+") as fh:
                fh.seek(0)
                fh.write(head)
        with open(self.path, "a") as fh:
            fh.write(_xyz_lines(coords, names))


ENERGY_COLUMNS = ("Time(fs)", "E_potential", "E_kinetic", "E_total", "Temperature", "E_bond", "E_angle",
                  "E_dihedral", "E_improper", "E_VdW", "E_coulomb", "E_other", "Volume(A^3)")


class EnergiesWriter:
    """allEnergiesFile: `path.header` with the column names and one line per
    call to write().  `pressure` adds the Pressure(bar) column (the
    reference's virial-on case)."""

    def __init__(self, path, pressure=False):
        self.path = path
        self.pressure = pressure
        cols = [f"{c:>14}" for c in ENERGY_COLUMNS]
        if pressure:
            cols.append(f"{'Pressure(bar)':>14}")
        with open(path + ".header", "w") as fh:
            fh.write(" ".join(cols) + " " + f"{'E_shadow':>20}" + "\n")
        open(path, "w").close()

    def write(self, time_fs, potential, kinetic, total, temperature, bond, angle, dihedral, improper, vdw,
              coulomb, other, volume, pressure=None, shadow=0.0):
        vals = [potential, kinetic, total, temperature, bond, angle, dihedral, improper, vdw, coulomb, other,
                volume]
        line = f"{time_fs:14.2f} " + " ".join(f"{v:14.8g}" for v in vals)
        if self.pressure:
            line += f" {pressure:14.8g}"
        line += f" {shadow:20.16g}\n"
        with open(self.path, "a") as fh:
            fh.write(line)


class DCDWriter:
    """CHARMM-format DCD trajectory (little endian), as DCDTrajectoryWriter
    writes it: header on the first frame, the set count at byte offsets 8 and
    20 updated on every later frame, then per frame three FP32 records."""

    def __init__(self, path, timestep_fs=1.0, first_step=0, comment=""):
        self.path = path
        self.timestep = timestep_fs
        self.first_step = first_step
        self.comment = comment
        self.sets = 0

    def _header(self, n_atoms, n_sets):
        csize = len(self.comment) + 9
        n2 = (csize + 80) // 80
        if csize % 80:
            n2 += 1
        n164 = n2 * 80 + 4
        ts = struct.pack("<f", np.float32(self.timestep * INV_TIMEFACTOR))
        h = struct.pack("<i4siii", 84, b"CORD", n_sets, self.first_step, 1)
        h += struct.pack("<i", n_sets) + struct.pack("<5i", 0, 0, 0, 0, 0) + ts
        h += struct.pack("<9i", *([0] * 9)) + struct.pack("<ii", 24, 84)
        remarks = f"Remarks: File '{self.path}'. protomol_amd"
        h += struct.pack("<ii", n164, n2)
        h += remarks.ljust(80)[:80].encode()
        h += ("Remarks: " + self.comment).ljust((n2 - 1) * 80)[:(n2 - 1) * 80].encode()
        h += struct.pack("<i", n164) + struct.pack("<iii", 4, n_atoms, 4)
        return h

    def write(self, coords):
        c = np.asarray(coords, np.float64).reshape(-1, 3)
        n = len(c)
        if self.sets == 0 or not os.path.exists(self.path):
            with open(self.path, "wb") as fh:
                fh.write(self._header(n, 1))
            self.sets = 1
        else:
            self.sets += 1
            with open(self.path, "r+b") as fh:
                fh.seek(8)
                fh.write(struct.pack("<i", self.sets))
                fh.seek(20)
                fh.write(struct.pack("<i", self.sets))
        rec = struct.pack("<i", 4 * n)
        with open(self.path, "ab") as fh:
            for k in range(3):
                fh.write(rec + c[:, k].astype("<f4").tobytes() + rec)


def read_xyz_frames(path):
    """Frames of an XYZ or XYZ-trajectory file: list of (names, (n, 3) array)."""
    lines = open(path).read().splitlines()
    i, frames = 0, []
    if len(lines) > 1 and lines[0].strip().isdigit() and lines[1].strip().isdigit():
        i = 1  # trajectory: frame count first
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        n = int(lines[i].split()[0])
        names, xyz = [], []
        for ln in lines[i + 1:i + 1 + n]:
            parts = ln.split()
            names.append(parts[0])
            xyz.append([float(v) for v in parts[1:4]])
        frames.append((names, np.array(xyz)))
        i += 1 + n
    return frames


def read_dcd(path):
    """(header dict, list of (n, 3) FP32 frames) of a little-endian DCD."""
    b = open(path, "rb").read()
    n_sets, first, _, _ = struct.unpack_from("<iiii", b, 8)
    ts = struct.unpack_from("<f", b, 44)[0]
    n164, n2 = struct.unpack_from("<ii", b, 92)
    off = 100 + n2 * 80 + 4
    n = struct.unpack_from("<iii", b, off)[1]
    off += 12
    frames = []
    for _ in range(n_sets):
        xyz = []
        for _k in range(3):
            ln = struct.unpack_from("<i", b, off)[0]
            xyz.append(np.frombuffer(b, "<f4", ln // 4, off + 4))
            off += ln + 8
        frames.append(np.stack(xyz, 1))
    return {"sets": n_sets, "first_step": first, "timestep": ts, "title_lines": n2, "atoms": n}, frames
