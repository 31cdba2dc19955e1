"""Record-file I/O for oracle/_ref/ref_driver, the reference's own force
classes compiled from its sources (see oracle/Makefile).

TEST INFRASTRUCTURE ONLY, and only in a container that holds the reference:
tests/golden/make_golden.py uses it to write the committed fixtures.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_ref", "ref_driver")


def write_records(path, recs: dict):
    with open(path, "wb") as fp:
        for name, val in recs.items():
            a = np.asarray(val)
            if a.dtype.kind in "iub":
                a = a.astype(np.int32).reshape(-1)
                dt = 1
            else:
                a = a.astype(np.float64).reshape(-1)
                dt = 0
            fp.write(name.encode()[:15].ljust(16, b"\0"))
            fp.write(struct.pack("<iq", dt, a.size))
            fp.write(a.tobytes())


def read_records(path) -> dict:
    out = {}
    with open(path, "rb") as fp:
        while True:
            head = fp.read(16)
            if len(head) < 16:
                break
            name = head.rstrip(b"\0").decode()
            dt, cnt = struct.unpack("<iq", fp.read(12))
            dtype = np.float64 if dt == 0 else np.int32
            out[name] = np.frombuffer(fp.read(cnt * np.dtype(dtype).itemsize), dtype=dtype).copy()
    return out


def system_records(sys) -> dict:
    r = {
        "n": np.int32(sys.n),
        "x": sys.x.reshape(-1),
        "charge": sys.charge,
        "mass": sys.mass,
        "molecule": sys.molecule.astype(np.int32),
        "type": sys.lj_types().astype(np.int32),
        "ntypes": np.int32(sys.lj_table_full().shape[0]),
        "lj": sys.lj_table_full().reshape(-1),
        "excl": sys.exclusions.reshape(-1).astype(np.int32),
        "box": sys.box,
        "origin": sys.origin,
        "csf": np.float64(sys.coulomb_scaling),
    }
    return r


def run(recs: dict) -> dict:
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} missing: run `make -C oracle ref`")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.rec"), os.path.join(d, "out.rec")
        write_records(fin, recs)
        p = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"ref_driver failed ({p.returncode}): {p.stderr}")
        out = read_records(fout)
    if "census_hash" in out:
        hi, lo = out["census_hash"]
        out["census"] = (int(out["census_count"][0]), (int(hi) << 32) | int(lo))
    return out
