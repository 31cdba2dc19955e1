"""ctypes binding of libprotomol_amd.so (include/protomol_amd.h).

The library is built in-tree by protomol_amd.build; importing this module
never falls back to anything else: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PROTOMOL_AMD_LIB names another build of the library (e.g. one of two builds timed side by side)
LIB_PATH = os.environ.get("PROTOMOL_AMD_LIB") or os.path.join(HERE, "libprotomol_amd.so")

PM_OK, PM_EINVAL, PM_ENOMEM, PM_EDEVICE, PM_ECAPACITY, PM_ENOTREADY, PM_EFULL = 0, -1, -2, -3, -4, -5, -6
SWITCH_NONE, SWITCH_CUTOFF, SWITCH_C1, SWITCH_C2, SWITCH_CN, SWITCH_CMPCNCN = range(6)
TERM_NONE, TERM_LJ, TERM_COULOMB, TERM_EWALD_REAL = range(4)
LJ_TABLE, LJ_PERATOM = 0, 1
FLAG_VIRIAL, FLAG_GB, FLAG_BONDED, FLAG_TIMING, FLAG_EWALD_RECIPROCAL, FLAG_GB_ACE, FLAG_MOLVIRIAL = 1, 2, 4, 8, 16, 32, 64
FLAG_ALL_PAIRS = 128
NSLOTS = 36

EXPORTED = ("pm_abi_version", "pm_device_info", "pm_ctx_create", "pm_ctx_destroy", "pm_ctx_reserve",
            "pm_ctx_memory", "pm_system_footprint", "pm_system_add", "pm_system_add_many", "pm_system_remove",
            "pm_last_error",
            "pm_evaluate", "pm_evaluate_batch", "pm_evaluate_systems", "pm_pair_census", "pm_md_set_state",
            "pm_md_run", "pm_md_get_state", "pm_md_set_state_systems", "pm_md_begin", "pm_md_advance",
            "pm_md_end", "pm_md_get_state_systems", "pm_md_set_langevin", "pm_system_set_seed", "pm_stats",
            "pm_random_velocity", "pm_evaluate_static")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SwitchDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("switchon", C.c_double), ("cutoff", C.c_double),
                ("switchoff", C.c_double), ("order", C.c_double)]


class SystemDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("scaled_charge", _dp), ("mass", _dp), ("type", _ip), ("molecule", _ip),
                ("lj_mode", C.c_int32), ("n_types", C.c_int32), ("lj_table", _dp), ("lj_sigma", _dp),
                ("lj_epsilon", _dp), ("lj_scaling", C.c_double), ("n_exclusions", C.c_int32),
                ("exclusions", _ip), ("coulomb_scaling", C.c_double), ("periodic", C.c_int32),
                ("box", C.c_double * 3), ("origin", C.c_double * 3), ("n_bonds", C.c_int32), ("bonds", _ip),
                ("bond_k", _dp), ("bond_r0", _dp), ("n_angles", C.c_int32), ("angles", _ip), ("angle_k", _dp),
                ("angle_theta0", _dp), ("gb_radius", _dp), ("gb_scaling", _dp)]


class ForceDesc(C.Structure):
    _fields_ = [("term_a", C.c_int32), ("switch_a", SwitchDesc), ("term_b", C.c_int32), ("switch_b", SwitchDesc),
                ("cutoff", C.c_double), ("ewald_alpha", C.c_double), ("ewald_accuracy", C.c_double),
                ("ewald_correction", C.c_int32), ("gb_offset", C.c_double), ("gb_alpha", C.c_double),
                ("gb_beta", C.c_double), ("gb_gamma", C.c_double), ("gb_solute_dielec", C.c_double),
                ("gb_solvent_dielec", C.c_double), ("skin", C.c_double), ("rebuild_interval", C.c_int32),
                ("flags", C.c_uint32), ("cell_scale", C.c_double), ("pair_waves", C.c_int32),
                ("pair_split", C.c_int32), ("gb_ace_sigma", C.c_double), ("gb_ace_rho_s", C.c_double)]


class PMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"protomol_amd error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load the in-tree HIP library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m protomol_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.pm_abi_version.restype = C.c_int
        L.pm_device_info.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), _dp]
        L.pm_ctx_create.argtypes = [C.POINTER(ForceDesc), C.c_int32, C.POINTER(SystemDesc), C.POINTER(C.c_void_p)]
        L.pm_ctx_destroy.argtypes = [C.c_void_p]
        L.pm_last_error.argtypes = [C.c_void_p]
        L.pm_last_error.restype = C.c_char_p
        L.pm_evaluate.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp]
        L.pm_evaluate_batch.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.pm_pair_census.argtypes = [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.pm_md_set_state.argtypes = [C.c_void_p, _dp, _dp]
        L.pm_md_run.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32]
        L.pm_md_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.pm_stats.argtypes = [C.c_void_p, _dp, C.c_int32]
        u64p = C.POINTER(C.c_uint64)
        L.pm_ctx_reserve.argtypes = [C.c_void_p, C.c_uint64]
        L.pm_ctx_memory.argtypes = [C.c_void_p, u64p, u64p]
        L.pm_system_footprint.argtypes = [C.c_void_p, C.POINTER(SystemDesc), u64p]
        L.pm_system_add.argtypes = [C.c_void_p, C.POINTER(SystemDesc), _ip]
        L.pm_system_add_many.argtypes = [C.c_void_p, C.c_int32, C.POINTER(SystemDesc), C.c_int32, _ip]
        L.pm_system_remove.argtypes = [C.c_void_p, C.c_int32]
        L.pm_evaluate_systems.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp, _dp]
        L.pm_md_set_state_systems.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp]
        L.pm_md_begin.argtypes = [C.c_void_p, C.c_int32, _ip]
        L.pm_md_advance.argtypes = [C.c_void_p, C.c_int32, _ip, C.c_int32, C.c_double, C.c_int32]
        L.pm_md_end.argtypes = [C.c_void_p, C.c_int32, _ip, C.c_double]
        L.pm_md_get_state_systems.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp, _dp, _dp]
        L.pm_md_set_langevin.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint32]
        L.pm_system_set_seed.argtypes = [C.c_void_p, C.c_int32, C.c_uint32]
        L.pm_random_velocity.argtypes = [C.c_int32, _dp, C.c_double, C.c_uint32, C.c_int32, _dp]
        L.pm_evaluate_static.argtypes = [C.POINTER(ForceDesc), C.POINTER(SystemDesc), _dp, _dp, _dp]
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != PM_OK:
        msg = lib().pm_last_error(ctx)
        raise PMError(rc, msg.decode() if msg else "")


def ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int32()
    hbm = C.c_double()
    check(lib().pm_device_info(name, 256, C.byref(ncu), C.byref(hbm)))
    return name.value.decode(), ncu.value, hbm.value


def random_velocity(mass, temperature: float, seed: int, randomtype: int = 0) -> np.ndarray:
    """ProtoMol's velocity start (pm_random_velocity): (n, 3) velocities."""
    m = np.ascontiguousarray(mass, np.float64).reshape(-1)
    v = np.zeros((len(m), 3))
    check(lib().pm_random_velocity(len(m), ptr(m), float(temperature), int(seed) & 0xFFFFFFFF, int(randomtype),
                                   ptr(v)))
    return v


def evaluate_static(force: ForceDesc, system, x, f, e):
    """pm_evaluate_static: one evaluation with no context and no device
    allocation (ADDS into f and e)."""
    keep = _Keep(system, bool(force.flags & FLAG_BONDED))
    xx = np.ascontiguousarray(x, np.float64)
    assert f.dtype == np.float64 and f.flags.c_contiguous and e.dtype == np.float64
    check(lib().pm_evaluate_static(C.byref(force), C.byref(keep.desc), ptr(xx), ptr(f), ptr(e)))


def switch(kind=SWITCH_NONE, switchon=0.0, cutoff=0.0, switchoff=0.0, order=0.0) -> SwitchDesc:
    return SwitchDesc(kind, switchon, cutoff, switchoff, order)


class _Keep:
    """Holds the numpy arrays a SystemDesc points into."""

    def __init__(self, sys, bonded: bool):
        if bonded and (len(sys.bond_k) != len(sys.bonds) or len(sys.angle_k) != len(sys.angles)):
            raise ValueError(f"{sys.name}: bonded terms need a constant per bond and per angle")
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        i32 = lambda a: None if a is None else np.ascontiguousarray(a, np.int32).reshape(-1)
        self.charge = f64(sys.charge)
        self.mass = f64(sys.mass)
        self.molecule = i32(sys.molecule)
        self.type = i32(sys.type)
        self.lj_table = f64(sys.lj_table)
        self.lj_sigma = f64(sys.lj_sigma)
        self.lj_epsilon = f64(sys.lj_epsilon)
        self.excl = i32(sys.exclusions)
        self.bonds = i32(sys.bonds)
        self.bond_k = f64(sys.bond_k)
        self.bond_r0 = f64(sys.bond_r0)
        self.angles = i32(sys.angles)
        self.angle_k = f64(sys.angle_k)
        self.angle_t0 = f64(sys.angle_t0)
        self.gb_radius = f64(sys.gb_radius)
        self.gb_scaling = f64(sys.gb_scaling)
        d = SystemDesc()
        d.n_atoms = sys.n
        d.scaled_charge = ptr(self.charge)
        d.mass = ptr(self.mass)
        d.type = ptr(self.type, _ip)
        d.molecule = ptr(self.molecule, _ip)
        if sys.lj_table is not None:
            d.lj_mode = LJ_TABLE
            d.n_types = sys.lj_table.shape[0]
            d.lj_table = ptr(self.lj_table)
        else:
            d.lj_mode = LJ_PERATOM
            d.lj_sigma = ptr(self.lj_sigma)
            d.lj_epsilon = ptr(self.lj_epsilon)
            d.lj_scaling = sys.lj_scaling
        d.n_exclusions = len(sys.exclusions)
        d.exclusions = ptr(self.excl, _ip)
        d.coulomb_scaling = sys.coulomb_scaling
        d.periodic = 1 if sys.periodic else 0
        for k in range(3):
            d.box[k] = float(sys.box[k])
            d.origin[k] = float(sys.origin[k])
        d.n_bonds = len(sys.bonds) if bonded else 0
        d.bonds = ptr(self.bonds, _ip)
        d.bond_k = ptr(self.bond_k)
        d.bond_r0 = ptr(self.bond_r0)
        d.n_angles = len(sys.angles) if bonded else 0
        d.angles = ptr(self.angles, _ip)
        d.angle_k = ptr(self.angle_k)
        d.angle_theta0 = ptr(self.angle_t0)
        d.gb_radius = ptr(self.gb_radius)
        d.gb_scaling = ptr(self.gb_scaling)
        self.desc = d


def _ids(ids):
    return np.ascontiguousarray(np.asarray(ids, np.int32).reshape(-1))


def _cat(arrs):
    return np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).reshape(-1, 3) for a in arrs]))


class Context:
    """One device context (pm_ctx) holding systems by id.  Systems passed at
    creation get ids 0..n-1; ``add``/``remove`` let systems join and leave."""

    def __init__(self, force: ForceDesc, systems=(), reserve_bytes: int | None = None):
        self.force = force
        self._bonded = bool(force.flags & FLAG_BONDED)
        systems = list(systems)
        keep = [_Keep(s, self._bonded) for s in systems]
        arr = (SystemDesc * max(1, len(systems)))(*[k.desc for k in keep])
        h = C.c_void_p()
        rc = lib().pm_ctx_create(C.byref(force), len(systems), arr, C.byref(h))
        if rc != PM_OK:
            raise PMError(rc, lib().pm_last_error(None).decode())
        self.h = h
        self.members = {k: s for k, s in enumerate(systems)}
        if reserve_bytes:
            check(lib().pm_ctx_reserve(self.h, int(reserve_bytes)), self.h)

    # ---- membership
    @property
    def systems(self):
        return [self.members[k] for k in sorted(self.members)]

    @property
    def offsets(self):
        return np.cumsum([0] + [s.n for s in self.systems])

    def footprint(self, system) -> int:
        b = C.c_uint64()
        check(lib().pm_system_footprint(self.h, C.byref(_Keep(system, self._bonded).desc), C.byref(b)), self.h)
        return b.value

    def memory(self):
        r, u = C.c_uint64(), C.c_uint64()
        check(lib().pm_ctx_memory(self.h, C.byref(r), C.byref(u)), self.h)
        return r.value, u.value

    def add(self, system) -> int:
        """Add a system; raises PMError with code PM_EFULL when the reservation has no room."""
        i = C.c_int32()
        check(lib().pm_system_add(self.h, C.byref(_Keep(system, self._bonded).desc), C.byref(i)), self.h)
        self.members[i.value] = system
        return i.value

    def add_many(self, systems, max_full: int = 1 << 30) -> list:
        """Add systems in order (pm_system_add_many): a list of ids, None for a
        system the reservation had no room for (and for every system after
        ``max_full`` of those, which are not tried)."""
        systems = list(systems)
        if not systems:
            return []
        keep = [_Keep(s, self._bonded) for s in systems]
        arr = (SystemDesc * len(systems))(*[k.desc for k in keep])
        ids = np.full(len(systems), -1, np.int32)
        check(lib().pm_system_add_many(self.h, len(systems), arr, max(1, int(max_full)), ptr(ids, _ip)), self.h)
        out = []
        for s, i in zip(systems, ids.tolist()):
            if i >= 0:
                self.members[i] = s
            out.append(i if i >= 0 else None)
        return out

    def remove(self, sid: int):
        check(lib().pm_system_remove(self.h, sid), self.h)
        del self.members[sid]

    def close(self):
        if getattr(self, "h", None):
            lib().pm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- SystemForce::evaluate semantics: ADD into forces and energies
    def evaluate(self, system: int, x: np.ndarray, f: np.ndarray, e: np.ndarray):
        xx = np.ascontiguousarray(x, np.float64)
        assert f.dtype == np.float64 and f.flags.c_contiguous and e.dtype == np.float64
        check(lib().pm_evaluate(self.h, system, ptr(xx), ptr(f), ptr(e)), self.h)

    def evaluate_batch(self, x: np.ndarray, f: np.ndarray, e: np.ndarray):
        xx = np.ascontiguousarray(x, np.float64)
        check(lib().pm_evaluate_batch(self.h, ptr(xx), ptr(f), ptr(e)), self.h)

    def evaluate_systems(self, ids, xs):
        """Evaluate the listed systems in one fused launch; returns (forces per system, energies (k, 36))."""
        ii = _ids(ids)
        x = _cat(xs)
        f = np.zeros_like(x)
        e = np.zeros((len(ii), NSLOTS))
        check(lib().pm_evaluate_systems(self.h, len(ii), ptr(ii, _ip), ptr(x), ptr(f), ptr(e)), self.h)
        return self._split(ii, f), e

    def census(self, system: int, x):
        """(count, hash) of the pair set at positions x (a fresh list), or with
        x None of the MD state's current list at its positions (no rebuild)."""
        xx = None if x is None else np.ascontiguousarray(x, np.float64)
        cnt, hsh = C.c_uint64(), C.c_uint64()
        check(lib().pm_pair_census(self.h, system, ptr(xx), C.byref(cnt), C.byref(hsh)), self.h)
        return cnt.value, hsh.value

    # ---- MD over all systems
    def md_set_state(self, x, v):
        xx = np.ascontiguousarray(x, np.float64)
        vv = np.ascontiguousarray(v, np.float64)
        check(lib().pm_md_set_state(self.h, ptr(xx), ptr(vv)), self.h)

    def md_run(self, n_steps: int, dt_fs: float, energies: bool = True, every: int = 0):
        """LeapfrogIntegrator::run of n_steps.  energies: energies and virial
        at the last step; every > 0: also on every every-th step."""
        cadence = every if every > 0 else (n_steps if energies else 0)
        check(lib().pm_md_run(self.h, n_steps, dt_fs, cadence), self.h)

    def md_get_state(self):
        n = int(self.offsets[-1])
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 3))
        e = np.zeros((len(self.members), NSLOTS))
        check(lib().pm_md_get_state(self.h, ptr(x), ptr(v), ptr(f), ptr(e)), self.h)
        return x, v, f, e

    # ---- MD over a subset (LeapfrogIntegrator::run = begin + advance + end)
    def _split(self, ii, a):
        out, off = [], 0
        for sid in ii:
            n = self.members[int(sid)].n
            out.append(a[off:off + n])
            off += n
        return out

    def set_state_systems(self, ids, xs, vs):
        ii = _ids(ids)
        x, v = _cat(xs), _cat(vs)
        check(lib().pm_md_set_state_systems(self.h, len(ii), ptr(ii, _ip), ptr(x), ptr(v)), self.h)

    def md_begin(self, ids):
        ii = _ids(ids)
        check(lib().pm_md_begin(self.h, len(ii), ptr(ii, _ip)), self.h)

    def md_advance(self, ids, n_steps: int, dt_fs: float, energies_last: bool = False):
        ii = _ids(ids)
        check(lib().pm_md_advance(self.h, len(ii), ptr(ii, _ip), n_steps, dt_fs, 1 if energies_last else 0), self.h)

    def md_end(self, ids, dt_fs: float):
        ii = _ids(ids)
        check(lib().pm_md_end(self.h, len(ii), ptr(ii, _ip), dt_fs), self.h)

    def get_state_systems(self, ids):
        """(x, v, f) lists per system and energies (k, 36)."""
        ii = _ids(ids)
        n = sum(self.members[int(k)].n for k in ii)
        x, v, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        e = np.zeros((len(ii), NSLOTS))
        check(lib().pm_md_get_state_systems(self.h, len(ii), ptr(ii, _ip), ptr(x), ptr(v), ptr(f), ptr(e)), self.h)
        return self._split(ii, x), self._split(ii, v), self._split(ii, f), e

    def set_langevin(self, temperature: float, gamma_per_ps: float, seed: int = 1234):
        """LangevinLeapfrogIntegrator for later MD calls (gamma 0: plain leapfrog)."""
        check(lib().pm_md_set_langevin(self.h, temperature, gamma_per_ps, seed), self.h)

    def set_seed(self, sid: int, seed: int):
        check(lib().pm_system_set_seed(self.h, sid, seed), self.h)

    def stats(self):
        s = np.zeros(10)
        check(lib().pm_stats(self.h, ptr(s), 10), self.h)
        return {"rebuilds": s[0], "regrow_reruns": s[1], "list_entries": s[2], "device_ms": s[3], "pair_ms": s[4],
                "pair_launches": s[5], "max_displacement": s[6], "skin": s[7], "reciprocal_ms": s[8],
                "kvectors": s[9]}
