"""LangevinLeapfrogIntegrator's half kick and its noise stream (CPU checks of
the test-side oracle; the device path is checked in test_gpu_parity.py)."""
import math

import numpy as np

import pmtest as T
from protomol_amd import systems as S


def test_philox_known_answers():
    # Philox-4x32-10 known-answer vectors (Random123 distribution, kat_vectors)
    cases = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
             ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
             ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
              (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for c, k, want in cases:
        got = T.philox4x32_10(*[np.array([v], np.uint64) for v in c], *k)
        assert tuple(int(g[0]) for g in got) == want


def test_noise_is_standard_normal_and_independent():
    g = np.concatenate([T.langevin_gauss(11, 20000, k) for k in range(5)])
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    c = np.corrcoef(g.T)
    assert np.max(np.abs(c - np.eye(3))) < 0.01
    a, b = T.langevin_gauss(11, 20000, 0), T.langevin_gauss(11, 20000, 1)
    assert abs(np.corrcoef(a[:, 0], b[:, 0])[0, 1]) < 0.03
    assert not np.array_equal(T.langevin_gauss(11, 10, 0), T.langevin_gauss(12, 10, 0))


def _free(n, mass=12.0):
    return S.System(x=np.zeros((n, 3)), charge=np.zeros(n), mass=np.full(n, mass),
                    molecule=np.arange(n, dtype=np.int32), box=np.full(3, 50.0), origin=np.zeros(3))


def test_half_kick_damping_closed_form():
    # T = 0, F = 0: each half kick scales v by exp(-gamma dt / 2) with gamma
    # in 1/ps and dt in fs (myGamma = gamma / (1000 INV_TIMEFACTOR))
    s = _free(4)
    v0 = np.arange(12, dtype=float).reshape(4, 3) + 1.0
    zero = lambda x: (np.zeros_like(x), np.zeros(36))
    gamma, dt, steps = 91.0, 1.0, 5
    _, v, _, _ = T.langevin_oracle(zero, s.x.copy(), v0.copy(), s.mass, dt, steps, 0.0, gamma, seed=3)
    assert np.allclose(v, v0 * math.exp(-0.5 * gamma * dt / 1000) ** (2 * steps), rtol=1e-14)


def test_fluctuation_dissipation():
    # F = 0: the velocity is an Ornstein-Uhlenbeck process whose stationary
    # variance per component is kB T / m (equipartition)
    s = _free(20000, mass=16.0)
    zero = lambda x: (np.zeros_like(x), np.zeros(36))
    temp = 300.0
    _, v, _, _ = T.langevin_oracle(zero, s.x.copy(), np.zeros((s.n, 3)), s.mass, 2.0, 200, temp, 50.0, seed=9)
    var = v.var()
    assert abs(var / (S.BOLTZMANN * temp / 16.0) - 1.0) < 0.02


def test_constant_force_exact_ornstein_uhlenbeck_mean():
    # T = 0 and a constant force: dv/dt = F/m - gamma v has the exact solution
    # v(t) = v0 e^{-gamma t} + F / (gamma m) (1 - e^{-gamma t}), and every half
    # kick of LangevinLeapfrogIntegrator::doHalfKick is that solution over
    # dt/2, so any number of steps lands on it (terminal velocity F/(gamma m))
    s = _free(3, mass=14.0)
    F = np.array([[3.0, -1.5, 0.25], [0.0, 7.0, -2.0], [-4.0, 0.5, 1.0]])
    const = lambda x: (F.copy(), np.zeros(36))
    v0 = np.array([[0.1, 0.0, -0.2], [0.0, 0.0, 0.0], [-0.3, 0.2, 0.05]])
    gamma, dt = 37.0, 2.0
    for steps in (1, 7, 400):
        _, v, _, _ = T.langevin_oracle(const, s.x.copy(), v0.copy(), s.mass, dt, steps, 0.0, gamma, seed=5)
        g_int = gamma / (1000.0 * S.INV_TIMEFACTOR)  # 1 / internal time unit
        t_int = steps * dt * S.INV_TIMEFACTOR
        decay = math.exp(-g_int * t_int)
        want = v0 * decay + F / (g_int * s.mass[:, None]) * (1.0 - decay)
        assert np.allclose(v, want, rtol=1e-12, atol=1e-15), (steps, v, want)
