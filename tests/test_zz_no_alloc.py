"""pm_evaluate_static, the evaluation that allocates no device memory (its
device arrays live in a static array each GPU holds), through the C ABI:
bitwise the result of pm_evaluate on a context, on two calls in a row, the
second with a smaller system (the array keeps the first call's contents),
and for both the cutoff path and the GB path; a system larger than the
array is refused with PM_ECAPACITY before any device work."""
import numpy as np
import pytest

from protomol_amd import _lib as L
from protomol_amd import systems as S
from protomol_amd import workloads as W

pytestmark = pytest.mark.gpu


def _ctx_eval(desc, s):
    with L.Context(desc, [s]) as ctx:
        f = np.zeros((s.n, 3))
        e = np.zeros(L.NSLOTS)
        ctx.evaluate(0, s.x, f, e)
    return f, e


def _static_eval(desc, s):
    f = np.zeros((s.n, 3))
    e = np.zeros(L.NSLOTS)
    L.evaluate_static(desc, s, s.x, f, e)
    return f, e


def test_static_equals_context_twice_smaller_second():
    print("case evaluate_static water 1,000 then 343 molecules", flush=True)
    desc = W.desc_water32k(flags=L.FLAG_VIRIAL)
    for s in (W.water1k(), S.water_box(343, 22.5, seed=5)):
        fc, ec = _ctx_eval(desc, s)
        fs, es = _static_eval(desc, s)
        assert np.array_equal(fs, fc) and np.array_equal(es, ec), s.name
    # adds into its outputs, as SystemForce::evaluate does
    s = S.water_box(343, 22.5, seed=5)
    f = np.ones((s.n, 3))
    e = np.ones(L.NSLOTS)
    L.evaluate_static(desc, s, s.x, f, e)
    assert np.array_equal(f - 1.0, _ctx_eval(desc, s)[0])


def test_static_gb_chain_twice_smaller_second():
    print("case evaluate_static GB chains 1,200 then 500 atoms", flush=True)
    desc = W.desc_chains()
    for n in (1200, 500):
        c = S.gb_chain(n, seed=60 + n)
        fc, ec = _ctx_eval(desc, c)
        fs, es = _static_eval(desc, c)
        assert np.array_equal(fs, fc) and np.array_equal(es, ec), n


def test_static_refuses_systems_beyond_the_array():
    print("case evaluate_static too large", flush=True)
    big = W.water32k()
    f = np.zeros((big.n, 3))
    e = np.zeros(L.NSLOTS)
    with pytest.raises(L.PMError) as err:
        L.evaluate_static(W.desc_water32k(), big, big.x, f, e)
    assert err.value.code == L.PM_ECAPACITY
    assert not f.any()
