"""The serving loop's scheduling on the CPU, with a stand-in context: every
request completes, admission never exceeds the device reservation, MD
requests run exactly their steps, arrivals wait for their time."""
import numpy as np
import pytest

from protomol_amd import _lib as L
from protomol_amd import serve


class FakeSystem:
    def __init__(self, n):
        self.n = n


class FakeCtx:
    """Records calls; a system's "result" is its positions moved by the steps it ran."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.used = 0
        self.members = {}
        self.steps = {}
        self.peak = 0
        self.advances = []

    def add(self, system):
        need = system.n * 100
        if self.used + need > self.capacity:
            raise L.PMError(L.PM_EFULL, "full")
        sid = next(i for i in range(10 ** 6) if i not in self.members)
        self.members[sid] = system
        self.used += need
        self.peak = max(self.peak, self.used)
        return sid

    def remove(self, sid):
        self.used -= self.members.pop(sid).n * 100

    def evaluate_systems(self, ids, xs):
        return [x * 2.0 for x in xs], np.ones((len(ids), L.NSLOTS))

    def set_state_systems(self, ids, xs, vs):
        for sid, x in zip(ids, xs):
            self.steps[sid] = [x.copy(), 0, False]

    def md_begin(self, ids):
        for sid in ids:
            self.steps[sid][2] = True

    def md_advance(self, ids, n, dt, energies_last=False):
        self.advances.append((tuple(ids), energies_last))
        for sid in ids:
            assert self.steps[sid][2]
            self.steps[sid][1] += n

    def md_end(self, ids, dt):
        pass

    def get_state_systems(self, ids):
        xs = [self.steps[s][0] + self.steps[s][1] for s in ids]
        return xs, xs, xs, np.zeros((len(ids), L.NSLOTS))


class FakeCtxMany(FakeCtx):
    """The same with pm_system_add_many's admission: in order, None for a
    system that does not fit, none tried after max_full of those."""

    def __init__(self, capacity):
        super().__init__(capacity)
        self.calls = 0

    def add_many(self, systems, max_full=1 << 30):
        self.calls += 1
        out, full = [], 0
        for s in systems:
            if full >= max_full:
                out.append(None)
                continue
            try:
                out.append(self.add(s))
            except L.PMError:
                out.append(None)
                full += 1
        return out


def _requests(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for k in range(n):
        size = int(rng.integers(10, 200))
        kind = "md" if rng.random() < 0.5 else "eval"
        x = rng.normal(size=(size, 3))
        out.append(serve.Request(rid=k, kind=kind, system=FakeSystem(size), x=x, v=x * 0,
                                 steps=int(rng.integers(1, 5)) * 10, arrival=t))
        t += rng.exponential(1e-4)
    return out


@pytest.mark.parametrize("make", [FakeCtx, FakeCtxMany])
def test_every_request_completes_within_capacity(make):
    ctx = make(capacity=200 * 100 * 4)
    srv = serve.Server(ctx, chunk=10)
    reqs = _requests(300, 1)
    done = srv.run(reqs, max_seconds=60)
    assert len(done) == 300 and srv.idle()
    assert ctx.peak <= ctx.capacity and ctx.used == 0
    for r in done:
        if r.kind == "md":
            assert np.array_equal(r.x_out, r.x + r.steps)
        else:
            assert np.array_equal(r.f, r.x * 2.0)
        assert r.done >= r.admitted >= r.arrival - 1e-9
    st = serve.latency_summary(done)
    assert st["requests"] == 300 and st["p99_ms"] >= st["p50_ms"]


def test_energies_summed_only_when_a_request_ends():
    ctx = FakeCtx(capacity=10 ** 9)
    srv = serve.Server(ctx, chunk=10)
    reqs = [serve.Request(rid=0, kind="md", system=FakeSystem(5), x=np.zeros((5, 3)), v=np.zeros((5, 3)), steps=30),
            serve.Request(rid=1, kind="md", system=FakeSystem(5), x=np.zeros((5, 3)), v=np.zeros((5, 3)), steps=10)]
    srv.run(reqs)
    assert [e for _, e in ctx.advances] == [True, False, True]


def test_md_steps_must_be_whole_chunks():
    srv = serve.Server(FakeCtx(10 ** 6), chunk=10)
    with pytest.raises(ValueError):
        srv.submit(serve.Request(rid=0, kind="md", system=FakeSystem(3), x=np.zeros((3, 3)), steps=15))


def test_admission_lookahead_with_add_many():
    """A request that does not fit waits; later ones are tried until
    ``lookahead`` have not fitted, and admission keeps arrival order."""
    ctx = FakeCtxMany(capacity=100 * 100)
    srv = serve.Server(ctx, chunk=10, lookahead=2)
    sizes = [60, 50, 45, 30, 20, 10]   # 60 fits; 50, 45 do not; 30 is not tried
    for k, n in enumerate(sizes):
        srv.submit(serve.Request(rid=k, kind="md", system=FakeSystem(n), x=np.zeros((n, 3)),
                                 v=np.zeros((n, 3)), steps=10))
    new_eval, new_md = srv._admit()
    assert [r.rid for r in new_md] == [0] and not new_eval
    assert [r.rid for r in srv.queue] == [1, 2, 3, 4, 5]
    assert ctx.calls == 1
