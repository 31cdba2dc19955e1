"""Continuous-batching request loop over one device context.

Requests are either a single force evaluation (ProtoMol's
``SystemForce::evaluate``) or a short MD segment (``LeapfrogIntegrator::run``
of n steps).  The server keeps every admitted system resident on the device
and works in ticks:

  1. admit waiting requests, in arrival order, while the context's device
     reservation has room for their footprint (a request that does not fit
     waits; later, smaller ones may pass it, at most ``lookahead`` places);
  2. evaluate all newly admitted evaluation requests in one fused launch and
     hand them back;
  3. set the state of newly admitted MD requests (F(x0), one fused launch)
     and arm their first half kick;
  4. advance every resident MD request by ``chunk`` steps in one fused
     launch sequence (graph-replayed), summing energies at the last step
     when some request ends there;
  5. close the requests that ended (half kick), read their state back and
     free their memory.

A system's arithmetic depends only on its own calls (fixed-point force
accumulation, per-system pair-list decisions), so each result is the same
whether the request ran alone or among others.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque

import numpy as np

from . import _lib as L


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                   # "eval" or "md"
    system: object              # systems.System
    x: np.ndarray
    v: np.ndarray | None = None
    steps: int = 0
    arrival: float = 0.0        # seconds, on the server clock
    admitted: float = -1.0
    done: float = -1.0
    sid: int = -1               # system id in the context while resident
    left: int = 0               # MD steps still to run
    # results
    f: np.ndarray | None = None
    e: np.ndarray | None = None
    x_out: np.ndarray | None = None
    v_out: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Server:
    """Serve requests on one context (``ctx``: an ``_lib.Context`` holding a
    device reservation, or any object with the same methods)."""

    def __init__(self, ctx, dt_fs: float = 1.0, chunk: int = 10, lookahead: int = 8, clock=time.perf_counter):
        self.ctx = ctx
        self.dt = dt_fs
        self.chunk = chunk
        self.lookahead = lookahead
        self.clock = clock
        self.t0 = clock()
        self.queue: deque[Request] = deque()
        self.resident: dict[int, Request] = {}   # sid -> MD request
        self.finished: list[Request] = []
        self.ticks = 0
        self.busy_s = 0.0

    def now(self) -> float:
        return self.clock() - self.t0

    def submit(self, req: Request):
        if req.kind == "md" and (req.steps < 1 or req.steps % self.chunk):
            raise ValueError(f"MD requests run whole chunks of {self.chunk} steps")
        req.left = req.steps
        self.queue.append(req)

    def idle(self) -> bool:
        return not self.queue and not self.resident

    # -- one scheduling round
    def _admit(self):
        """Admit waiting requests in arrival order in one pm_system_add_many
        call: a request that does not fit waits, later ones are still tried
        until ``lookahead`` requests have not fitted."""
        new_eval, new_md = [], []
        if not self.queue:
            return new_eval, new_md
        cand = list(self.queue)
        if hasattr(self.ctx, "add_many"):
            ids = self.ctx.add_many([r.system for r in cand], max_full=self.lookahead)
        else:
            ids, full = [], 0
            for r in cand:
                if full >= self.lookahead:
                    ids.append(None)
                    continue
                try:
                    ids.append(self.ctx.add(r.system))
                except L.PMError as err:
                    if err.code != L.PM_EFULL:
                        raise
                    ids.append(None)
                    full += 1
        pending = deque()
        now = self.now()
        for req, sid in zip(cand, ids):
            if sid is None:
                pending.append(req)
                continue
            req.sid = sid
            req.admitted = now
            (new_eval if req.kind == "eval" else new_md).append(req)
        self.queue = pending
        return new_eval, new_md

    def _finish(self, req: Request):
        self.ctx.remove(req.sid)
        req.done = self.now()
        self.finished.append(req)

    def tick(self):
        t_start = self.clock()
        new_eval, new_md = self._admit()
        if new_eval:
            fs, es = self.ctx.evaluate_systems([r.sid for r in new_eval], [r.x for r in new_eval])
            for r, f, e in zip(new_eval, fs, es):
                r.f, r.e = f, e
                self._finish(r)
        if new_md:
            ids = [r.sid for r in new_md]
            self.ctx.set_state_systems(ids, [r.x for r in new_md], [r.v for r in new_md])
            self.ctx.md_begin(ids)
            for r in new_md:
                self.resident[r.sid] = r
        if self.resident:
            ids = sorted(self.resident)
            ending = [sid for sid in ids if self.resident[sid].left == self.chunk]
            self.ctx.md_advance(ids, self.chunk, self.dt, energies_last=bool(ending))
            for sid in ids:
                self.resident[sid].left -= self.chunk
            if ending:
                self.ctx.md_end(ending, self.dt)
                xs, vs, fs, es = self.ctx.get_state_systems(ending)
                for k, sid in enumerate(ending):
                    r = self.resident.pop(sid)
                    r.x_out, r.v_out, r.f, r.e = xs[k], vs[k], fs[k], es[k]
                    self._finish(r)
        self.ticks += 1
        self.busy_s += self.clock() - t_start

    def run(self, requests, max_seconds: float = 600.0):
        """Feed ``requests`` (sorted by ``arrival``, seconds from now) as the
        clock passes their arrival times; return them when all are done."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.t0 = self.clock()
        while pending or not self.idle():
            now = self.now()
            if now > max_seconds:
                raise RuntimeError("serving run exceeded its time limit")
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if self.idle():
                # nothing resident: wait for the next arrival
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            self.tick()
        return self.finished


def latency_summary(reqs) -> dict:
    lat = np.array([r.latency for r in reqs])
    out = {"requests": len(reqs), "p50_ms": float(np.percentile(lat, 50) * 1e3),
           "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3)}
    for kind in ("eval", "md"):
        sub = np.array([r.latency for r in reqs if r.kind == kind])
        if len(sub):
            out[f"{kind}_p50_ms"] = float(np.percentile(sub, 50) * 1e3)
            out[f"{kind}_p99_ms"] = float(np.percentile(sub, 99) * 1e3)
    return out
