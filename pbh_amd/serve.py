"""Serving collapse runs with continuous batching on one GPU.

Requests (a peak to run to its first shell crossing) arrive over time.  Each admitted request gets a
slot: a PM3D handle of its mesh size, resident in HBM, its initial conditions built on the device,
one plain step, then a device step schedule.  All active slots advance together, one captured
hipGraph replay per step (pbh_amd.batch); every `poll` replays the host reads the per-slot
min det(dx/dq) histories, completes the requests that crossed, frees their slots and admits waiting
requests (the graph is re-captured whenever the active set changes).  Admission is by an HBM budget:
a new slot is allocated only while the bytes held by all slots and initial-condition engines stay
within it; otherwise the request waits for a free slot of its size.

Threshold-bisection chains search mu_th such that the peak crosses before a_target: each probe is a
request whose completion submits the next probe (8 probes per chain).

No counterpart in the reference (its driver runs one box per process, pbhgr/scripts/pm3d_peak.py,
and parameter scans run as a process pool, pbhgr/pbhgr/scan.py:67-76).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import ic
from .batch import Batch, Member, _finish
from .pm3d import PM3D


@dataclass
class Request:
    rid: int
    spec: ic.PeakSpec
    arrival: float
    da_frac: float = 0.02
    a_max: float = 1.0
    chain: "BisectionChain | None" = None
    # filled by the server
    t_admit: float = float("nan")
    t_done: float = float("nan")
    result: object = None

    @property
    def latency(self):
        return self.t_done - self.arrival


@dataclass
class BisectionChain:
    """Bisection on mu for 'crosses before a_target' at fixed (e, p); `probes` requests in series."""
    e: float
    p: float
    N: int
    a_target: float
    lo: float = 0.02
    hi: float = 0.5
    probes: int = 8
    seed: int = 0
    done: int = 0
    history: list = field(default_factory=list)

    def spec(self) -> ic.PeakSpec:
        mu = 0.5 * (self.lo + self.hi)
        ax = 0.1 * ic.zeldovich_axes(self.e, self.p)
        return ic.PeakSpec(N=self.N, mu=mu, sigma=tuple(ax), seed=self.seed, name=f"bisect_e{self.e:g}_{self.done}",
                           meta={"e": self.e, "p": self.p})

    def update(self, req: Request):
        crossed = bool(req.result.crossed and req.result.a_sc <= self.a_target)
        self.history.append((req.spec.mu, crossed, req.result.a_sc))
        if crossed:
            self.hi = req.spec.mu
        else:
            self.lo = req.spec.mu
        self.done += 1
        return self.done < self.probes

    @property
    def mu_threshold(self):
        return 0.5 * (self.lo + self.hi)


class CollapseServer:
    def __init__(self, precision: str = "mixed", hbm_budget: float = 200e9, poll: int = 8, max_steps: int = 4000):
        self.precision = precision
        self.budget = hbm_budget
        self.poll = poll
        self.max_steps = max_steps
        self.free = {}        # N -> [PM3D]
        self.engines = {}     # N -> FP64 PM3D building initial conditions
        self.held = 0         # bytes held by slots and engines
        self.batch = Batch()
        self.active = []      # (Request, Member)
        self.queue = deque()
        self.completed = []
        self.graph_builds = 0
        self.max_active = 0

    # -- slots
    def _engine(self, N):
        if N not in self.engines:
            e = PM3D(N, 1.0, N, precision="fp64")
            self.engines[N] = e
            self.held += e.device_bytes()
        return self.engines[N]

    def _slot(self, N):
        lst = self.free.setdefault(N, [])
        if lst:
            return lst.pop()
        probe_bytes = getattr(self, f"_bytes_{N}", None)
        if probe_bytes is not None and self.held + probe_bytes > self.budget:
            return None
        pm = PM3D(N, 1.0, N, precision=self.precision)
        b = pm.device_bytes()
        setattr(self, f"_bytes_{N}", b)
        if self.held + b > self.budget:
            del pm
            return None
        self.held += b
        return pm

    # -- request lifecycle
    def _admit(self, req: Request, now: float) -> bool:
        spec = req.spec
        pm = self._slot(spec.N)
        if pm is None:
            return False
        ic.device_ic(pm, spec, engine=self._engine(spec.N))
        m = Member(pm, spec.a_i, req.da_frac, req.a_max, tag=str(req.rid))
        m.j0 = pm.min_jacobian()
        a1, _ = pm.step(m.a0, min(m.da_frac * m.a0, m.a_max - m.a0))
        m.a_vals = [a1]
        m.jhist = [pm.min_jacobian()]
        req.t_admit = now
        if m.jhist[0] <= 0.0:
            _finish(m, 0)
            self._complete(req, m, now)
            return True
        ents, avals = pm.schedule_entries(a1, m.da_frac, m.a_max, self.max_steps, pm.last_kick)
        pm.set_schedule(ents, len(avals))
        m._sched_a = avals
        self.active.append((req, m))
        return True

    def _complete(self, req: Request, m: Member, now: float):
        req.result = m.result
        req.t_done = now
        self.completed.append(req)
        self.free.setdefault(req.spec.N, []).append(m.pm)
        if req.chain is not None and req.chain.update(req):
            nxt = Request(rid=req.rid * 100 + req.chain.done, spec=req.chain.spec(), arrival=now, da_frac=req.da_frac,
                          a_max=req.a_max, chain=req.chain)
            self.queue.append(nxt)

    def run(self, arrivals):
        """Serve a list of (arrival_time_s, Request) to completion; returns the completed requests."""
        arrivals = sorted(arrivals, key=lambda x: x[0])
        pend = deque(arrivals)
        t0 = time.perf_counter()
        dirty = True
        while pend or self.queue or self.active:
            now = time.perf_counter() - t0
            while pend and pend[0][0] <= now:
                t_arr, req = pend.popleft()
                req.arrival = t_arr
                self.queue.append(req)
            # admit in arrival order while slots allow (requests of other sizes may pass a blocked one)
            waiting = deque()
            while self.queue:
                req = self.queue.popleft()
                if self._admit(req, time.perf_counter() - t0):
                    dirty = True
                else:
                    waiting.append(req)
            self.queue = waiting
            if not self.active:
                if pend:
                    time.sleep(max(0.0, min(pend[0][0] - (time.perf_counter() - t0), 0.01)))
                continue
            if dirty:
                self.batch.set([m.pm for _, m in self.active])
                self.graph_builds += 1
                dirty = False
            self.max_active = max(self.max_active, len(self.active))
            self.batch.run(self.poll)
            self.batch.synchronize()
            now = time.perf_counter() - t0
            still = []
            for req, m in self.active:
                done, hist = m.pm.schedule_status()
                m.jhist = [m.jhist[0]] + list(hist)
                m.a_vals = [m.a_vals[0]] + m._sched_a[:done]
                neg = np.nonzero(np.asarray(m.jhist) <= 0.0)[0]
                if len(neg) or done >= len(m._sched_a):
                    _finish(m, int(neg[0]) if len(neg) else None)
                    m.pm.schedule_end(m.a_vals[-1])
                    self._complete(req, m, now)
                    dirty = True
                else:
                    still.append((req, m))
            self.active = still
        return self.completed


def poisson_arrivals(rate: float, n: int, seed: int = 0):
    """n arrival times of a Poisson process of the given rate (per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def config_requests(n: int, N: int = 256, seed0: int = 1000, noise: float = 0.01):
    """Independent requests drawn like the mixed-precision workload: mu ~ U[0.05, 0.5],
    e ~ U[0, 0.3], p ~ U[-e, e], axes 0.1 L, random field rms noise*mu (per-request noise stream)."""
    out = []
    for k in range(n):
        rng = np.random.default_rng(seed0 + k)
        mu = float(rng.uniform(0.05, 0.5))
        e = float(rng.uniform(0.0, 0.3))
        p = float(rng.uniform(-e, e))
        ax = 0.1 * ic.zeldovich_axes(e, p)
        out.append(ic.PeakSpec(N=N, mu=mu, sigma=tuple(ax), seed=seed0 + k, noise_rms_frac=noise, name=f"req{k}",
                               meta={"e": e, "p": p}))
    return out


def summarize(done, t_wall=None):
    lat = np.array([r.latency for r in done])
    t_end = max(r.t_done for r in done)
    t_first = min(r.arrival for r in done)
    return {
        "completed": len(done),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "solves_per_s": len(done) / (t_end - t_first) if t_end > t_first else float("nan"),
        "crossed": int(sum(bool(r.result.crossed) for r in done)),
    }
