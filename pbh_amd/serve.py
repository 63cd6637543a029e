"""Serving collapse runs with continuous batching on one GPU.

Requests (a peak to run to its first shell crossing) arrive over time.  Each admitted request gets a
slot: a PM3D handle of its mesh size, resident in HBM, its initial conditions built on the device,
one plain step, then a device step schedule.  All active slots advance together, one captured
hipGraph replay per step (pbh_amd.batch).

The loop overlaps admission with the replays: it queues `poll` replays of the current graph on the
batch stream (asynchronous), and while they run it admits waiting requests, building their initial
conditions and first step on each new slot's own stream, so new work runs on the device next to the
running batch.  Then it waits for the replays, reads the per-slot min det(dx/dq) histories,
completes the requests that crossed (or ran out of schedule), and re-captures the graph with the
admitted requests and without the finished ones.  A slot freed by a completion is reused only after
that re-capture, so no replay in flight ever touches a slot being rebuilt.

Admission is by an HBM budget: a new slot (or an initial-condition engine) is allocated only while
the bytes held by all slots and engines stay within it; idle slots of other mesh sizes and idle
engines are released to make room.  A request that could never fit raises instead of waiting
forever.  A request that fails (an error building its state, or non-finite particles reported at
completion) completes with `error` set and its slot discarded; the others are not affected.

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
from .batch import Member, _finish


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
    error: str | None = None
    rejected: bool = False  # invalid parameters, turned away at admission (no slot, no device work)

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


def _device_pm(N: int, precision: str):
    from .pm3d import PM3D
    return PM3D(N, 1.0, N, precision=precision)


def _device_ic(pm, spec, engine):
    return ic.device_ic(pm, spec, engine=engine)


def _device_batch():
    from .batch import Batch
    return Batch()


class CollapseServer:
    """Continuous-batching server.  `make_pm(N, precision)`, `build_ic(pm, spec, engine)` and
    `batch` default to the device library (PM3D handles, pm3d_make_peak_ic, pbh_amd.batch.Batch);
    tests of the scheduling logic pass stand-ins with the same methods."""

    def __init__(self, precision: str = "mixed", hbm_budget: float = 200e9, poll: int = 8, max_steps: int = 4000,
                 make_pm=None, build_ic=None, batch=None):
        self.precision = precision
        self.budget = hbm_budget
        self.poll = poll
        self.max_steps = max_steps
        self.make_pm = make_pm or _device_pm
        self.build_ic = build_ic or _device_ic
        self.batch = batch if batch is not None else _device_batch()
        self.free = {}          # N -> idle slots
        self.engines = {}       # N -> FP64 handle building initial conditions
        self.slot_bytes = {}    # N -> bytes of one slot
        self.engine_bytes = {}  # N -> bytes of one engine
        self.held = 0           # bytes held by slots and engines
        self.active = []        # (Request, Member) in the captured graph
        self.graph = []         # their Members in the order of the last capture
        self.joining = []       # admitted while replays ran; captured at the next re-capture
        self.queue = deque()
        self.completed = []
        self.failed = []
        self.graph_builds = 0
        self.max_active = 0
        self.evicted = 0

    # -- memory
    def _release(self, pm):
        self.held -= pm.device_bytes()
        pm.close()

    def _make_room(self, need: int, keep_N: int) -> bool:
        """Release idle slots of other mesh sizes, then engines of other sizes, until `need` more
        bytes fit in the budget."""
        for N in sorted(self.free, key=lambda k: -len(self.free[k])):
            while N != keep_N and self.free[N] and self.held + need > self.budget:
                self._release(self.free[N].pop())
                self.evicted += 1
        for N in list(self.engines):
            if self.held + need <= self.budget:
                break
            if N != keep_N and self.engines[N] is not None:
                self._release(self.engines.pop(N))
                self.evicted += 1
        return self.held + need <= self.budget

    def _allocate(self, N: int, precision: str, known: int | None):
        """A new handle within the budget, or None.  The size of a handle is known after the first
        one of its kind; the first is created, measured and released again if it does not fit."""
        if known is not None and not self._make_room(known, N):
            return None
        pm = self.make_pm(N, precision)
        b = pm.device_bytes()
        if self.held + b > self.budget and not self._make_room(b, N):
            pm.close()
            return None
        self.held += b
        return pm

    def _engine(self, N: int):
        if N not in self.engines:
            if self.precision == "fp64":
                self.engines[N] = None  # FP64 slots build their own initial conditions
            else:
                eng = self._allocate(N, "fp64", self.engine_bytes.get(N))
                if eng is None:
                    return False
                self.engine_bytes[N] = eng.device_bytes()
                self.engines[N] = eng
        return self.engines[N]

    def _slot(self, N: int):
        idle = self.free.get(N)
        if idle:
            return idle.pop()
        pm = self._allocate(N, self.precision, self.slot_bytes.get(N))
        if pm is not None:
            self.slot_bytes[N] = pm.device_bytes()
        return pm

    # -- request lifecycle
    def _admit(self, req: Request, now: float) -> bool:
        spec = req.spec
        try:
            ic.validate(spec)  # an invalid request fails here, before any slot or device work
        except ValueError as err:
            req.t_admit = now
            req.rejected = True
            self._fail(req, None, f"ValueError: {err}", now)
            return True
        eng = self._engine(spec.N)
        if eng is False:
            return False
        pm = self._slot(spec.N)
        if pm is None:
            return False
        req.t_admit = now
        try:
            self.build_ic(pm, spec, eng)
            m = Member(pm, spec.a_i, req.da_frac, req.a_max, tag=str(req.rid))
            m.j0 = pm.min_jacobian()
            a1, _ = pm.step(m.a0, min(m.da_frac * m.a0, m.a_max - m.a0))
            m.a_vals = [a1]
            m.jhist = [pm.min_jacobian()]
            if m.jhist[0] <= 0.0:
                _finish(m, 0)
                self._complete(req, m, now)
                return True
            ents, avals = pm.schedule_entries(a1, m.da_frac, m.a_max, self.max_steps, pm.last_kick)
            if not avals:  # the first step already reached a_max
                _finish(m, None)
                self._complete(req, m, now)
                return True
            pm.set_schedule(ents, len(avals))
        except Exception as err:  # noqa: BLE001 -- this request fails alone
            self._fail(req, pm, f"{type(err).__name__}: {err}", now)
            return True
        m._sched_a = avals
        self.joining.append((req, m))
        return True

    def _fail(self, req: Request, pm, why: str, now: float):
        req.error = why
        req.t_done = now
        self.failed.append(req)
        self.completed.append(req)
        if pm is not None:
            self._release(pm)  # its state is not trusted: never reused

    def _complete(self, req: Request, m: Member, now: float):
        bad = m.pm.synchronize()
        if bad:
            self._fail(req, m.pm, f"{bad} particles with non-finite position or momentum", now)
            return
        req.result = m.result
        req.t_done = now
        self.completed.append(req)
        self.free.setdefault(req.spec.N, []).append(m.pm)
        if req.chain is not None and req.chain.update(req):
            nxt = Request(rid=req.rid * 100 + req.chain.done, spec=req.chain.spec(), arrival=now, da_frac=req.da_frac,
                          a_max=req.a_max, chain=req.chain)
            self.queue.append(nxt)

    def _admit_queue(self, clock) -> int:
        """Admit queued requests in arrival order while slots allow (requests of other sizes may
        pass a blocked one).  Returns the number admitted."""
        waiting = deque()
        n = 0
        while self.queue:
            req = self.queue.popleft()
            if self._admit(req, clock()):
                n += 1
            else:
                waiting.append(req)
        self.queue = waiting
        return n

    def _collect(self, now: float) -> bool:
        """Status of every replayed member (one device pass, Batch.status); completes those whose
        schedule ended (at the first crossing or its last entry).  True if any finished."""
        st = self.batch.status(len(self.graph))
        where = {id(m): k for k, m in enumerate(self.graph)}
        still = []
        for req, m in self.active:
            done, nsteps, _ = st[where[id(m)]]
            if done < nsteps:
                still.append((req, m))
                continue
            done, hist = m.pm.schedule_status()
            m.jhist = [m.jhist[0]] + list(hist)
            m.a_vals = [m.a_vals[0]] + m._sched_a[:done]
            neg = np.nonzero(np.asarray(m.jhist) <= 0.0)[0]
            _finish(m, int(neg[0]) if len(neg) else None)
            m.pm.schedule_end(m.a_vals[-1])
            self._complete(req, m, now)
        changed = len(still) != len(self.active)
        self.active = still
        return changed

    def run(self, arrivals):
        """Serve a list of (arrival_time_s, Request) to completion; returns the completed requests."""
        pend = deque(sorted(arrivals, key=lambda x: x[0]))
        t0 = time.perf_counter()

        def clock():
            return time.perf_counter() - t0

        replaying = False
        dirty = False
        while pend or self.queue or self.active or self.joining or replaying:
            now = clock()
            while pend and pend[0][0] <= now:
                t_arr, req = pend.popleft()
                req.arrival = t_arr
                self.queue.append(req)
            self._admit_queue(clock)  # runs next to the replays queued on the previous pass
            if replaying:
                self.batch.synchronize()
                replaying = False
                dirty |= self._collect(clock())
            if self.joining:
                self.active.extend(self.joining)
                self.joining = []
                dirty = True
            if not self.active:
                # nothing on the device: slots just freed may admit the queue now
                if self.queue and self._admit_queue(clock):
                    continue
                if self.queue and not pend:
                    head = self.queue[0]
                    raise RuntimeError(f"request {head.rid} (mesh {head.spec.N}^3) does not fit the HBM budget of "
                                       f"{self.budget / 1e9:.3g} GB even with every other slot released")
                if pend:
                    time.sleep(max(0.0, min(pend[0][0] - clock(), 0.01)))
                continue
            if dirty:
                self.graph = [m for _, m in self.active]
                self.batch.set([m.pm for m in self.graph])
                self.graph_builds += 1
                dirty = False
            self.max_active = max(self.max_active, len(self.active))
            self.batch.run(self.poll)
            replaying = True
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
        with np.errstate(invalid="ignore"):  # 1 - 3e + p <= 0: undefined axes, rejected at admission
            ax = 0.1 * ic.zeldovich_axes(e, p)
        out.append(ic.PeakSpec(N=N, mu=mu, sigma=tuple(ax), seed=seed0 + k, noise_rms_frac=noise, name=f"req{k}",
                               meta={"e": e, "p": p}))
    return out


def summarize(done, t_wall=None):
    """completed: requests with a result; rejected: requests whose parameters describe no peak
    (turned away at admission); failed: requests that errored after admission."""
    ok = [r for r in done if r.error is None]
    rejected = sum(1 for r in done if r.error is not None and r.rejected)
    lat = np.array([r.latency for r in ok]) if ok else np.array([np.nan])
    t_end = max(r.t_done for r in done)
    t_first = min(r.arrival for r in done)
    return {
        "completed": len(ok),
        "rejected_invalid": rejected,
        "failed": len(done) - len(ok) - rejected,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "solves_per_s": len(ok) / (t_end - t_first) if t_end > t_first else float("nan"),
        "crossed": int(sum(bool(r.result.crossed) for r in ok)),
    }
