"""Scheduling logic of the continuous-batching server (pbh_amd.serve) on the CPU, with stand-in
handles: HBM-budget admission across mesh sizes (idle slots of another size are released, a request
that can never fit raises instead of spinning), a request that fails alone (error building its
state, non-finite particles at completion) while the others complete as they would alone, an empty
schedule, and chains.  The stand-ins follow the PM3D/Batch methods the server calls; a stand-in box
crosses after a fixed number of steps, and its schedule ends on the "device" at the crossing as the
library's does."""
import numpy as np
import pytest

from pbh_amd import ic
from pbh_amd.serve import BisectionChain, CollapseServer, Request, summarize


class StubPM:
    """A box that crosses after `cross` steps: min det(dx/dq) after step k is 1 - k/(cross - 0.5)."""

    live = 0

    def __init__(self, N, precision, nbytes, cross_of):
        self.N, self.precision = N, precision
        self.nbytes = nbytes
        self.cross_of = cross_of
        self.cross = None
        self.k = 0
        self.bad = 0
        self.closed = False
        self.sched = None
        self.done = 0
        StubPM.live += 1

    def device_bytes(self):
        return self.nbytes

    def close(self):
        assert not self.closed
        self.closed = True
        StubPM.live -= 1

    def _j(self, k):
        return 1.0 - k / (self.cross - 0.5)

    def min_jacobian(self):
        return self._j(self.k)

    def step(self, a, da):
        self.k += 1
        self.last_kick = (0.5 * da, 1.0)
        return a + da, None

    def schedule_entries(self, a, da_frac, a_max, max_steps, prev):
        avals = []
        while len(avals) < max_steps and a < a_max:
            a = a + min(da_frac * a, a_max - a)
            avals.append(a)
        return list(range(len(avals))), avals

    def set_schedule(self, ents, n):
        self.sched, self.done = n, 0

    def schedule_status(self, with_history=True):
        return self.done, [self._j(1 + i + 1) for i in range(self.done)]

    def schedule_end(self, a):
        self.k += self.done
        return self.done

    def synchronize(self):
        return self.bad

    def advance(self, n):
        for _ in range(n):
            if self.done >= self.sched or self._j(1 + self.done) <= 0.0:
                return  # the table ended at the crossing
            self.done += 1


class StubBatch:
    def __init__(self):
        self.members = []
        self.sets = 0

    def set(self, pms):
        assert all(not pm.closed for pm in pms)
        self.members = list(pms)
        self.sets += 1

    def run(self, n):
        for pm in self.members:
            assert not pm.closed
            pm.advance(n)

    def synchronize(self):
        pass

    def status(self, n):
        """(done, nsteps, jmin) per member; a table ends (nsteps = done) at the crossing."""
        assert n == len(self.members)
        out = []
        for pm in self.members:
            crossed = pm.done > 0 and pm._j(1 + pm.done) <= 0.0
            out.append((pm.done, pm.done if crossed else pm.sched, pm._j(1 + pm.done) if pm.done else 1e300))
        return out


def make_server(budget, sizes, fail_mu=None, bad_mu=None, cross=None, **kw):
    """sizes: N -> bytes of a slot (engines take the same); cross: mu -> steps to crossing."""
    cross = cross or {}

    def make_pm(N, precision):
        return StubPM(N, precision, sizes[N], cross)

    def build_ic(pm, spec, engine):
        if fail_mu is not None and spec.mu == fail_mu:
            raise RuntimeError("initial conditions failed")
        pm.cross = cross.get(spec.mu, 12)
        pm.k = 0
        pm.bad = 7 if (bad_mu is not None and spec.mu == bad_mu) else 0
        return pm

    return CollapseServer(precision="mixed", hbm_budget=budget, poll=4, make_pm=make_pm, build_ic=build_ic,
                          batch=StubBatch(), **kw)


def spec(N, mu, seed=0):
    return ic.PeakSpec(N=N, mu=mu, sigma=(0.1, 0.1, 0.1), seed=seed)


@pytest.mark.timeout(30)
def test_budget_full_of_one_size_admits_the_other():
    """Two slots of 32^3 (+ its engine) fill the budget; a 64^3 request waits until they finish,
    then the idle 32^3 slots and engine are released for it."""
    StubPM.live = 0
    srv = make_server(budget=300, sizes={32: 100, 64: 150})
    reqs = [(0.0, Request(rid=k, spec=spec(32, 0.1 + k / 10), arrival=0.0)) for k in range(2)]
    reqs.append((0.0, Request(rid=9, spec=spec(64, 0.3), arrival=0.0)))
    done = srv.run(reqs)
    assert sorted(r.rid for r in done) == [0, 1, 9]
    assert all(r.error is None and r.result.crossed for r in done)
    assert srv.held <= srv.budget and srv.evicted >= 2
    assert done[-1].rid == 9 and done[-1].t_admit >= max(r.t_done for r in done[:2])


@pytest.mark.timeout(30)
def test_request_larger_than_budget_raises():
    srv = make_server(budget=150, sizes={32: 100, 64: 200})
    with pytest.raises(RuntimeError, match="does not fit"):
        srv.run([(0.0, Request(rid=1, spec=spec(64, 0.2), arrival=0.0))])


@pytest.mark.timeout(30)
def test_failed_request_does_not_disturb_the_others():
    """One request fails while building its state, one ends with non-finite particles; the other
    requests complete with the results they get alone, and the failed slots are discarded."""
    mus = [0.1, 0.2, 0.3, 0.4, 0.5]
    cross = {0.1: 9, 0.2: 13, 0.3: 17, 0.4: 21, 0.5: 25}
    alone = {}
    for mu in mus:
        s = make_server(budget=10_000, sizes={32: 100}, cross=cross)
        r = s.run([(0.0, Request(rid=0, spec=spec(32, mu), arrival=0.0))])[0]
        alone[mu] = r.result
    StubPM.live = 0
    srv = make_server(budget=10_000, sizes={32: 100}, fail_mu=0.2, bad_mu=0.4, cross=cross)
    done = {r.rid: r for r in srv.run([(0.0, Request(rid=k, spec=spec(32, mu), arrival=0.0))
                                       for k, mu in enumerate(mus)])}
    assert done[1].error.startswith("RuntimeError") and done[1].result is None
    assert "non-finite" in done[3].error
    for k in (0, 2, 4):
        got, ref = done[k].result, alone[mus[k]]
        assert done[k].error is None
        assert (got.crossed, got.steps, got.a_sc) == (ref.crossed, ref.steps, ref.a_sc)
    assert len(srv.failed) == 2
    # the two failed slots were released; the three good ones are idle, plus the engine
    assert StubPM.live == 3 + 1
    s = summarize(list(done.values()))
    assert s["completed"] == 3 and s["failed"] == 2 and s["rejected_invalid"] == 0


@pytest.mark.timeout(30)
def test_empty_schedule_completes():
    """A request whose first plain step already reaches a_max has no schedule: it completes
    (not crossed) instead of raising, and its slot is reused."""
    srv = make_server(budget=10_000, sizes={32: 100})
    reqs = [(0.0, Request(rid=k, spec=spec(32, 0.3), arrival=0.0, da_frac=0.02, a_max=0.0101)) for k in range(3)]
    done = srv.run(reqs)
    assert len(done) == 3 and all(r.error is None and not r.result.crossed for r in done)
    assert len(srv.free[32]) == 1


@pytest.mark.timeout(30)
def test_chain_probes_follow_completions():
    cross = {}
    ch = BisectionChain(e=0.1, p=0.0, N=32, a_target=0.2, probes=5)
    srv = make_server(budget=10_000, sizes={32: 100}, cross=cross)
    done = srv.run([(0.0, Request(rid=1, spec=ch.spec(), arrival=0.0, chain=ch))])
    assert len(done) == 5 and ch.done == 5
    assert len({h[0] for h in ch.history}) == 5  # every probe bisects to a new mu
    assert np.all(np.diff([r.t_done for r in done]) >= 0)


@pytest.mark.timeout(30)
def test_invalid_request_fails_before_any_slot():
    """A request whose Zel'dovich axes are undefined (1 - 3e + p <= 0 gives NaN widths) fails at
    admission without taking a slot or building anything; the valid one completes."""
    StubPM.live = 0
    srv = make_server(budget=10_000, sizes={32: 100})
    with np.errstate(invalid="ignore"):
        ax = tuple(0.1 * ic.zeldovich_axes(0.3, -0.3))
    bad = ic.PeakSpec(N=32, mu=0.2, sigma=ax, meta={"e": 0.3, "p": -0.3})
    done = {r.rid: r for r in srv.run([(0.0, Request(rid=1, spec=bad, arrival=0.0)),
                                       (0.0, Request(rid=2, spec=spec(32, 0.2), arrival=0.0))])}
    assert "Zel'dovich axes undefined" in done[1].error and done[1].result is None
    assert done[2].error is None and done[2].result.crossed
    assert StubPM.live == 2  # one slot and its engine: the invalid request allocated nothing
    assert done[1].rejected and not done[2].rejected
    s = summarize(list(done.values()))
    assert (s["completed"], s["rejected_invalid"], s["failed"]) == (1, 1, 0)
