"""Many independent collapse runs stepped together on one GPU.

The reference runs independent parameter points as separate processes (pbhgr/pbhgr/scan.py:67-76,
a multiprocessing Pool over spherical runs); pm3d itself runs one box per process
(pbhgr/scripts/pm3d_peak.py).  Here every box (PM3D handle, any mesh size, FP64 or mixed) keeps its
state in HBM with a device-resident step schedule, and one captured hipGraph holds one step of every
active box as independent branches; a replay advances all of them.  Each box records min det(dx/dq)
per step on the device and ends its own schedule at its first crossing; every `poll` replays the host
reads the status of all boxes in one device pass, retires those whose schedule ended and re-captures
the graph for the rest.

A box gives bitwise the same trajectory in a batch as alone: the kernels and the host scalars are
the same (tests/test_gpu_batch.py).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .collapse import CollapseResult, crossing_a
from .pm3d import T_I, PM3D


@dataclass
class Member:
    pm: PM3D
    a0: float
    da_frac: float
    a_max: float
    tag: str = ""
    # filled by the runner
    a_vals: list = field(default_factory=list)   # scale factor after each step (step 1 = plain step)
    jhist: list = field(default_factory=list)    # min det(dx/dq) after each step
    j0: float = 1.0
    result: CollapseResult | None = None
    t_done: float = 0.0


class Batch:
    """ctypes wrapper of the pm3d_batch_* entry points."""

    def __init__(self, steps_per_graph: int = 8):
        """steps_per_graph: scheduled steps one graph launch advances.  Each mesh-size class runs
        its steps of a launch back to back, without waiting for the other classes between steps
        (64-peak set: 244-252 solves/s at 1, 257-262 at 2-16, 257 at 32)."""
        self._L = _lib.load()
        b = ctypes.c_void_p()
        _lib.check(self._L.pm3d_batch_create(ctypes.byref(b)))
        self._b = b
        self.steps_per_graph = steps_per_graph
        self._chk(self._L.pm3d_batch_set_steps(self._b, steps_per_graph))

    def _chk(self, rc):
        if rc < 0:
            raise _lib.Pm3dError(f"pm3d batch: {self._L.pm3d_batch_last_error(self._b).decode()}")
        return rc

    def set(self, pms):
        arr = (ctypes.c_void_p * max(len(pms), 1))(*[pm._h.value for pm in pms])
        self._chk(self._L.pm3d_batch_set(self._b, arr, len(pms)))

    def run(self, nsteps: int):
        self._chk(self._L.pm3d_batch_run(self._b, nsteps))

    def synchronize(self):
        self._chk(self._L.pm3d_batch_synchronize(self._b))

    def status(self, count: int):
        """(done, nsteps, jmin) of every member, in the order of set(), after the queued replays."""
        out = (_lib.Pm3dMemberStatus * max(count, 1))()
        self._chk(self._L.pm3d_batch_status(self._b, out, count))
        return [(out[i].done, out[i].nsteps, out[i].jmin) for i in range(count)]

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            try:
                self._L.pm3d_batch_destroy(b)
            except Exception:
                pass


def _finish(m: Member, crossed_at: int | None):
    """crossed_at: index into m.jhist of the first value <= 0."""
    if crossed_at is None:
        a_last = m.a_vals[-1] if m.a_vals else m.a0
        m.result = CollapseResult(False, float("nan"), float("nan"), len(m.jhist), a_last,
                                  m.jhist[-1] if m.jhist else m.j0, list(zip(m.a_vals, m.jhist)))
        return
    j_prev = m.j0 if crossed_at == 0 else m.jhist[crossed_at - 1]
    a_prev = m.a0 if crossed_at == 0 else m.a_vals[crossed_at - 1]
    a_sc = crossing_a(a_prev, j_prev, m.a_vals[crossed_at], m.jhist[crossed_at])
    m.result = CollapseResult(True, a_sc, T_I * a_sc ** 1.5, crossed_at + 1, m.a_vals[crossed_at],
                              m.jhist[crossed_at], list(zip(m.a_vals[: crossed_at + 1], m.jhist[: crossed_at + 1])))


def run_batched(members: list[Member], max_steps: int = 4000, poll: int = 32, batch: Batch | None = None):
    """Step every member to its first shell crossing (or a_max) in one graph-replayed batch.  A member's
    schedule ends on the device at its crossing (its later steps do nothing), so `poll` replays run
    between host status reads without wasted work; a status read is one device pass over all members
    (Batch.status), and a member's history is read once, when its table has ended."""
    batch = batch or Batch()
    t0 = time.perf_counter()
    # the first (plain) step of every member is queued before any result is read, so the members'
    # streams run their first steps side by side
    for m in members:
        m.j0 = m.pm.min_jacobian()
    for m in members:
        da = min(m.da_frac * m.a0, m.a_max - m.a0)
        a1, _ = m.pm.step(m.a0, da)
        m.a_vals = [a1]
    active = []
    for m in members:
        m.jhist = [m.pm.min_jacobian()]
        if m.jhist[0] <= 0.0:
            _finish(m, 0)
            continue
        ents, avals = m.pm.schedule_entries(m.a_vals[0], m.da_frac, m.a_max, max_steps, m.pm.last_kick)
        if not avals:
            _finish(m, None)
            continue
        m.pm.set_schedule(ents, len(avals))
        m._sched_a = avals
        active.append(m)
    graph = list(active)  # members of the captured graph, in its order
    batch.set([m.pm for m in graph])
    while active:
        batch.run(poll)
        st = batch.status(len(graph))
        where = {id(m): k for k, m in enumerate(graph)}
        still = []
        for m in active:
            done, nsteps, _ = st[where[id(m)]]
            if done < nsteps:
                still.append(m)
                continue
            # the table ended: at the first crossing (the device shortened it) or at its last entry
            done, hist = m.pm.schedule_status()
            m.jhist = [m.jhist[0]] + list(hist)
            m.a_vals = [m.a_vals[0]] + m._sched_a[:done]
            neg = np.nonzero(np.asarray(m.jhist) <= 0.0)[0]
            _finish(m, int(neg[0]) if len(neg) else None)
            m.t_done = time.perf_counter() - t0
        # a finished member's steps do nothing on the device (its schedule ends at the crossing), so
        # the graph is rebuilt without them only once a quarter of its members have finished
        if still and len(graph) - len(still) >= max(1, len(graph) // 4):
            graph = list(still)
            batch.set([m.pm for m in graph])
        active = still
    for m in members:
        if getattr(m, "_sched_a", None) is not None:
            m.pm.schedule_end(m.a_vals[-1])
    return [m.result for m in members]


def run_plain(members: list[Member], max_steps: int = 4000):
    """The single-run GPU path: each member alone, PM3D.step + min_jacobian per step."""
    from .collapse import run_to_crossing
    out = []
    for m in members:
        out.append(run_to_crossing(m.pm, m.a0, m.da_frac, m.a_max, max_steps=max_steps))
    return out
