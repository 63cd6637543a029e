"""Batched collapse solves/s against the plain single-run GPU path.

Workload: one triaxial peak per seed (100..163 by default; pbh_amd.ic.batch_spec), mesh sizes
drawn from {32, 64, 96, 128}^3 with Np = N, mixed precision, initial conditions built on the GPU,
each run from a_i = 0.01 with da = 0.02 a to its first shell crossing (pbh_amd.collapse).

* plain  : one box at a time, PM3D.step + min det(dx/dq) per step (the driver-loop pattern of
           pbhgr/scripts/pm3d_peak.py:26-33 on the GPU path);
* batched: all boxes in one captured hipGraph per step (pbh_amd.batch.run_batched).

Both include only the stepping to collapse (handles and initial conditions are built first).
"""
from __future__ import annotations

import time

from . import ic
from .batch import Member, run_batched
from .collapse import run_to_crossing
from .pm3d import PM3D


def _members(seeds, precision="mixed", da_frac=0.02, engines=None):
    engines = engines if engines is not None else {}
    out = []
    for sd in seeds:
        spec = ic.batch_spec(sd)
        if spec.N not in engines:
            engines[spec.N] = PM3D(spec.N, 1.0, spec.N, precision="fp64")
        pm = ic.device_ic(PM3D(spec.N, 1.0, spec.N, precision=precision), spec, engine=engines[spec.N])
        out.append(Member(pm, spec.a_i, da_frac, 1.0, tag=spec.name))
    return out, engines


def batched_vs_plain(seeds=range(100, 164), precision="mixed", poll=32):
    seeds = list(seeds)
    engines = {}
    plain, engines = _members(seeds, precision, engines=engines)
    t0 = time.perf_counter()
    rp = [run_to_crossing(m.pm, m.a0, m.da_frac, m.a_max) for m in plain]
    t_plain = time.perf_counter() - t0
    steps_plain = sum(r.steps for r in rp)
    del plain
    batched, engines = _members(seeds, precision, engines=engines)
    t0 = time.perf_counter()
    rb = run_batched(batched, poll=poll)
    t_batch = time.perf_counter() - t0
    same = all(a.crossed == b.crossed and a.steps == b.steps and (a.t_c == b.t_c or not a.crossed)
               for a, b in zip(rp, rb))
    n = len(seeds)
    # plain FP64 one-at-a-time (the reference-precision GPU path)
    p64, engines = _members(seeds, "fp64", engines=engines)
    t0 = time.perf_counter()
    r64 = [run_to_crossing(m.pm, m.a0, m.da_frac, m.a_max) for m in p64]
    t_plain64 = time.perf_counter() - t0
    del p64
    dtc = max(abs(a.t_c - b.t_c) / a.t_c for a, b in zip(r64, rb) if a.crossed and b.crossed)
    return {
        "plain_fp64_s": t_plain64, "plain_fp64_solves_per_s": n / t_plain64,
        "speedup_vs_plain_fp64": t_plain64 / t_batch, "max_rel_t_c_mixed_vs_fp64": dtc,
        "members": n, "precision": precision, "poll_steps": poll,
        "plain_s": t_plain, "batched_s": t_batch,
        "plain_solves_per_s": n / t_plain, "batched_solves_per_s": n / t_batch,
        "speedup": t_plain / t_batch,
        "crossed": sum(r.crossed for r in rb), "steps_total": steps_plain,
        "identical_results": bool(same),
        "t_c_range": [min(r.t_c for r in rb if r.crossed), max(r.t_c for r in rb if r.crossed)],
    }


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--poll", type=int, default=32, help="graph replays between host status reads")
    print(json.dumps(batched_vs_plain(poll=ap.parse_args().poll)))
