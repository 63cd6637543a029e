"""Serving benchmark: Poisson arrivals of collapse requests, p50/p99 time-to-collapse latency and
solves/s, next to plain FP64 one-at-a-time GPU runs of the same requests.

    python -m pbh_amd.bench_serve [--N 256] [--rates 5,20,50,200] [--n 16]
"""
from __future__ import annotations

import argparse
import json
import time

from .collapse import run_to_crossing
from .pm3d import PM3D
from . import ic
from .serve import BisectionChain, CollapseServer, Request, config_requests, poisson_arrivals, summarize


def _valid(spec) -> bool:
    try:
        ic.validate(spec)
        return True
    except ValueError:
        return False


def plain_fp64(specs):
    """One request at a time, FP64, PM3D.step + min det(dx/dq) per step (initial conditions on the device)."""
    t0 = time.perf_counter()
    out = []
    eng = {}
    specs = [s for s in specs if _valid(s)]
    for s in specs:
        pm = ic.device_ic(PM3D(s.N, 1.0, s.N, precision="fp64"), s)
        out.append(run_to_crossing(pm, s.a_i, 0.02, 1.0))
        del pm
    dt = time.perf_counter() - t0
    return {"requests": len(specs), "wall_s": dt, "solves_per_s": len(specs) / dt}


def sweep(N=256, rates=(5, 20, 50, 200), n=16, precision="mixed", poll=8):
    rows = []
    for r in rates:
        srv = CollapseServer(precision=precision, poll=poll)
        specs = config_requests(n, N=N, seed0=1000)
        times = poisson_arrivals(r, n, seed=int(r))
        reqs = [(float(t), Request(rid=k, spec=s, arrival=float(t))) for k, (t, s) in enumerate(zip(times, specs))]
        done = srv.run(reqs)
        row = summarize(done)
        row.update(rate_req_per_s=r, max_active=srv.max_active, graph_builds=srv.graph_builds,
                   held_GB=srv.held / 1e9)
        rows.append(row)
        print(json.dumps(row), flush=True)
    return rows


def chains(N=256, precision="mixed", probes=8, a_target=0.3):
    srv = CollapseServer(precision=precision)
    reqs = []
    chs = []
    for k, e in enumerate((0.0, 0.1, 0.2)):
        ch = BisectionChain(e=e, p=0.0, N=N, a_target=a_target, probes=probes, seed=2000 + k)
        chs.append(ch)
        reqs.append((0.0, Request(rid=k + 1, spec=ch.spec(), arrival=0.0, chain=ch)))
    t0 = time.perf_counter()
    done = srv.run(reqs)
    dt = time.perf_counter() - t0
    out = summarize(done)
    out.update(wall_s=dt, mu_threshold={f"e={c.e:g}": c.mu_threshold for c in chs}, probes=probes, a_target=a_target)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--rates", default="5,20,50,200")
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--plain", type=int, default=4)
    ap.add_argument("--chains", action="store_true")
    ap.add_argument("--poll", type=int, default=8, help="graph replays between admissions and status reads")
    args = ap.parse_args()
    rates = [float(x) for x in args.rates.split(",")]
    res = {"N": args.N, "poll": args.poll, "sweep": sweep(args.N, rates, args.n, poll=args.poll)}
    if args.plain:
        res["plain_fp64"] = plain_fp64(config_requests(args.plain, N=args.N, seed0=1000))
        print(json.dumps(res["plain_fp64"]), flush=True)
    if args.chains:
        res["chains"] = chains(args.N)
        print(json.dumps(res["chains"]), flush=True)
    print(json.dumps(res))
