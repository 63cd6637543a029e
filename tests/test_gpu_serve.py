"""Continuous-batching server (pbh_amd.serve): every request's result equals the same request run
alone, requests that arrive while others run are admitted as slots free up, and bisection chains
feed probes back into the queue.  Calls go through the C-ABI of _pm3d_hip.so."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def test_served_results_equal_alone_and_slots_are_reused():
    from pbh_amd import ic
    from pbh_amd.collapse import run_to_crossing
    from pbh_amd.pm3d import PM3D
    from pbh_amd.serve import CollapseServer, Request, config_requests, poisson_arrivals
    say("serve vs alone")
    specs = config_requests(6, N=32, seed0=500) + config_requests(4, N=48, seed0=600)
    times = poisson_arrivals(400.0, len(specs), seed=3)
    # a small HBM budget forces waiting for free slots
    one = PM3D(32, 1.0, 32, precision="mixed").device_bytes()
    srv = CollapseServer(precision="mixed", hbm_budget=12 * one + 2 * PM3D(48, 1.0, 48).device_bytes(), poll=4)
    reqs = [(float(t), Request(rid=k, spec=s, arrival=float(t))) for k, (t, s) in enumerate(zip(times, specs))]
    done = {r.rid: r for r in srv.run(reqs)}
    assert len(done) == len(specs)
    assert srv.held <= srv.budget
    for k, s in enumerate(specs):
        pm = ic.device_ic(PM3D(s.N, 1.0, s.N, precision="mixed"), s)
        alone = run_to_crossing(pm, s.a_i, 0.02, 1.0)
        got = done[k].result
        assert got.crossed == alone.crossed and got.steps == alone.steps
        assert got.t_c == alone.t_c
        assert done[k].t_done >= done[k].t_admit >= done[k].arrival


def test_bisection_chain_converges():
    from pbh_amd.serve import BisectionChain, CollapseServer, Request
    say("bisection chain")
    ch = BisectionChain(e=0.1, p=0.0, N=32, a_target=0.2, probes=6, seed=9)
    srv = CollapseServer(precision="mixed", poll=4)
    done = srv.run([(0.0, Request(rid=1, spec=ch.spec(), arrival=0.0, chain=ch))])
    assert len(done) == 6 and ch.done == 6
    # crossing before a_target is monotone in mu: every probe above the final threshold crossed
    for mu, crossed, a_sc in ch.history:
        assert crossed == (mu >= ch.hi)
    assert ch.lo < ch.hi
