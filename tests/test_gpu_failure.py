"""Failure detection on the device.  The reference's native kernels return a count of bad cells or
shells (pbhgr/csrc/kernels.c:36,73,105,166); here every particle kernel counts the particles whose
momentum or position became non-finite in a sticky device counter, returned by pm3d_synchronize
(PM3D.synchronize).  A non-finite particle keeps its box memory-safe (its cell and row-list row
are clamped to the mesh; the weights it deposits are garbage, not out-of-bounds writes), it is
counted, and boxes stepped beside it are unaffected.  A served request whose state turns
non-finite completes as failed while the others keep the results they get alone.  Every call goes
through the C-ABI of _pm3d_hip.so."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("what", ["x", "p"])
def test_nonfinite_particle_is_counted_and_isolated(prec, what):
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("non-finite particle", prec, what)
    spec = ic.PeakSpec(N=32, mu=0.3, sigma=(0.12, 0.11, 0.10), seed=0)
    st = ic.make(spec)
    x_bad, p_bad = st["x"].copy(), st["p"].copy()
    if what == "x":
        x_bad[12345, 1] = np.nan
    else:
        p_bad[777, 2] = np.inf
    boxes = {k: PM3D(32, 1.0, 32, precision=prec) for k in ("clean", "bad", "alone")}
    boxes["clean"].set_state(st["x"], st["p"])
    boxes["bad"].set_state(x_bad, p_bad)
    boxes["alone"].set_state(st["x"], st["p"])
    for name in ("clean", "bad", "alone"):
        assert boxes[name].synchronize() == 0  # setting a state counts nothing
    a = st["a_i"]
    for _ in range(3):
        da = 0.02 * a
        for name in ("clean", "bad"):  # queued side by side
            boxes[name].step(a, da)
        a_new, _ = boxes["alone"].step(a, da)
        a = a_new
    nbad = boxes["bad"].synchronize()
    say(f"  flagged {nbad} particle updates; clean box flagged {boxes['clean'].synchronize()}")
    assert nbad > 0
    assert boxes["clean"].synchronize() == 0
    assert np.array_equal(boxes["clean"].x, boxes["alone"].x)
    assert np.array_equal(boxes["clean"].p, boxes["alone"].p)
    for b in boxes.values():
        b.close()


def test_served_request_with_nonfinite_state_fails_alone():
    """Three 32^3 requests; the middle one's amplitude (1e300, finite, so admitted) overflows its
    initial displacement.  It completes as failed; the other two equal their runs alone."""
    from pbh_amd import ic
    from pbh_amd.collapse import run_to_crossing
    from pbh_amd.pm3d import PM3D
    from pbh_amd.serve import CollapseServer, Request, summarize
    say("served request with a non-finite state")
    specs = [ic.PeakSpec(N=32, mu=mu, sigma=(0.1, 0.09, 0.08), seed=40 + k, noise_rms_frac=0.01)
             for k, mu in enumerate((0.3, 1e300, 0.25))]
    srv = CollapseServer(precision="mixed", poll=4)
    done = {r.rid: r for r in srv.run([(0.0, Request(rid=k, spec=s, arrival=0.0)) for k, s in enumerate(specs)])}
    assert len(done) == 3
    say(f"  request 1: {done[1].error}")
    assert done[1].error is not None and done[1].result is None and not done[1].rejected
    # its schedule ended on the device at the first step with a flagged particle: far fewer flagged
    # updates than a whole schedule of 32^3 particles (about 230 steps)
    assert int(done[1].error.split()[0]) <= 2 * 32 ** 3
    for k in (0, 2):
        pm = ic.device_ic(PM3D(32, 1.0, 32, precision="mixed"), specs[k])
        alone = run_to_crossing(pm, specs[k].a_i, 0.02, 1.0)
        got = done[k].result
        assert done[k].error is None
        assert (got.crossed, got.steps, got.t_c) == (alone.crossed, alone.steps, alone.t_c)
        pm.close()
    s = summarize(list(done.values()))
    assert (s["completed"], s["failed"], s["rejected_invalid"]) == (2, 1, 0)
