"""Batched, graph-replayed stepping (pbh_amd.batch) and the mixed-radix FFT sizes it needs
(N = 48, 96).  Every call goes through the C-ABI of _pm3d_hip.so.

A box must give bitwise the same results alone (PM3D.step) and in a batch (device schedule,
one hipGraph replay per step for all boxes)."""
import sys

import numpy as np
import pytest

from oracle import pm3d_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b)
    return float(np.abs(np.asarray(a) - b).max() / np.abs(b).max())


@pytest.mark.parametrize("N", [48, 96])
def test_mixed_radix_fft_parity(N):
    from pbh_amd.pm3d import PM3D
    say("mixed radix fft", N)
    pm = PM3D(N, 1.0, 8)
    rng = np.random.default_rng(N)
    a = rng.standard_normal((N, N, N))
    assert rel(pm.rfftn(a), np.fft.rfftn(a)) <= 1e-13
    c = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    assert rel(pm.irfftn(c), np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))) <= 1e-13
    ref = O.OraclePM3D(N, 1.0, 8, smooth=1.0)
    pm2 = PM3D(N, 1.0, 8, smooth=1.0)
    g = pm2.poisson_gradient(a, 1.3)
    gr = ref.gradient_meshes(ref.potential_k(a, 1.3))
    for c in range(3):
        assert rel(g[..., c], gr[c]) <= 1e-12


def test_trajectory_n96_vs_oracle():
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("trajectory 96^3")
    spec = ic.PeakSpec(N=96, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    pm = PM3D(96, 1.0, 96)
    pm.set_state(st["x"], st["p"])
    ref = O.OraclePM3D(96, 1.0, 96)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    a = 0.01
    for _ in range(3):
        pm.step(a, 0.02 * a)
        a, dr = ref.step(a, 0.02 * a)
    d = (pm.x - ref.x + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() <= 1e-10
    assert rel(pm.p, ref.p) <= 1e-10


def _members(seeds, precision, da_frac=0.02):
    from pbh_amd import ic
    from pbh_amd.batch import Member
    from pbh_amd.pm3d import PM3D
    out = []
    for sd in seeds:
        spec = ic.batch_spec(sd)
        st = ic.make(spec)
        pm = PM3D(spec.N, 1.0, spec.N, precision=precision)
        pm.set_state(st["x"], st["p"])
        out.append(Member(pm, st["a_i"], da_frac, 1.0, tag=spec.name))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_state_equals_alone(precision):
    """K steps of four boxes in one graph == the same K steps of each box alone, bitwise."""
    from pbh_amd.batch import Batch
    say("batch == alone", precision)
    seeds = [100, 101, 102, 103]
    alone = _members(seeds, precision)
    mix = _members(seeds, precision)
    K = 6
    for m in alone:
        a = m.a0
        for _ in range(K):
            a, _ = m.pm.step(a, m.da_frac * a)
        m.a_end = a
    b = Batch()
    for m in mix:
        a1, _ = m.pm.step(m.a0, m.da_frac * m.a0)
        ents, avals = m.pm.schedule_entries(a1, m.da_frac, 10.0, K - 1, m.pm.last_kick)
        m.pm.set_schedule(ents, len(avals))
        m.a_end = avals[-1]
    b.set([m.pm for m in mix])
    b.run(K - 1)
    b.synchronize()
    for u, v in zip(alone, mix):
        assert v.pm.schedule_end(v.a_end) == K - 1
        assert u.a_end == v.a_end
        assert np.array_equal(u.pm.x, v.pm.x), u.tag
        assert np.array_equal(u.pm.p, v.pm.p), u.tag
    # the state continues normally after the schedule
    for u, v in zip(alone, mix):
        u.pm.step(u.a_end, 0.01)
        v.pm.step(v.a_end, 0.01)
        assert np.array_equal(u.pm.x, v.pm.x)


@pytest.mark.parametrize("steps_per_graph", [1, 8])
def test_batched_collapse_equals_plain(steps_per_graph):
    """First shell crossing of boxes of mesh 32-128 (mixed): batched == plain, bitwise t_c, with one
    step or eight steps per graph launch."""
    from pbh_amd.batch import Batch, run_batched, run_plain
    say("batched collapse == plain, steps per graph", steps_per_graph)
    seeds = [100, 104, 108, 112, 116, 120]
    rp = run_plain(_members(seeds, "mixed"))
    rb = run_batched(_members(seeds, "mixed"), poll=8, batch=Batch(steps_per_graph=steps_per_graph))
    for x, y in zip(rp, rb):
        assert x.crossed and y.crossed
        assert x.steps == y.steps
        assert x.t_c == y.t_c


@pytest.mark.parametrize("N,noise", [(32, 0.01), (64, 0.0), (96, 0.02)])
def test_device_ic_matches_host(N, noise):
    """pm3d_make_peak_ic (field, random stream, FFT displacement, lattice) == pbh_amd.ic.make."""
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("device ic", N, noise)
    spec = ic.PeakSpec(N=N, mu=0.2, sigma=(0.1, 0.09, 0.08), seed=5, noise_rms_frac=noise)
    st = ic.make(spec)
    pm = ic.device_ic(PM3D(N, 1.0, N), spec)
    d = (pm.x - st["x"] + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() <= 1e-12
    assert rel(pm.p, st["p"]) <= 1e-11
    mx = ic.device_ic(PM3D(N, 1.0, N, precision="mixed"), spec)
    d = (mx.x - st["x"] + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() <= 1.0 / N * 2 ** -23
