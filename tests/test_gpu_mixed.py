"""Mixed-precision mode (int32 cell + FP32 offset, FP32 momenta/FFT/forces, int64 fixed-point
deposit, FP64 scalars) against the FP64 oracle, and the first-shell-crossing time against the
oracle's.  Every call goes through the C-ABI of _pm3d_hip.so.

Tolerances: density max|drho|/max|rho| <= 1e-5 (the mixed-precision density bound); positions
(minimum image, / L) and momenta <= 1e-5 after 20 steps; collapse time |dt_c| <= 1e-4 t_c.
"""
import sys

import numpy as np
import pytest

from oracle import pm3d_oracle as O

pytestmark = pytest.mark.gpu


class OracleStepper:
    """The CPU oracle behind the step/min_jacobian interface that run_to_crossing drives."""

    def __init__(self, oracle_pm):
        self.pm = oracle_pm

    def step(self, a, da):
        return self.pm.step(a, da)

    def min_jacobian(self):
        return O.min_jacobian(self.pm.x, self.pm.Np, self.pm.L)


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b)
    return float(np.abs(np.asarray(a) - b).max() / np.abs(b).max())


def pos_err(x, y, L):
    d = (np.asarray(x) - np.asarray(y) + L / 2) % L - L / 2
    return float(np.abs(d).max() / L)


def test_mixed_state_roundtrip_and_keys():
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("mixed state roundtrip")
    spec = ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    pm = PM3D(64, 1.0, 64, precision="mixed")
    pm.set_state(st["x"], st["p"])
    assert pos_err(pm.x, st["x"], 1.0) <= 2 ** -24 / 64 * 1.01
    assert rel(pm.p, st["p"]) <= 1e-7
    # The stored base cell is the oracle's CIC base index i0 = floor(x/dx - 0.5) of the FP64 input
    # (pm3d.py:77-79), except where the fraction f = g - i0 rounds to 1.0 in FP32: then the cell is
    # i0 + 1 with offset 0.  Reading x back gives (cell + 0.5 + off) * dx exactly (dx = 2^-6), so
    # the oracle's index of the returned positions is the stored cell itself.
    xb = pm.x
    n_up = 0
    for d in range(3):
        lo, hi, f = O.cic_index(st["x"][:, d], 1.0 / 64, 64)
        up = f.astype(np.float32) >= np.float32(1.0)
        n_up += int(up.sum())
        assert np.all(f[up] > 1.0 - 2.0 ** -24)
        expect = np.where(up, hi, lo)
        got = O.cic_index(xb[:, d], 1.0 / 64, 64)[0]
        np.testing.assert_array_equal(got, expect)
    assert n_up <= 1e-4 * st["x"].size
    # and the flat key of the returned positions matches the oracle's key wherever no axis rounded up
    k_dev, k_ref = O.cic_key(xb, 1.0 / 64, 64), O.cic_key(st["x"], 1.0 / 64, 64)
    assert int((k_dev != k_ref).sum()) <= n_up


@pytest.mark.parametrize("N", [32, 64])
def test_mixed_forces_vs_oracle(N):
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("mixed forces", N)
    spec = ic.PeakSpec(N=N, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    pm = PM3D(N, 1.0, N, precision="mixed")
    pm.set_state(st["x"], st["p"])
    ref = O.OraclePM3D(N, 1.0, N)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    acc, delta = pm.forces(0.01)
    acc_r, delta_r = ref.forces(0.01)
    rho, rho_r = delta + 1.0, delta_r + 1.0
    assert rel(rho, rho_r) <= 1e-5
    assert rel(acc, acc_r) <= 1e-4


def test_mixed_trajectory_vs_oracle():
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("mixed trajectory 64^3, 20 steps")
    spec = ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    pm = PM3D(64, 1.0, 64, precision="mixed")
    pm.set_state(st["x"], st["p"])
    ref = O.OraclePM3D(64, 1.0, 64)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    a = 0.01
    for _ in range(20):
        a_g, dm = pm.step(a, 0.02 * a)
        a, dr = ref.step(a, 0.02 * a)
    assert rel(np.asarray(dm) + 1.0, dr + 1.0) <= 1e-5
    assert pos_err(pm.x, ref.x, 1.0) <= 1e-5
    assert rel(pm.p, ref.p) <= 1e-5
    assert dm.max() == pytest.approx(dr.max(), rel=1e-5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_collapse_time_vs_oracle(precision):
    """First shell crossing of a 32^3 triaxial peak: t_c within 1e-4 t_c of the FP64 oracle's."""
    from pbh_amd import ic
    from pbh_amd.collapse import run_to_crossing
    from pbh_amd.pm3d import PM3D
    say("collapse time", precision)
    spec = ic.PeakSpec(N=32, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    pm = PM3D(32, 1.0, 32, precision=precision)
    pm.set_state(st["x"], st["p"])
    ref = O.OraclePM3D(32, 1.0, 32)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    r_gpu = run_to_crossing(pm, 0.01, 0.02, 1.0)
    r_ref = run_to_crossing(OracleStepper(ref), 0.01, 0.02, 1.0)
    assert r_ref.crossed and r_gpu.crossed
    assert r_gpu.steps == r_ref.steps
    assert abs(r_gpu.t_c - r_ref.t_c) <= 1e-4 * r_ref.t_c
    if precision == "fp64":
        assert abs(r_gpu.t_c - r_ref.t_c) <= 1e-9 * r_ref.t_c


def test_mixed_vs_fp64_at_256():
    """The mixed bench workload at its own size (peak256: 256^3 mesh and particles, random field):
    from one initial state (built on the device), three mixed steps agree with three FP64 steps
    (the FP64 path is pinned to the oracle at the sizes the oracle runs in seconds) within the
    mixed-precision bounds, and the fixed-point deposit conserves mass."""
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D
    say("mixed vs fp64 at 256^3")
    spec = ic.named("peak256")
    pm64 = PM3D(spec.N, spec.L, spec.N, precision="fp64")
    ic.device_ic(pm64, spec)
    pmx = PM3D(spec.N, spec.L, spec.N, precision="mixed")
    ic.device_ic(pmx, spec, engine=pm64)
    a = spec.a_i
    for _ in range(3):
        a_next, d64 = pm64.step(a, 0.0125 * a)
        _, dmx = pmx.step(a, 0.0125 * a)
        a = a_next
    r64, rmx = np.asarray(d64) + 1.0, np.asarray(dmx) + 1.0
    say("256^3 mixed vs fp64: rho %.2e  x %.2e  p %.2e  mass %.1e" % (
        rel(rmx, r64), pos_err(pmx.x, pm64.x, spec.L), rel(pmx.p, pm64.p), abs(rmx.mean() - 1.0)))
    assert rel(rmx, r64) <= 1e-5
    assert abs(rmx.mean() - 1.0) <= 1e-8 and abs(r64.mean() - 1.0) <= 1e-12
    assert pos_err(pmx.x, pm64.x, spec.L) <= 1e-5
    assert rel(pmx.p, pm64.p) <= 1e-5


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("N,Np", [(32, 32), (32, 28), (32, 30), (48, 48)])
def test_mixed_min_jacobian_matches_oracle_on_state(N, Np, precision):
    """min det(dx/dq) of the state against the oracle's on the positions read back from it
    (Np % 4 == 0: four lattice sites per thread; Np = 28 also covers rows of 7 quads across
    waves; otherwise one site per thread).  FP64: the four-site kernel's division-free entries and
    the one-site kernel's IEEE quotients give the oracle's minimum bit for bit."""
    from pbh_amd.pm3d import PM3D
    say(f"{precision} min_jacobian N={N} Np={Np}")
    L = 1.0
    q = (np.arange(Np) + 0.5) * L / Np
    X = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(7)
    amp = 0.45 * L / Np  # displacements up to about half a lattice spacing: folds somewhere
    X = (X + amp * np.sin(2 * np.pi * X[:, [1, 2, 0]] / L) + amp * 0.3 * rng.standard_normal(X.shape)) % L
    pm = PM3D(N, L, Np, precision=precision)
    pm.set_state(X, np.zeros_like(X))
    got = pm.min_jacobian()
    ref = O.min_jacobian(pm.x, Np, L)
    say(f"  gpu {got:.17e} oracle {ref:.17e}")
    if precision == "fp64":
        assert got == ref
    else:
        assert got == pytest.approx(ref, rel=1e-9, abs=1e-12)
    pm.close()
