"""Device analyse (pm3d.py:132-159) against the oracle's host analyse on the same state.

Centre, enclosed counts: exact (same expressions, same order); set axes / hoop: rtol 1e-9 (the
reference's r.T @ r goes through BLAS, whose summation order is its own); density maximum: rtol
1e-10 in FP64, 1e-5 in the mixed mode (FP32 CIC weights).
"""
import sys

import numpy as np
import pytest

from oracle import pm3d_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _compare(out, ref, rho_rtol):
    for key, val in ref.items():
        if key.startswith("m_in_") or key.startswith("M_") or key == "a":
            assert out[key] == val, key
        elif key == "rho_max_over_mean":
            assert out[key] == pytest.approx(val, rel=rho_rtol), key
        else:
            np.testing.assert_allclose(np.asarray(out[key]), np.asarray(val), rtol=1e-9, atol=1e-14, err_msg=key)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_analyse_driver_config(prec):
    """A driver-shaped run (64^3 mesh and particles, L = 8 r_m, mu = 0.3, e = 0.1), 25 steps."""
    from pbh_amd.ltb import LTBGaussian, profile_table
    from pbh_amd.pm3d import PM3D, analyse
    say("analyse driver config", prec)
    rm = LTBGaussian(0.3).r_m
    spline = profile_table("yoo")["spline"]
    pm = PM3D(64, 8.0 * rm, 64, smooth=2.0, precision=prec).make_ic(0.3, spline, 0.1, 0.0, 0.0, (2.0, 10.0))
    a = 1.0
    for _ in range(25):
        a, dm = pm.step(a, 0.05 * a)
    before = np.asarray(dm).copy() if prec == "fp64" else None
    out = analyse(pm, a)
    ref_pm = O.OraclePM3D(64, 8.0 * rm, 64, smooth=2.0).make_ic(0.3, spline, 0.1, 0.0, 0.0, (2.0, 10.0))
    assert np.array_equal(ref_pm.r_ell_q, pm.r_ell_q)
    ref_pm.x = pm.x
    ref = O.analyse(ref_pm, a)
    _compare(out, ref, 1e-10 if prec == "fp64" else 1e-5)
    if before is not None:  # the step's density mesh is still on the device, unchanged
        from pbh_amd.pm3d import DeltaMesh
        assert np.array_equal(np.asarray(DeltaMesh(pm, dm._serial)), before)


def test_analyse_large_sets():
    """Sets of 10^4-10^5 particles (several blocks per set) on a 128^3 peak, with synthetic
    Lagrangian radii (distance of the lattice site from the box centre)."""
    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D, analyse
    say("analyse large sets")
    spec = ic.named("peak128")
    st = ic.make(spec)
    pm = PM3D(spec.N, spec.L, spec.n_side)
    pm.set_state(st["x"], st["p"])
    a = st["a_i"]
    for _ in range(5):
        a, _ = pm.step(a, 0.0125 * a)
    q = (np.arange(spec.n_side) + 0.5) / spec.n_side - 0.5
    Q = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    pm.r_ell_q = np.sqrt(np.sum((Q / np.array([1.0, 0.9, 0.8])) ** 2, axis=1)) * 2.0
    out = analyse(pm, a, fracs=(0.1, 0.2, 0.3), radii_cells=(1.0, 2.0, 4.0, 8.0, 16.0))
    assert len(pm.analysis_sets((0.1, 0.2, 0.3))[3]) > 50000
    ref_pm = O.OraclePM3D(spec.N, spec.L, spec.n_side)
    ref_pm.x = pm.x
    ref_pm.r_ell_q = pm.r_ell_q
    ref = O.analyse(ref_pm, a, fracs=(0.1, 0.2, 0.3), radii_cells=(1.0, 2.0, 4.0, 8.0, 16.0))
    _compare(out, ref, 1e-10)
