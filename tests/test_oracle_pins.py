"""Pins of the CPU oracle (oracle/pm3d_oracle.py) by closed forms, invariants, brute force and the
reference's own outputs (golden vectors written by tests/golden/make_golden.py).  CPU only.

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong sign or
index, a transposed operand) fails it.
"""
import itertools

import numpy as np
import pytest

from oracle import pm3d_oracle as O


# ---------------------------------------------------------------- cic_index (pm3d.py:76-79)
def test_cic_index_closed_form_table():
    N, L = 8, 1.0
    dx = L / N
    X = np.array([0.0, 0.5 * dx - 1e-12, 0.5 * dx, 3.5 * dx, 3.75 * dx, L - 1e-12, L])
    lo, hi, f = O.cic_index(X, dx, N)
    # x < dx/2 wraps to the last cell with f in [1/2, 1); a cell centre gives f = 0; x == L gives
    # i0 = N-1, i1 = 0, f = 1/2
    assert lo.tolist() == [7, 7, 0, 3, 3, 7, 7]
    assert hi.tolist() == [0, 0, 1, 4, 4, 0, 0]
    np.testing.assert_allclose(f, [0.5, 1.0 - 1e-12 / dx, 0.0, 0.0, 0.25, 0.5 - 1e-12 / dx, 0.5], atol=1e-9)


def test_cic_key_brute_force():
    rng = np.random.default_rng(3)
    N, L = 6, 2.5
    dx = L / N
    X = rng.uniform(0, L, (200, 3))
    key = O.cic_key(X, dx, N)
    for i in range(200):
        base = [int(np.floor(X[i, d] / dx - 0.5)) % N for d in range(3)]
        assert key[i] == (base[0] * N + base[1]) * N + base[2]


# ------------------------------------------------------------------ deposit (pm3d.py:81-89)
def test_deposit_brute_force_n4():
    rng = np.random.default_rng(0)
    N, L = 4, 1.0
    dx = L / N
    X = rng.uniform(0, L, (37, 3))
    rho = np.zeros((N, N, N))
    for x in X:
        g = x / dx - 0.5
        i0 = np.floor(g).astype(int)
        f = g - i0
        for c in itertools.product((0, 1), repeat=3):
            w = np.prod([f[d] if c[d] else 1 - f[d] for d in range(3)])
            rho[tuple((i0[d] + c[d]) % N for d in range(3))] += w
    np.testing.assert_allclose(O.deposit(X, dx, N), rho, rtol=0, atol=1e-14)


def test_deposit_lattice_invariants():
    N, L = 8, 1.0
    dx = L / N
    q = (np.arange(N) + 0.5) * dx
    X = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    assert np.array_equal(O.deposit(X, dx, N), np.ones((N, N, N)))           # Np = N: rho == 1 exactly
    q2 = (np.arange(2 * N) + 0.5) * (L / (2 * N))
    X2 = np.stack(np.meshgrid(q2, q2, q2, indexing="ij"), -1).reshape(-1, 3)
    assert np.array_equal(O.deposit(X2, dx, N), 8 * np.ones((N, N, N)))      # Np = 2N: f in {1/4,3/4}
    rng = np.random.default_rng(1)
    Xr = rng.uniform(0, L, (1000, 3))
    assert abs(O.deposit(Xr, dx, N).sum() - 1000) < 1e-10                    # mass conservation


# ------------------------------------------------------------------- interp (pm3d.py:91-99)
def test_interp_properties():
    rng = np.random.default_rng(2)
    N, L = 8, 1.0
    dx = L / N
    X = rng.uniform(0, L, (300, 3))
    np.testing.assert_allclose(O.interp(np.full((N, N, N), 3.25), X, dx, N), 3.25, rtol=1e-15)
    F = rng.standard_normal((N, N, N))
    centres = (np.array([[1, 2, 3], [7, 0, 5]]) + 0.5) * dx
    np.testing.assert_allclose(O.interp(F, centres, dx, N), [F[1, 2, 3], F[7, 0, 5]], rtol=1e-14)
    # adjointness: sum_p interp(F, X)_p == sum_cells F * deposit(X)
    np.testing.assert_allclose(O.interp(F, X, dx, N).sum(), (F * O.deposit(X, dx, N)).sum(), rtol=1e-12)
    # a field linear in x inside one cell (no wrap) is reproduced exactly
    Fl = np.broadcast_to(np.arange(N, dtype=float)[:, None, None], (N, N, N)).copy()
    Xl = np.array([[2.7 * dx, 0.5, 0.5], [4.1 * dx, 0.3, 0.9]])
    np.testing.assert_allclose(O.interp(Fl, Xl, dx, N), Xl[:, 0] / dx - 0.5, rtol=1e-13)


# --------------------------------------------------- FFT / Poisson / gradient (pm3d.py:101-112)
@pytest.mark.parametrize("smooth", [0.0, 1.5])
def test_single_mode_gradient_closed_form(smooth):
    N, L, a = 16, 2.0, 1.7
    pm = O.OraclePM3D(N, L, N, smooth=smooth)
    dx = L / N
    xc = (np.arange(N) + 0.5) * dx
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
    m = np.array([2, -3, 1])
    k0 = 2 * np.pi * m / L
    eps = 1e-3
    delta = eps * np.cos(k0[0] * X + k0[1] * Y + k0[2] * Z)
    g = pm.gradient_meshes(pm.potential_k(delta, a))
    k2 = float(k0 @ k0)
    filt = np.exp(-0.5 * k2 * (smooth * dx) ** 2) if smooth > 0 else 1.0
    amp = -1.5 * O.H_I ** 2 / a * eps / k2 * filt
    # Phi = amp cos(k.x) -> g_c = -dPhi/dx_c = amp k_c sin(k.x)
    s = np.sin(k0[0] * X + k0[1] * Y + k0[2] * Z)
    for c in range(3):
        np.testing.assert_allclose(g[c], amp * k0[c] * s, atol=1e-14 * abs(amp) * 10 + 1e-16)
        if k0[c] != 0:
            assert np.abs(g[c]).max() > 0.5 * abs(amp * k0[c])


def test_potential_parseval_and_dc():
    rng = np.random.default_rng(5)
    N, L = 8, 1.0
    pm = O.OraclePM3D(N, L, N)
    delta = rng.standard_normal((N, N, N))
    phik = pm.potential_k(delta, 1.0)
    assert phik[0, 0, 0] == 0
    # lap Phi = 1.5 H^2/a delta on the mesh, except for the mean (removed) and the Nyquist planes
    dk = np.fft.rfftn(delta)
    np.testing.assert_allclose(-phik[1:3, 1:3, 1:3] * pm.k2[1:3, 1:3, 1:3], 1.5 * 25.0 * dk[1:3, 1:3, 1:3], rtol=1e-13)


# ----------------------------------------------------------------------- step (pm3d.py:117-129)
def test_step_zero_force_closed_form():
    N, L = 8, 1.0
    pm = O.OraclePM3D(N, L, N)
    q = (np.arange(N) + 0.5) * (L / N)
    pm.x = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    pm.p = np.zeros_like(pm.x)
    a, da = 1.0, 0.05
    a1, delta = pm.step(a, da)
    assert a1 == a + da
    assert np.abs(delta).max() < 1e-15
    assert np.abs(pm.p).max() < 1e-15
    # uniform momentum: drift x += da*p/(a_mid^3 H_mid), no force
    pm2 = O.OraclePM3D(N, L, N)
    p0 = np.array([0.01, -0.02, 0.005])
    x0 = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    pm2.x = x0.copy()
    pm2.p = np.tile(p0, (N ** 3, 1))
    pm2.step(a, da)
    am = a + 0.5 * da
    np.testing.assert_allclose(pm2.x, (x0 + da * p0 / (am ** 3 * O.H_I * am ** -1.5)) % L, rtol=0, atol=1e-15)
    np.testing.assert_allclose(pm2.p, np.tile(p0, (N ** 3, 1)), atol=1e-16)


def test_step_time_reversible():
    rng = np.random.default_rng(7)
    N, L = 8, 1.0
    pm = O.OraclePM3D(N, L, N)
    q = (np.arange(N) + 0.5) * (L / N)
    x0 = (np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3) + 0.01 * rng.standard_normal((N ** 3, 3))) % L
    p0 = 1e-3 * rng.standard_normal((N ** 3, 3))
    pm.x, pm.p = x0.copy(), p0.copy()
    a, da = 1.0, 0.01
    a1, _ = pm.step(a, da)
    pm._acc = None
    a2, _ = pm.step(a1, -da)
    assert abs(a2 - a) < 1e-15
    d = (pm.x - x0 + L / 2) % L - L / 2
    assert np.abs(d).max() < 1e-12
    assert np.abs(pm.p - p0).max() < 1e-12


# ---------------------------------------------------------------- reference's own outputs
def _yoo_spline():
    from pbh_amd.ltb import profile_table
    return profile_table("yoo")["spline"]


def test_oracle_matches_reference_golden(golden):
    arrays, meta = golden
    sp = _yoo_spline()
    for c in meta["cases"]:
        nm = c["name"]
        pm = O.OraclePM3D(c["N"], c["L"], c["Np"], seed=c["seed"], smooth=c["smooth"]).make_ic(
            c["mu"], sp, c["e"], c["p"], c["tail"], (2.0, 10.0))
        np.testing.assert_array_equal(pm.x, arrays[nm + "/x0"])
        np.testing.assert_array_equal(pm.p, arrays[nm + "/p0"])
        np.testing.assert_array_equal(pm.r_ell_q, arrays[nm + "/r_ell_q"])
        assert pm.tail_rms_grid == c["tail_rms_grid"]
        acc, delta = pm.forces(1.0)
        np.testing.assert_array_equal(acc, arrays[nm + "/acc0"])
        np.testing.assert_array_equal(delta, arrays[nm + "/delta0"])
        a = 1.0
        for k, st in enumerate(c["steps"]):
            a, delta = pm.step(a, st["da"])
            assert a == st["a"]
            if f"{nm}/x{k + 1}" in arrays:
                np.testing.assert_array_equal(pm.x, arrays[f"{nm}/x{k + 1}"])
                np.testing.assert_array_equal(pm.p, arrays[f"{nm}/p{k + 1}"])
                np.testing.assert_array_equal(delta, arrays[f"{nm}/delta{k + 1}"])
        ana = O.analyse(pm, a)
        for key, val in c["analyse"].items():
            np.testing.assert_array_equal(np.asarray(ana[key]), np.asarray(val), err_msg=key)


def test_ltb_scalars_match_reference_golden(golden):
    from pbh_amd.ltb import LTBGaussian
    _, meta = golden
    for row in meta["ltb"]:
        g = LTBGaussian(row["mu"])
        assert g.r_m == row["r_m"] and g.t_H == row["t_H"]
        assert g.tC(1e-6) == pytest.approx(row["tC0"], rel=1e-14)
        assert g.tC(g.r_m) == pytest.approx(row["tC_rm"], rel=1e-14)
    sp = _yoo_spline()
    np.testing.assert_allclose(sp(np.asarray(meta["yoo_spline_r"])), meta["yoo_spline_v"], rtol=1e-15, atol=1e-17)


# ---------------------------------------------------- shell-crossing detector (build-defined)
def test_min_jacobian_plane_wave_closed_form():
    """1D Zel'dovich plane wave x = q + A sin(k q): det(dx/dq) = 1 + A k cos(k q) in the continuum;
    with forward differences over dq the minimum is 1 + A*(sin(k(q+dq)) - sin(kq))/dq, whose zero
    gives the crossing amplitude A_c = dq / max_q (sin(kq) - sin(k(q+dq)))."""
    Np, L = 16, 1.0
    dq = L / Np
    q1 = (np.arange(Np) + 0.5) * dq
    k = 2 * np.pi / L
    Q = np.stack(np.meshgrid(q1, q1, q1, indexing="ij"), -1).reshape(-1, 3)
    for A in (0.0, 0.05, 0.2):
        x = Q.copy()
        x[:, 0] = (Q[:, 0] + A * np.sin(k * Q[:, 0])) % L
        expect = 1 + A * np.min((np.sin(k * (q1 + dq)) - np.sin(k * q1)) / dq)
        assert O.min_jacobian(x, Np, L) == pytest.approx(expect, rel=1e-12, abs=1e-12)
    # linear interpolation of the crossing
    assert O.crossing_scale_factor(1.0, 0.5, 2.0, -0.5) == pytest.approx(1.5)
