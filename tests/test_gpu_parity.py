"""GPU parity of the HIP path against the CPU FP64 oracle (oracle/pm3d_oracle.py) and the reference's
golden vectors.  Every call goes through the C-ABI of _pm3d_hip.so.

Tolerances (FP64 mode): cell keys and CIC fractions bit-exact; everything else
max|gpu - ref| / max|ref| <= 1e-10 per field (positions: minimum-image difference over L).
"""
import sys

import numpy as np
import pytest

from oracle import pm3d_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def pos_err(x, y, L):
    d = (np.asarray(x) - np.asarray(y) + L / 2) % L - L / 2
    return float(np.abs(d).max() / L)


@pytest.fixture(scope="module")
def PM3D():
    from pbh_amd.pm3d import PM3D as cls
    return cls


def lattice(N, L, jitter=0.0, seed=0):
    q = (np.arange(N) + 0.5) * (L / N)
    X = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    if jitter:
        X = (X + jitter * (L / N) * np.random.default_rng(seed).standard_normal(X.shape)) % L
    return X


# ------------------------------------------------------------------------------ CIC keys
@pytest.mark.parametrize("N,L", [(16, 1.0), (128, 1.0), (64, 8 * np.sqrt(6.0))])
def test_cic_keys_bit_exact(PM3D, N, L):
    say("cic keys", N, L)
    pm = PM3D(N, L, 8)
    dx = L / N
    rng = np.random.default_rng(N)
    X = rng.uniform(0, L, (200000, 3))
    edge = np.array([0.0, L, np.nextafter(L, 0), 0.5 * dx, np.nextafter(0.5 * dx, 0), np.nextafter(0.5 * dx, 1),
                     np.remainder(-1e-20, L), np.remainder(-0.0, L), 3.5 * dx, (N - 0.5) * dx])
    X[: len(edge), 0] = edge
    X[: len(edge), 1] = edge[::-1]
    X[: len(edge), 2] = np.roll(edge, 3)
    key, frac = pm.cic_keys(X)
    assert np.array_equal(key, O.cic_key(X, dx, N))
    ref = np.stack([O.cic_index(X[:, d], dx, N)[2] for d in range(3)], -1)
    assert np.array_equal(frac, ref)


# ------------------------------------------------------------------------------- deposit
@pytest.mark.parametrize("N,n,fb", [(8, 1, 0), (16, 5000, 0), (64, 300000, 0), (128, 2097152, 0),
                                    (16, 5000, 40), (128, 2097152, 40)])
def test_deposit_parity(PM3D, N, n, fb):
    """fb 0: 128-bit cells at 2^-52; fb 40: 64-bit cells at 2^-40 (about 1e-12 of a particle)."""
    say("deposit", N, n, fb)
    L = 1.0
    pm = PM3D(N, L, 8, frac_bits=fb)
    X = np.random.default_rng(n).uniform(0, L, (n, 3))
    rho = pm.deposit(X)
    ref = O.deposit(X, L / N, N)
    assert rel(rho, ref) <= TOL
    assert abs(rho.sum() - n) <= 1e-9 * n


def test_deposit_lattice_exact(PM3D):
    say("deposit lattice")
    N = 32
    pm = PM3D(N, 1.0, N)
    assert np.array_equal(pm.deposit(lattice(N, 1.0)), np.ones((N, N, N)))
    assert np.array_equal(pm.deposit(lattice(2 * N, 1.0)), 8 * np.ones((N, N, N)))
    assert np.array_equal(pm.deposit(np.zeros((0, 3))), np.zeros((N, N, N)))


# -------------------------------------------------------------------------------- interp
@pytest.mark.parametrize("N", [8, 32, 128])
def test_interp_parity(PM3D, N):
    say("interp", N)
    L = 2.0
    pm = PM3D(N, L, 8)
    rng = np.random.default_rng(N + 1)
    F = rng.standard_normal((N, N, N))
    X = rng.uniform(0, L, (100000, 3))
    assert rel(pm.interp(F, X), O.interp(F, X, L / N, N)) <= 1e-14


# ----------------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [8, 16, 32, 64, 128, 256])
def test_rfftn_parity(PM3D, N):
    say("rfftn", N)
    pm = PM3D(N, 1.0, 8)
    a = np.random.default_rng(N).standard_normal((N, N, N))
    assert rel(pm.rfftn(a), np.fft.rfftn(a)) <= 1e-13


@pytest.mark.parametrize("N", [8, 16, 64, 128])
def test_irfftn_parity_non_hermitian(PM3D, N):
    """numpy's c2r drops the imaginary parts of the kz=0 and kz=N/2 bins after the x,y inverse
    transforms; random (non-Hermitian) input exercises exactly that convention."""
    say("irfftn", N)
    pm = PM3D(N, 1.0, 8)
    rng = np.random.default_rng(N + 7)
    c = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    assert rel(pm.irfftn(c), np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))) <= 1e-13


@pytest.mark.parametrize("N,smooth", [(16, 0.0), (64, 2.0), (128, 0.0)])
def test_poisson_gradient_parity(PM3D, N, smooth):
    say("poisson", N, smooth)
    L, a = 3.0, 1.3
    pm = PM3D(N, L, 8, smooth=smooth)
    ref = O.OraclePM3D(N, L, 8, smooth=smooth)
    delta = np.random.default_rng(N).standard_normal((N, N, N))
    g = pm.poisson_gradient(delta, a)
    gr = ref.gradient_meshes(ref.potential_k(delta, a))
    for c in range(3):
        assert rel(g[..., c], gr[c]) <= 1e-12, c


# ------------------------------------------------------------- reference golden vectors
def _spline():
    from pbh_amd.ltb import profile_table
    return profile_table("yoo")["spline"]


def test_make_ic_matches_golden(PM3D, golden):
    arrays, meta = golden
    for c in meta["cases"]:
        say("make_ic", c["name"])
        pm = PM3D(c["N"], c["L"], c["Np"], seed=c["seed"], smooth=c["smooth"]).make_ic(
            c["mu"], _spline(), c["e"], c["p"], c["tail"], (2.0, 10.0))
        assert pos_err(pm.x, arrays[c["name"] + "/x0"], c["L"]) <= TOL
        assert rel(pm.p, arrays[c["name"] + "/p0"]) <= TOL
        assert rel(pm.delta_lin, arrays[c["name"] + "/delta_lin"]) <= TOL
        assert pm.tail_rms_grid == pytest.approx(c["tail_rms_grid"], rel=1e-12)


def test_forces_and_steps_match_golden(PM3D, golden):
    arrays, meta = golden
    for c in meta["cases"]:
        nm = c["name"]
        say("golden steps", nm)
        pm = PM3D(c["N"], c["L"], c["Np"], seed=c["seed"], smooth=c["smooth"])
        pm.set_state(arrays[nm + "/x0"], arrays[nm + "/p0"])
        acc, delta = pm.forces(1.0)
        assert rel(acc, arrays[nm + "/acc0"]) <= TOL
        assert rel(delta, arrays[nm + "/delta0"]) <= TOL
        a = 1.0
        for k, st in enumerate(c["steps"]):
            a, dm = pm.step(a, st["da"])
            assert a == st["a"]
            if f"{nm}/x{k + 1}" in arrays:
                assert pos_err(pm.x, arrays[f"{nm}/x{k + 1}"], c["L"]) <= TOL
                assert rel(pm.p, arrays[f"{nm}/p{k + 1}"]) <= TOL
                ref_delta = arrays[f"{nm}/delta{k + 1}"]
                assert rel(np.asarray(dm), ref_delta) <= TOL
                assert dm.max() == pytest.approx(ref_delta.max(), rel=1e-10, abs=1e-10)


def test_analyse_matches_golden(PM3D, golden):
    from pbh_amd.pm3d import analyse
    arrays, meta = golden
    for c in meta["cases"]:
        nm = c["name"]
        K = c["nsteps"]
        say("analyse", nm)
        pm = PM3D(c["N"], c["L"], c["Np"], seed=c["seed"], smooth=c["smooth"]).make_ic(
            c["mu"], _spline(), c["e"], c["p"], c["tail"], (2.0, 10.0))
        pm.set_state(arrays[f"{nm}/x{K}"], arrays[f"{nm}/p{K}"])
        out = analyse(pm, c["steps"][-1]["a"])
        for key, val in c["analyse"].items():
            np.testing.assert_allclose(np.asarray(out[key]), np.asarray(val), rtol=1e-9, atol=1e-12, err_msg=key)


# ------------------------------------------------------------ trajectories on the workloads
def _trajectory_case(PM3D, spec, nsteps, da_frac, frac_bits=0):
    from pbh_amd import ic
    st = ic.make(spec)
    L = spec.L
    pm = PM3D(spec.N, L, spec.n_side, frac_bits=frac_bits)
    pm.set_state(st["x"], st["p"])
    ref = O.OraclePM3D(spec.N, L, spec.n_side)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    a = st["a_i"]
    # keys bit-exact on identical positions
    key, _ = pm.cic_keys(st["x"])
    assert np.array_equal(key, O.cic_key(st["x"], L / spec.N, spec.N))
    for k in range(nsteps):
        da = da_frac * a
        a_g, dm = pm.step(a, da)
        a_r, dr = ref.step(a, da)
        assert a_g == a_r
        a = a_r
    assert rel(np.asarray(dm), dr) <= TOL
    assert pos_err(pm.x, ref.x, L) <= TOL
    assert rel(pm.p, ref.p) <= TOL
    return pm, ref


@pytest.mark.parametrize("fb", [0, 40])
def test_trajectory_peak32_parity(PM3D, fb):
    from pbh_amd import ic
    say("trajectory peak32", fb)
    spec = ic.PeakSpec(N=32, mu=0.3, sigma=(0.12, 0.11, 0.10), seed=0)
    pm, ref = _trajectory_case(PM3D, spec, 40, 0.02, fb)
    assert pm.min_jacobian() == pytest.approx(O.min_jacobian(ref.x, 32, 1.0), rel=1e-9)


@pytest.mark.parametrize("fb", [0, 40])
def test_trajectory_peak128_parity(PM3D, fb):
    """128^3 mesh, 2.1M particles, FP64 on the GPU vs the CPU oracle (3 steps from the IC)."""
    from pbh_amd import ic
    say("trajectory peak128", fb)
    _trajectory_case(PM3D, ic.named("peak128"), 3, 0.0125, fb)


def test_step_is_deterministic(PM3D):
    from pbh_amd import ic
    say("determinism")
    st = ic.make(ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10)))
    outs = []
    for _ in range(2):
        pm = PM3D(64, 1.0, 64)
        pm.set_state(st["x"], st["p"])
        a = 0.01
        for k in range(5):
            a, dm = pm.step(a, 0.02 * a)
        outs.append((pm.x, pm.p, np.asarray(dm)))
    for u, v in zip(*outs):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("fb", [0, 40])
def test_tiled_deposit_equals_atomic(PM3D, fb):
    """The step's owner-computes deposit and the reference-shaped atomic deposit give the same mesh."""
    from pbh_amd import ic
    say("tiled vs atomic", fb)
    spec = ic.PeakSpec(N=64, mu=0.5, sigma=(0.08, 0.07, 0.06))
    st = ic.make(spec)
    pm = PM3D(64, 1.0, 64, frac_bits=fb)
    pm.set_state(st["x"], st["p"])
    a = 0.01
    for k in range(60):          # run well into the non-linear regime so that 'far' particles exist
        a, dm = pm.step(a, 0.03 * a)
    d_step = np.asarray(dm)
    rho = pm.deposit(pm.x)
    assert np.array_equal(d_step, rho / 1.0 - 1.0)


@pytest.mark.parametrize("N,squeeze,shape", [(32, 0.5, "0"), (32, 0.5, "2"), (256, 0.4, "auto")])
def test_crowded_columns_keep_the_128bit_sums(PM3D, N, squeeze, shape, monkeypatch):
    """FP64 column deposit: a cell takes near particles from at most 64 sites and the entries of the
    four row lists at and below it (at most 6N each).  Meshes with 64 + 24N < 4096 sum each 128-bit
    cell in one 64-bit accumulator, larger ones as low and high 32-bit halves.  A lattice squeezed
    around the box centre puts every particle in the central rows: at 32^3 (one accumulator) the
    rows are full, at 256^3 (halves) they overflow into the global pass.  The step's density equals
    the reference-shaped atomic deposit of the same positions bit for bit."""
    monkeypatch.setenv("PM3D_DEPOSIT_SHAPE", shape)
    say("crowded columns", N, squeeze, shape)
    L = 1.0
    q = (np.arange(N) + 0.5) * L / N
    X = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    X = 0.5 * L + squeeze * (X - 0.5 * L)
    pm = PM3D(N, L, N)
    pm.set_state(X, np.zeros_like(X))
    del X
    a, dm = pm.step(0.01, 1e-6)  # its second deposit takes the far particles from the row lists
    a, dm = pm.step(a, 1e-6)
    assert np.array_equal(np.asarray(dm), pm.deposit(pm.x) / 1.0 - 1.0)
    pm.close()


def test_driver_rows_match_oracle_loop(tmp_path):
    """The drop-in driver (16^3 mesh, 8^3 particles, to t = 3 t_H) against the same loop on the oracle."""
    import json
    from pbh_amd.ltb import LTBGaussian, profile_table
    from pbh_amd.pm3d_peak import main
    say("driver")
    out = main(["--N", "16", "--Np", "8", "--tend", "3", "--mu", "0.3", "--e", "0.1", "--out", str(tmp_path)])
    ref_ltb = LTBGaussian(0.3)
    rm, t_H = ref_ltb.r_m, ref_ltb.t_H
    pm = O.OraclePM3D(16, 8.0 * rm, 8, seed=0, smooth=2.0).make_ic(0.3, profile_table("yoo")["spline"], 0.1, 0.0, 0.0, (2.0, 10.0))
    a, step, next_diag, rows = 1.0, 0, 0.0, []
    a_end = (3.0 * t_H / O.T_I) ** (2.0 / 3.0)
    while a < a_end:
        a, delta = pm.step(a, min(0.02 * a, a_end - a))
        step += 1
        t = O.T_I * a ** 1.5
        if t / t_H >= next_diag or a >= a_end:
            d = O.analyse(pm, a)
            d.update(step=step, delta_max=float(delta.max()))
            rows.append(d)
            next_diag = t / t_H + max(0.5, 0.02 * ref_ltb.tC(1e-6) / t_H)
    assert len(rows) == len(out["rows"])
    for r_ref, r_gpu in zip(rows, out["rows"]):
        assert r_gpu["step"] == r_ref["step"]
        for key in ("axes_com_0.2", "hoop_0.1", "m_in_2c", "rho_max_over_mean", "delta_max"):
            np.testing.assert_allclose(np.asarray(r_gpu[key]), np.asarray(r_ref[key]), rtol=1e-8, err_msg=key)
    assert (tmp_path / "mu0.3_e0.1_p0_N16_Np8_L8_sm2.json").exists()
    rec = json.loads((tmp_path / "mu0.3_e0.1_p0_N16_Np8_L8_sm2_provenance.json").read_text())
    assert rec["gpu"]["arch"].startswith("gfx950") and rec["steps"] == rows[-1]["step"]


@pytest.mark.parametrize("N,prec", [(128, "fp64"), (128, "mixed"), (96, "fp64"), (256, "mixed")])
def test_fft_tile_variants(PM3D, N, prec, monkeypatch):
    """Every FFT tile configuration (pencils per workgroup x one or two LDS buffers) transforms
    each pencil with the same operations, so a step's positions, momenta and density contrast
    agree bit for bit across all four."""
    from pbh_amd import ic
    say("fft tiles", N, prec)
    spec = ic.PeakSpec(N=N, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    outs = {}
    for tiles in ("0", "1", "2", "3"):
        monkeypatch.setenv("PM3D_FFT_TILES", tiles)
        pm = PM3D(N, 1.0, N, precision=prec)
        assert set(pm.fft_tiles().values()) == {int(tiles)}
        pm.set_state(st["x"], st["p"])
        a = st["a_i"]
        for _ in range(3):
            a, dm = pm.step(a, 0.02 * a)
        outs[tiles] = (pm.x, pm.p, np.asarray(dm))
        del pm
    for t in ("1", "2", "3"):
        for u, v in zip(outs["0"], outs[t]):
            assert np.array_equal(u, v), t


@pytest.mark.parametrize("N,prec", [(128, "fp64"), (256, "mixed"), (96, "fp64"), (96, "mixed"), (64, "mixed"),
                                    (16, "fp64")])
def test_deposit_shape_variants(PM3D, N, prec, monkeypatch):
    """Every deposit column shape sums the same fixed-point corner weights, so a step's positions,
    momenta and density contrast agree bit for bit across the shapes that fit, and the automatic
    choice is one of them."""
    from pbh_amd import ic
    say("deposit shapes", N, prec)
    spec = ic.PeakSpec(N=N, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    outs, shapes = {}, {}
    for v in ("0", "1", "2", "3", "auto"):
        monkeypatch.setenv("PM3D_DEPOSIT_SHAPE", v)
        pm = PM3D(N, 1.0, N, precision=prec)
        shapes[v] = pm.deposit_tiles()
        pm.set_state(st["x"], st["p"])
        a = st["a_i"]
        for _ in range(3):
            a, dm = pm.step(a, 0.02 * a)
        outs[v] = (pm.x, pm.p, np.asarray(dm))
        del pm
    cands = [shapes[v] for v in ("0", "1", "2", "3")]
    assert cands[0][:2] in ((8, 8), (4, 8), (4, 4), (2, 4)) and all(s[2] == N for s in cands)
    assert shapes["auto"] in cands
    assert len(set(cands)) >= 2
    for v in ("1", "2", "3", "auto"):
        for u, w in zip(outs["0"], outs[v]):
            assert np.array_equal(u, w), (v, shapes[v])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_live_shape_trials_keep_bits(PM3D, prec, monkeypatch):
    """The step deposits try every candidate column shape every 64 steps (PM3D_DEPOSIT_SHAPE=all
    makes every fitting shape a candidate) and may move to another; the trajectory and every
    step's density maximum agree bit for bit with a run held to one shape."""
    from pbh_amd import ic
    say("live shape trials", prec)
    spec = ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    outs, trials = {}, {}
    for v in ("0", "all"):
        monkeypatch.setenv("PM3D_DEPOSIT_SHAPE", v)
        pm = PM3D(64, 1.0, 64, precision=prec)
        pm.set_state(st["x"], st["p"])
        a = st["a_i"]
        maxes = []
        for _ in range(150):
            a, dm = pm.step(a, 0.004 * a)
            maxes.append(dm.max())
        outs[v] = (pm.x, pm.p, np.asarray(dm), np.array(maxes))
        trials[v] = pm.deposit_trials()
        pm.close()
    assert trials["0"]["trials"] == 0
    assert trials["all"]["trials"] >= 2 and bin(trials["all"]["candidates"]).count("1") >= 2, trials
    for u, w in zip(outs["0"], outs["all"]):
        assert np.array_equal(u, w)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_deferred_gather_is_invisible(PM3D, prec):
    """A step leaves its forces as gradient meshes and the next step gathers them in its first
    kernel.  Reading p after every step (gather + kick applied on their own), calling forces()
    between steps (it replaces the gradient meshes, so the next step kicks with the stored
    acceleration), assigning x (FP64: an exact round trip; the acceleration of the old positions is
    kept, as the reference keeps _acc), or none of these, must give the same trajectory bit for bit."""
    from pbh_amd import ic
    say("deferred gather", prec)
    spec = ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    outs = []
    for mode in ("plain", "read_p", "forces") + (("set_x",) if prec == "fp64" else ()):
        pm = PM3D(64, 1.0, 64, precision=prec)
        pm.set_state(st["x"], st["p"])
        a = st["a_i"]
        for _ in range(4):
            a, _dm = pm.step(a, 0.03 * a)
            if mode == "read_p":
                _ = pm.p
            elif mode == "forces":
                pm.forces(a)
            elif mode == "set_x":
                pm.x = pm.x
        outs.append((pm.x, pm.p))
        del pm
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][1], o[1])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_kernel_timing_is_invisible_and_consistent(PM3D, prec):
    """Per-kernel timing (each step kernel's dispatch stamping its own start/stop events) leaves the
    trajectory bit for bit unchanged, times every step kernel once per step, and its durations are
    positive and add up to no more than the wall time of the steps they were taken in."""
    import time
    from pbh_amd import ic
    say("kernel timing", prec)
    spec = ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    outs = []
    for timed in (False, True):
        pm = PM3D(64, 1.0, 64, precision=prec)
        pm.set_state(st["x"], st["p"])
        a = st["a_i"]
        a, _ = pm.step(a, 0.03 * a)  # first step: forces of the initial state, then the step
        pm.synchronize()
        pm.set_timing(timed)
        pm.timings()  # reset the accumulators
        steps = 5
        t0 = time.perf_counter()
        for _ in range(steps):
            a, _ = pm.step(a, 0.03 * a)
        pm.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3
        tm = pm.timings()
        pm.set_timing(False)
        outs.append((pm.x, pm.p))
        if timed:
            step_kernels = ("gather_kick", "deposit", "deposit_far", "fwd_zy", "x_poisson", "inv_yz")
            for k in step_kernels:
                cnt, ms = tm[k]
                assert cnt == steps, (k, cnt)
                assert np.isfinite(ms) and 0.0 < ms < wall_ms, (k, ms, wall_ms)
            assert sum(tm[k][1] for k in step_kernels) <= wall_ms
        else:
            assert all(c == 0 for c, _ in tm.values())
        del pm
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("prec,fb", [("fp64", 0), ("fp64", 40), ("mixed", 0)])
def test_global_sums(PM3D, prec, fb):
    """pm3d_totals: the deposit's mass is n plus the exact deviation of the fixed-point total
    (FP64 mode: recomputed on the host as the integer sum over particles of sum_c rint(w_c 2^fb) -
    2^fb from the oracle's corner weights, which equal the kernel's on identical positions, fb the
    handle's fraction bits (40 on 64-bit cells, 52 on 128-bit); the weights of a particle need not sum
    to exactly 1 in FP64), and the total momentum matches the oracle's sum of p; the same state gives
    the same bits twice."""
    from pbh_amd import ic
    say("global sums", prec, fb)
    spec = ic.PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10))
    st = ic.make(spec)
    pm = PM3D(64, 1.0, 64, precision=prec, frac_bits=fb)
    pm.set_state(st["x"], st["p"])
    ref = O.OraclePM3D(64, 1.0, 64)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    a = st["a_i"]
    for _ in range(5):
        pm.step(a, 0.03 * a)
        a, _ = ref.step(a, 0.03 * a)
    t1, t2 = pm.totals(), pm.totals()
    assert t1["mass"] == t2["mass"] and t1["mass_dev"] == t2["mass_dev"]
    assert np.array_equal(t1["momentum"], t2["momentum"])
    n = 64 ** 3
    assert t1["mass"] == pytest.approx(n + t1["mass_dev"], rel=1e-15)
    scale = np.abs(ref.p).sum(axis=0)
    dmom = np.abs(t1["momentum"] - ref.p.sum(axis=0)) / scale
    say(f"{prec}: mass - n = {t1['mass_dev']:.3e}, momentum {t1['momentum']}, rel diff {dmom.max():.2e}")
    if prec == "fp64":
        x, bits = pm.x, pm.frac_bits
        assert bits == (fb or 52)
        units = np.zeros(len(x), np.int64)  # per particle sum_c q_c <= 8 * 2^bits: exact in int64
        for ix, iy, iz, wx, wy, wz in O._corners(x, 1.0 / 64, 64):
            units += np.rint((wx * wy * wz) * 2.0 ** bits).astype(np.int64)
        dev_units = int((units - (1 << bits)).sum())
        assert t1["mass_dev"] == dev_units * 2.0 ** -bits
        assert dmom.max() <= 1e-12
    else:
        assert abs(t1["mass_dev"]) <= 8 * n * 2.0 ** -32
        assert dmom.max() <= 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_far_row_lists_are_exact(PM3D, prec, monkeypatch):
    """Far particles (base cell more than one cell from the lattice site) reach the column deposit
    through per-row lists built by the gather-kick-drift kernel, with an overflow list past each
    row's capacity.  A peak run well past the point where a third of the particles are far gives
    the same positions, momenta and density bit for bit with the default capacity, with capacity 1
    (nearly every far particle through the overflow pass) and with no lists (all through the
    global-atomic pass): every particle is deposited exactly once whichever path takes it."""
    from pbh_amd import ic
    say("far row lists", prec)
    spec = ic.PeakSpec(N=64, mu=0.5, sigma=(0.08, 0.07, 0.06))
    st = ic.make(spec)
    outs = {}
    for cap in ("", "1", "0"):
        monkeypatch.setenv("PM3D_ROW_CAP", cap) if cap else monkeypatch.delenv("PM3D_ROW_CAP", raising=False)
        pm = PM3D(64, 1.0, 64, precision=prec)
        pm.set_state(st["x"], st["p"])
        a = st["a_i"]
        for _ in range(60):
            a, dm = pm.step(a, 0.03 * a)
        outs[cap] = (pm.x, pm.p, np.asarray(dm))
        del pm
    x = outs[""][0]
    i0 = np.floor(x / (1.0 / 64) - 0.5).astype(np.int64) % 64
    q = np.arange(64)
    site = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    o = (i0 - site) % 64
    far = ~np.all((o == 0) | (o == 63), axis=1)
    say(f"{prec}: {far.mean():.1%} of the particles far after 60 steps")
    assert far.mean() > 0.05
    for cap in ("1", "0"):
        for u, v in zip(outs[""], outs[cap]):
            assert np.array_equal(u, v), cap
