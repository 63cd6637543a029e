"""Host-side logic of the drop-in facade (CPU): step scalars, IC generator, schedule."""
import numpy as np
import pytest

from pbh_amd import ic
from pbh_amd.pm3d import H_I, step_scalars, zeldovich_axes
from oracle import pm3d_oracle as O


def test_step_scalars_follow_reference_expressions():
    a, da, smooth, dx = 1.37, 0.0274, 2.0, 0.15
    sc = step_scalars(a, da, smooth, dx, True)
    H = H_I * a ** -1.5
    a_mid = a + 0.5 * da
    a_new = a + da
    assert sc.c1 == 0.5 * da and sc.den1 == a * H
    assert sc.den_d == a_mid ** 3 * (H_I * a_mid ** -1.5)
    assert sc.den2 == a_new * (H_I * a_new ** -1.5)
    assert sc.s_new == -1.5 * H_I ** 2 / a_new and sc.s_old == -1.5 * H_I ** 2 / a
    assert sc.sm2 == (smooth * dx) ** 2 and sc.need_forces == 1


def test_zeldovich_axes_matches_oracle():
    for e, p in [(0, 0), (0.2, 0.05), (0.3, -0.1)]:
        np.testing.assert_array_equal(zeldovich_axes(e, p), O.zeldovich_axes(e, p))
        assert np.prod(zeldovich_axes(e, p)) == pytest.approx(1.0, rel=1e-14)


def test_ic_zeldovich_divergence_recovers_delta():
    """delta = -div psi on the mesh (spectral), psi the Zel'dovich displacement per unit D."""
    spec = ic.PeakSpec(N=16, mu=0.1, sigma=(0.2, 0.18, 0.16))
    delta = ic.linear_delta(spec)
    x, p, q = ic.zeldovich(spec, delta)
    L, N = spec.L, spec.N
    disp = ((x - q - L / 2) + L / 2) % L - L / 2     # x = (q + disp + L/2) % L
    # Np == N: lattice sites are cell centres, so the CIC gather returns mesh values exactly
    psi = disp.reshape(N, N, N, 3)
    kf = 2 * np.pi * np.fft.fftfreq(N, d=L / N)
    div = sum(np.fft.ifftn(1j * np.meshgrid(kf, kf, kf, indexing="ij")[c] * np.fft.fftn(psi[..., c])).real
              for c in range(3))
    ref = delta - delta.mean()
    # the Nyquist planes of the spectral derivative are lost; compare the smooth field
    assert np.abs(-div - ref).max() < 2e-3 * spec.mu
    a = spec.a_i
    np.testing.assert_allclose(p, a ** 2 * ic.hubble(a) * disp, rtol=1e-12, atol=1e-18)


def test_ic_named_workloads_and_noise_rms():
    s = ic.named("peak256")
    assert s.N == 256 and s.noise_rms_frac == 0.01
    small = ic.PeakSpec(N=32, mu=0.1, sigma=(0.1, 0.1, 0.1), noise_rms_frac=0.01, seed=2)
    d0 = ic.linear_delta(ic.PeakSpec(N=32, mu=0.1, sigma=(0.1, 0.1, 0.1)))
    d1 = ic.linear_delta(small)
    assert np.std(d1 - d0) == pytest.approx(0.001, rel=1e-12)


def test_schedule_matches_driver_loop():
    sch = ic.schedule(1.0, 0.02, 1.5)
    a = 1.0
    for (a0, da) in sch:
        assert a0 == a
        assert da == min(0.02 * a, 1.5 - a)
        a = a + da
    assert a >= 1.5


def test_driver_flags_and_tag_match_reference_driver():
    """pbh_amd.pm3d_peak keeps the flags, defaults and output tag of pm3d_peak.py:8-13,21."""
    from pbh_amd.pm3d_peak import parser, run_tag
    a = parser().parse_args([])
    assert (a.mu, a.e, a.p, a.N, a.Np, a.L, a.tend, a.tail, a.k1, a.k2, a.profile, a.seed, a.smooth, a.da_frac,
            a.out) == (0.1, 0.0, 0.0, 128, 64, 8.0, 40.0, 0.0, 2.0, 10.0, "yoo", 0, 2.0, 0.02, "runs/pm3d")
    assert run_tag(a) == "mu0.1_e0_p0_N128_Np64_L8_sm2"
    b = parser().parse_args(["--tail", "0.05", "--seed", "3", "--mu", "0.25"])
    assert run_tag(b) == "mu0.25_e0_p0_N128_Np64_L8_sm2_tail0.05_k2-10_s3"


def test_run_record_has_reference_provenance_fields(tmp_path):
    """The run record carries every field of the reference's provenance record
    (provenance.py:48-76) plus the solver extras."""
    import json
    from pbh_amd.provenance import RunRecord
    rec = RunRecord(tmp_path, {"mu": 0.1, "N": 16}, 3, name="r_provenance.json")
    out = rec.finish(extra={"collapse": {"t_c": np.float64(1.5)}, "steps": np.int64(7)})
    on_disk = json.loads((tmp_path / "r_provenance.json").read_text())
    for key in ("git", "python", "numpy", "compiler", "compile_flags", "host", "cpu", "n_cpu", "mpi", "params",
                "seed", "start_utc", "end_utc", "wall_time_s", "peak_rss_MB"):
        assert key in on_disk, key
    assert on_disk["collapse"]["t_c"] == 1.5 and on_disk["steps"] == 7 and out["seed"] == 3
    assert "--offload-arch=gfx950" in on_disk["compile_flags"]
