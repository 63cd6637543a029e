"""Initial conditions of the workloads this library is measured on (host, FP64, once per run).

The reference builds its initial data from a curvature profile (pm3d.py:40-73, `make_ic`).  The
workloads here are specified directly in the density contrast, so this module generalises the same
Zel'dovich construction (pm3d.py:59-68) to an explicit linear density field at a start scale
factor a_i in EdS:

    delta_i(x) = mu * exp(-0.5 * sum_c (x_c / sigma_c)^2)   [+ Gaussian random field]
    psi = -grad lap^-1 delta_i  (psi_k = i k delta_k / k^2)          displacement at a_i
    x = (q + CIC(psi)(q) + L/2) % L,   p = a_i^2 H(a_i) psi,   H(a) = H_I a^-3/2

``mu`` is the linear density contrast at a_i (not extrapolated to a = 1).  Random fields come from
the SplitMix64 counter stream of `white_noise`, which the device generator reproduces cell by cell,
so `device_ic` builds the same initial data on the GPU.  The peak is centred on
the box-centre mesh vertex, as in pm3d.py:42,66, and particles sit on the lattice
q = (i + 1/2) L/Np - L/2 in C order (pm3d.py:64-65).  Oracle and device receive the same (x, p).

Named workloads (periodic box L = 1, Np = N):
  peak64    N=64,  mu=0.3, sigma=(0.12,0.11,0.10),             seed 0   (CPU reference case)
  peak128   N=128, mu=0.2, sigma=(0.10,0.09,0.08),             seed 1   (FP64 on one GPU)
  peak256   N=256, mu=0.1, e=0.2, p=0, axes 0.1 L via zeldovich_axes,
            + GRF P(k) ~ k^-2 for 0 < k <= k_Nyq/2, rms 0.01 mu, noise stream seed 2  (mixed)
  batch     one peak per seed (100..163): draws from default_rng(seed) in the order
            N ~ U{32, 64, 96, 128}, mu ~ U[0.05, 0.5], e ~ U[0, 0.3], p ~ U[-e, e];
            axes 0.1 L * zeldovich_axes(e, p)                                  (batched, mixed)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .pm3d import H_I, zeldovich_axes


@dataclass
class PeakSpec:
    N: int
    mu: float
    sigma: tuple                  # Gaussian widths along x, y, z in units of L
    L: float = 1.0
    Np: int = 0                   # 0: Np = N
    a_i: float = 0.01
    seed: int = 0
    noise_rms_frac: float = 0.0   # rms of the added random field as a fraction of mu
    noise_kmax_frac: float = 0.5  # field cut at this fraction of k_Nyquist
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n_side(self) -> int:
        return self.Np or self.N


def named(name: str) -> PeakSpec:
    if name == "peak64":
        return PeakSpec(N=64, mu=0.3, sigma=(0.12, 0.11, 0.10), seed=0, name=name)
    if name == "peak128":
        return PeakSpec(N=128, mu=0.2, sigma=(0.10, 0.09, 0.08), seed=1, name=name)
    if name == "peak256":
        ax = 0.1 * zeldovich_axes(0.2, 0.0)
        return PeakSpec(N=256, mu=0.1, sigma=tuple(ax), seed=2, noise_rms_frac=0.01, name=name,
                        meta={"e": 0.2, "p": 0.0})
    raise KeyError(name)


def hubble(a: float) -> float:
    return H_I * a ** -1.5


def _k_grids(N: int, L: float):
    dx = L / N
    kf = 2 * np.pi * np.fft.fftfreq(N, d=dx)
    kh = 2 * np.pi * np.fft.rfftfreq(N, d=dx)
    kx, ky, kz = np.meshgrid(kf, kf, kh, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    return kx, ky, kz, k2


_U64 = np.uint64


def splitmix64(z):
    """SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014) on uint64 arrays (wrapping arithmetic);
    the same stream as the device generator in csrc/ic.hpp."""
    z = np.asarray(z, dtype=np.uint64) + _U64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> _U64(30))) * _U64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> _U64(27))) * _U64(0x94D049BB133111EB)
    return z ^ (z >> _U64(31))


def white_noise(seed: int, shape) -> np.ndarray:
    """Standard normal field from the counter stream `seed`: Box-Muller of two 53-bit uniforms per
    cell, u1 in (0, 1], u2 in [0, 1).  Cell i of the flattened field is independent of the shape."""
    M = int(np.prod(shape))
    key = splitmix64(np.array([seed], dtype=np.uint64))[0]
    i = np.arange(M, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a = splitmix64(key + _U64(2) * i)
        b = splitmix64(key + _U64(2) * i + _U64(1))
    u1 = ((a >> _U64(11)) + _U64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> _U64(11)).astype(np.float64) * 2.0 ** -53
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)).reshape(shape)


def linear_delta(spec: PeakSpec) -> np.ndarray:
    """delta_i on the mesh (cell centres relative to the box-centre vertex)."""
    N, L = spec.N, spec.L
    dx = L / N
    xc = (np.arange(N) + 0.5) * dx - L / 2
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
    sx, sy, sz = (s * L for s in spec.sigma)
    delta = spec.mu * np.exp(-0.5 * ((X / sx) ** 2 + (Y / sy) ** 2 + (Z / sz) ** 2))
    if spec.noise_rms_frac > 0:
        white = np.fft.rfftn(white_noise(spec.seed, (N, N, N)))
        kx, ky, kz, k2 = _k_grids(N, L)
        kk = np.sqrt(k2)
        kk[0, 0, 0] = 0.0
        kcut = spec.noise_kmax_frac * np.pi * N / L
        amp = np.where((kk > 0) & (kk <= kcut), 1.0 / np.maximum(kk, 1e-30), 0.0)   # sqrt(P), P ~ k^-2
        field_ = np.fft.irfftn(white * amp, s=(N, N, N))
        field_ *= spec.noise_rms_frac * spec.mu / np.std(field_)
        delta = delta + field_
    return delta


def zeldovich(spec: PeakSpec, delta: np.ndarray | None = None):
    """(x, p, q) of the Zel'dovich growing mode at a_i for the linear field delta_i."""
    N, L = spec.N, spec.L
    Np = spec.n_side
    if delta is None:
        delta = linear_delta(spec)
    kx, ky, kz, k2 = _k_grids(N, L)
    dk = np.fft.rfftn(delta)
    dk[0, 0, 0] = 0.0
    psi = [np.fft.irfftn(1j * kc * dk / k2, s=(N, N, N)) for kc in (kx, ky, kz)]
    q1 = (np.arange(Np) + 0.5) * (L / Np) - L / 2
    Q = np.stack(np.meshgrid(q1, q1, q1, indexing="ij"), axis=-1).reshape(-1, 3)
    Qw = (Q + L / 2) % L
    disp = np.stack([_cic_gather(f, Qw, L / N, N) for f in psi], axis=-1)
    x = (Q + disp + L / 2) % L
    p = spec.a_i ** 2 * hubble(spec.a_i) * disp
    return x, p, Q


def _cic_gather(F, X, dx, N):
    """Host CIC gather used only to place the initial lattice (same stencil as pm3d.py:91-99)."""
    g = X / dx - 0.5
    i0 = np.floor(g).astype(np.int64)
    f = g - i0
    lo, hi = i0 % N, (i0 + 1) % N
    out = np.zeros(len(X))
    for cx in (0, 1):
        ix = hi[:, 0] if cx else lo[:, 0]
        wx = f[:, 0] if cx else 1 - f[:, 0]
        for cy in (0, 1):
            iy = hi[:, 1] if cy else lo[:, 1]
            wy = f[:, 1] if cy else 1 - f[:, 1]
            for cz in (0, 1):
                iz = hi[:, 2] if cz else lo[:, 2]
                wz = f[:, 2] if cz else 1 - f[:, 2]
                out += F[ix, iy, iz] * wx * wy * wz
    return out


def make(spec: PeakSpec):
    """Return dict(x, p, q, a_i, N, Np, L) for a workload."""
    x, p, q = zeldovich(spec)
    return {"x": x, "p": p, "q": q, "a_i": spec.a_i, "N": spec.N, "Np": spec.n_side, "L": spec.L}


def schedule(a0: float, da_frac: float, a_end: float):
    """Geometric step schedule of the driver loop (pm3d_peak.py:26-28): da = min(da_frac*a, a_end-a)."""
    a = a0
    out = []
    while a < a_end:
        da = min(da_frac * a, a_end - a)
        out.append((a, da))
        a = a + da
    return out


def batch_spec(seed: int) -> PeakSpec:
    """One member of the batched workload (draw order: N, mu, e, p from default_rng(seed))."""
    rng = np.random.default_rng(seed)
    N = int(rng.choice([32, 64, 96, 128]))
    mu = float(rng.uniform(0.05, 0.5))
    e = float(rng.uniform(0.0, 0.3))
    p = float(rng.uniform(-e, e))
    ax = 0.1 * zeldovich_axes(e, p)
    return PeakSpec(N=N, mu=mu, sigma=tuple(ax), seed=seed, name=f"batch{seed}", meta={"e": e, "p": p})


def validate(spec: PeakSpec) -> None:
    """Raise ValueError for parameters that cannot describe a peak: a non-finite amplitude, widths
    that are not finite and positive (the Zel'dovich axes are undefined when the smallest
    eigenvalue (1 - 3e + p)/3 is not positive), a non-finite random field, or a bad start."""
    if not np.isfinite(spec.mu):
        raise ValueError(f"{spec.name or 'peak'}: amplitude mu={spec.mu} is not finite")
    sig = np.asarray(spec.sigma, dtype=float)
    if sig.shape != (3,) or not np.all(np.isfinite(sig)) or not np.all(sig > 0):
        e, p = spec.meta.get("e"), spec.meta.get("p")
        raise ValueError(f"{spec.name or 'peak'}: widths {tuple(sig)} are not finite and positive"
                         + (f" (Zel'dovich axes undefined for e={e:g}, p={p:g})" if e is not None else ""))
    if not (np.isfinite(spec.noise_rms_frac) and spec.noise_rms_frac >= 0 and np.isfinite(spec.noise_kmax_frac)):
        raise ValueError(f"{spec.name or 'peak'}: random field parameters are not finite")
    if not (np.isfinite(spec.a_i) and spec.a_i > 0 and np.isfinite(spec.L) and spec.L > 0):
        raise ValueError(f"{spec.name or 'peak'}: a_i and L must be finite and positive")


def device_ic(pm, spec: PeakSpec, engine=None):
    """Build the workload's initial state on the GPU (pm3d_make_peak_ic) in `pm`.

    An FP64 PM3D builds it in place; a mixed PM3D receives it from an FP64 `engine` handle of the
    same N (created if not given; reuse one across requests to avoid re-allocation)."""
    import ctypes
    from . import _lib
    from .pm3d import PM3D
    validate(spec)
    if spec.n_side != spec.N:
        raise ValueError("device initial conditions need Np == N")
    ic = _lib.Pm3dPeakIC()
    ic.mu = spec.mu
    ic.sx, ic.sy, ic.sz = (s * spec.L for s in spec.sigma)
    ic.noise_rms = spec.noise_rms_frac * spec.mu if spec.noise_rms_frac > 0 else 0.0
    ic.noise_kcut = spec.noise_kmax_frac * np.pi * spec.N / spec.L
    ic.coef = spec.a_i ** 2 * hubble(spec.a_i)
    ic.seed = spec.seed
    if pm.precision == "fp64":
        _lib.check(pm._L.pm3d_make_peak_ic(pm._h, ctypes.byref(ic)), pm._h)
    else:
        eng = engine if engine is not None else PM3D(spec.N, spec.L, spec.N, precision="fp64")
        _lib.check(eng._L.pm3d_make_peak_ic(eng._h, ctypes.byref(ic)), eng._h)
        _lib.check(pm._L.pm3d_copy_state(pm._h, eng._h), pm._h)
    pm._acc_a = None
    return pm
