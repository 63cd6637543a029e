"""CPU FP64 oracle for the pm3d particle-mesh step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
this module, and only as the checker (or as the timed CPU baseline).  The product path in ``pbh_amd``
never imports it.

It restates, in plain numpy FP64 and in this file's own structure, the algorithm of
``pbhgr/pbhgr/pm3d.py`` (reference repository uniquehuman/pbh).  Each function cites the reference
lines it follows.  Where rounding order matters for parity it is kept as the reference evaluates it:

* CIC index   ``g = X/dx - 0.5; i0 = floor(g); f = g - i0``            pm3d.py:76-79
* deposit     8 corners in (x0|x1, y0|y1, z0|z1) order, weight (wx*wy)*wz, bincount per corner
                                                                           pm3d.py:81-89
* interp      out += ((F*wx)*wy)*wz over the same corners               pm3d.py:91-99
* forces      delta = rho/n_mean - 1; dk = rfftn(delta);
              phik = ((-1.5*H_I**2/a) * dk) / k2   [numpy: complex/real == z * (1/k2)]
              [* exp(-0.5*k2*(smooth*dx)**2)]; phik[0,0,0] = 0;
              g_c = irfftn(-1j*k_c*phik)                                   pm3d.py:101-114
* step        KDK in a with the cached opening acceleration            pm3d.py:117-129

Parity pinning (what checks this oracle, see tests/test_oracle_*.py):

* cic_index    closed-form table (cell centres, wrap at 0 and L, x == L)           -> pinned
* deposit      brute-force triple loop on N=4, mass conservation, lattice rho==1/8 -> pinned
* interp       constant / cell-centre / adjointness with deposit                    -> pinned
* forces       single Fourier mode closed form; golden vectors of the reference     -> pinned
* step         zero-force closed form, time reversibility; golden vectors           -> pinned
* make_ic      golden vectors written by importing the reference (tests/golden)     -> pinned
* analyse      golden vectors written by importing the reference (tests/golden)     -> pinned
* shell_cross  not in the reference; closed form on a 1D Zel'dovich plane wave      -> pinned
               (definition is this build's own, see ``min_jacobian``)
"""
from __future__ import annotations

import numpy as np

# pm3d.py:15-17 -- units G = c = 1, k = 1, a_i = 1
H_I = 5.0
T_I = 2.0 / (3.0 * H_I)
RHO_A3 = 3.0 * H_I ** 2 / (8.0 * np.pi)


def zeldovich_axes(e: float, p: float) -> np.ndarray:
    """Axis ratios with unit product from the Zel'dovich eigenvalues (pm3d.py:20-23)."""
    lam = np.array([(1 + 3 * e + p), (1 - 2 * p), (1 - 3 * e + p)]) / 3.0
    axes = lam ** -0.5
    return axes / np.prod(axes) ** (1.0 / 3.0)


def k_axes(N: int, dx: float):
    """1D wavenumbers with numpy's fftfreq/rfftfreq conventions (pm3d.py:31-32).

    fftfreq puts the Nyquist index N/2 at the NEGATIVE frequency; rfftfreq's last index is +N/2.
    """
    kfull = 2 * np.pi * np.fft.fftfreq(N, d=dx)
    khalf = 2 * np.pi * np.fft.rfftfreq(N, d=dx)
    return kfull, khalf


def cic_index(X1: np.ndarray, dx: float, N: int):
    """One axis of the cell-centred CIC stencil (pm3d.py:76-79).

    Returns (i0 mod N, (i0+1) mod N, f) with i0 = floor(X/dx - 0.5) as int64 and f = g - i0.
    """
    g = X1 / dx - 0.5
    base = np.floor(g).astype(np.int64)
    frac = g - base
    return base % N, (base + 1) % N, frac


def cic_key(X: np.ndarray, dx: float, N: int) -> np.ndarray:
    """Flat base-cell key (ix0*N + iy0)*N + iz0 of every particle (pm3d.py:88 flat index, corner 0)."""
    ix0 = cic_index(X[:, 0], dx, N)[0]
    iy0 = cic_index(X[:, 1], dx, N)[0]
    iz0 = cic_index(X[:, 2], dx, N)[0]
    return (ix0 * N + iy0) * N + iz0


def _corners(X: np.ndarray, dx: float, N: int):
    """Yield (ix, iy, iz, wx, wy, wz) for the 8 CIC corners in the reference's loop order (pm3d.py:85-87)."""
    sx = cic_index(X[:, 0], dx, N)
    sy = cic_index(X[:, 1], dx, N)
    sz = cic_index(X[:, 2], dx, N)
    pairs = []
    for s in (sx, sy, sz):
        lo, hi, f = s
        pairs.append(((lo, 1 - f), (hi, f)))
    for ix, wx in pairs[0]:
        for iy, wy in pairs[1]:
            for iz, wz in pairs[2]:
                yield ix, iy, iz, wx, wy, wz


def deposit(X: np.ndarray, dx: float, N: int) -> np.ndarray:
    """CIC particle counts per cell, C order (N,N,N) (pm3d.py:81-89)."""
    rho = np.zeros(N ** 3)
    for ix, iy, iz, wx, wy, wz in _corners(X, dx, N):
        rho += np.bincount((ix * N + iy) * N + iz, weights=wx * wy * wz, minlength=N ** 3)
    return rho.reshape(N, N, N)


def interp(F: np.ndarray, X: np.ndarray, dx: float, N: int) -> np.ndarray:
    """CIC gather of mesh F at positions X (pm3d.py:91-99)."""
    out = np.zeros(len(X))
    for ix, iy, iz, wx, wy, wz in _corners(X, dx, N):
        out += F[ix, iy, iz] * wx * wy * wz
    return out


def remainder(v, L):
    """np.remainder, restated for the record: fmod with the sign fixed to L's; can return L itself
    for tiny negative v (pm3d.py:67,124)."""
    return np.remainder(v, L)


class OraclePM3D:
    """FP64 restatement of ``PM3D`` (pm3d.py:26-129)."""

    def __init__(self, N: int, L: float, Np: int, seed: int = 0, smooth: float = 0.0):
        # pm3d.py:27-37
        self.N, self.L, self.Np, self.smooth = N, L, Np, smooth
        self.dx = L / N
        kf, kh = k_axes(N, self.dx)
        self.kx, self.ky, self.kz = np.meshgrid(kf, kf, kh, indexing="ij")
        self.k2 = self.kx ** 2 + self.ky ** 2 + self.kz ** 2
        self.k2[0, 0, 0] = 1.0
        self.rng = np.random.default_rng(seed)
        self.m_part = RHO_A3 * L ** 3 / Np ** 3
        self._acc = None
        self._acc_a = None

    # -- initial data (pm3d.py:40-73)
    def make_ic(self, mu, profile_spline, e=0.0, p=0.0, tail_rms=0.0, tail_k=(2.0, 10.0)):
        N, L = self.N, self.L
        xc = (np.arange(N) + 0.5) * self.dx - L / 2
        X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
        A = zeldovich_axes(e, p)
        r_ell = np.sqrt((X / A[0]) ** 2 + (Y / A[1]) ** 2 + (Z / A[2]) ** 2)
        zeta = mu * profile_spline(np.minimum(r_ell, profile_spline.x[-1]))
        zk = np.fft.rfftn(zeta)
        if tail_rms > 0:
            kmag = np.sqrt(self.k2)
            kmag[0, 0, 0] = 0.0
            k_lo, k_hi = tail_k
            inband = (kmag >= k_lo) & (kmag <= k_hi)
            power = np.where(inband, 2 * np.pi ** 2 * tail_rms ** 2
                             / (np.maximum(kmag, 1e-30) ** 3 * np.log(k_hi / k_lo)), 0.0)
            white = np.fft.rfftn(self.rng.standard_normal((N, N, N)))
            tail = white * np.sqrt(power * N ** 3 / L ** 3)
            self.tail_rms_grid = float(np.std(np.fft.irfftn(tail, s=(N, N, N))))
            zk = zk + tail
        else:
            self.tail_rms_grid = 0.0
        dk = 0.4 * self.k2 / H_I ** 2 * zk
        dk[0, 0, 0] = 0.0
        phik = -dk / self.k2
        psi = [np.fft.irfftn(1j * kc * phik, s=(N, N, N)) * -1.0 for kc in (self.kx, self.ky, self.kz)]
        q1 = (np.arange(self.Np) + 0.5) * (L / self.Np) - L / 2
        Q = np.stack(np.meshgrid(q1, q1, q1, indexing="ij"), axis=-1).reshape(-1, 3)
        disp = np.stack([self.interp(f, (Q + L / 2) % L) for f in psi], axis=-1)
        self.x = (Q + disp + L / 2) % L
        self.p = H_I * disp
        self.q = Q
        self.delta_lin = np.fft.irfftn(dk, s=(N, N, N))
        self.r_ell_q = np.sqrt(np.sum((Q / A) ** 2, axis=1))
        self.A = A
        return self

    # -- mesh operations (pm3d.py:76-99)
    def deposit(self, X):
        return deposit(X, self.dx, self.N)

    def interp(self, F, X):
        return interp(F, X, self.dx, self.N)

    def density_contrast(self, X):
        """delta = rho/n_mean - 1 (pm3d.py:103-104)."""
        n_mean = self.Np ** 3 / self.N ** 3
        return self.deposit(X) / n_mean - 1.0

    def potential_k(self, delta, a):
        """Phi_k on the half spectrum (pm3d.py:105-109)."""
        dk = np.fft.rfftn(delta)
        phik = -1.5 * H_I ** 2 / a * dk / self.k2
        if self.smooth > 0:
            phik = phik * np.exp(-0.5 * self.k2 * (self.smooth * self.dx) ** 2)
        phik[0, 0, 0] = 0.0
        return phik

    def gradient_meshes(self, phik):
        """g_c = -dPhi/dx_c on the mesh, spectral (pm3d.py:111-112)."""
        N = self.N
        return [np.fft.irfftn(-1j * kc * phik, s=(N, N, N)) for kc in (self.kx, self.ky, self.kz)]

    def forces(self, a):
        """(acc (n,3), delta) at the current positions (pm3d.py:101-114)."""
        delta = self.density_contrast(self.x)
        g = self.gradient_meshes(self.potential_k(delta, a))
        acc = np.stack([self.interp(gc, self.x) for gc in g], axis=-1)
        return acc, delta

    # -- KDK leapfrog in a (pm3d.py:117-129)
    def step(self, a, da):
        H = H_I * a ** -1.5
        if self._acc is None or self._acc_a != a:
            self._acc, _ = self.forces(a)
        self.p += 0.5 * da * self._acc / (a * H)
        a_mid = a + 0.5 * da
        H_mid = H_I * a_mid ** -1.5
        self.x = (self.x + da * self.p / (a_mid ** 3 * H_mid)) % self.L
        a_new = a + da
        H_new = H_I * a_new ** -1.5
        acc, delta = self.forces(a_new)
        self.p += 0.5 * da * acc / (a_new * H_new)
        self._acc, self._acc_a = acc, a_new
        return a_new, delta


def analyse(pm, a, fracs=(0.1, 0.2, 0.3), radii_cells=(1.0, 2.0, 4.0, 8.0)):
    """Lagrangian-set axes, hoop compaction and enclosed-mass compaction (pm3d.py:132-159)."""
    L = pm.L
    x = np.asarray(pm.x)
    out = {"a": a}
    rm = np.sqrt(6.0)
    order = np.argsort(pm.r_ell_q)
    core = np.zeros(len(x), bool)
    core[order[:27]] = True
    xs = x[core]
    rel = (xs - xs[0] + L / 2) % L - L / 2
    centre = (xs[0] + rel.mean(axis=0)) % L
    rel_all = (x - centre + L / 2) % L - L / 2
    for f in fracs:
        sel = pm.r_ell_q < f * rm
        if sel.sum() < 30:
            sel = np.zeros(len(x), bool)
            sel[order[:30]] = True
        r = rel_all[sel]
        r = r - r.mean(axis=0)
        inertia = r.T @ r / len(r)
        ev = np.sqrt(np.maximum(np.linalg.eigvalsh(inertia), 0.0)) * np.sqrt(5.0)
        M = pm.m_part * sel.sum()
        axes_phys = a * ev[::-1]
        R_hoop = np.sqrt(0.5 * (axes_phys[0] ** 2 + axes_phys[1] ** 2))
        out[f"axes_{f}"] = axes_phys.tolist()
        out[f"axes_com_{f}"] = ev[::-1].tolist()
        out[f"M_{f}"] = M
        out[f"hoop_{f}"] = 2 * M / max(R_hoop, 1e-30)
    r_com = np.sqrt(np.sum(rel_all ** 2, axis=1))
    for nc in radii_cells:
        R = a * nc * pm.dx
        M = pm.m_part * np.count_nonzero(r_com < nc * pm.dx)
        out[f"m_in_{nc:g}c"] = M
        out[f"C_{nc:g}c"] = 2 * M / R
    out["rho_max_over_mean"] = float(np.max(pm.deposit(x)) / (pm.Np ** 3 / pm.N ** 3))
    return out


# ------------------------------------------------------------------ shell crossing (build-defined)
def min_jacobian(x: np.ndarray, Np: int, L: float) -> float:
    """min over Lagrangian lattice cells of det(dx/dq), forward differences along the q-lattice.

    Not in the reference (pm3d has no shell-crossing detector).  x is (Np^3, 3) in Lagrangian C order.
    Differences are unwrapped to the minimum image, and scaled by the lattice spacing dq = L/Np, so
    an undisplaced lattice gives exactly 1 and the first shell crossing is where the minimum reaches 0.
    """
    X = x.reshape(Np, Np, Np, 3)
    dq = L / Np

    def fwd(axis):
        d = np.roll(X, -1, axis=axis) - X
        d = d - L * np.rint(d / L)
        return d / dq

    ex, ey, ez = fwd(0), fwd(1), fwd(2)
    det = (ex[..., 0] * (ey[..., 1] * ez[..., 2] - ey[..., 2] * ez[..., 1])
           - ex[..., 1] * (ey[..., 0] * ez[..., 2] - ey[..., 2] * ez[..., 0])
           + ex[..., 2] * (ey[..., 0] * ez[..., 1] - ey[..., 1] * ez[..., 0]))
    return float(det.min())


def crossing_scale_factor(a0, j0, a1, j1):
    """Linear interpolation in a of the zero of min det(dx/dq) between two steps (j0 > 0 >= j1)."""
    return a0 + (a1 - a0) * j0 / (j0 - j1)
