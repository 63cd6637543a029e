"""Drop-in `pm3d` for MI355X: the call surface of the reference module pbhgr/pbhgr/pm3d.py
(uniquehuman/pbh) with the hot path on hand-written gfx950 HIP kernels.

    from pbh_amd.pm3d import PM3D, analyse, H_I, T_I      # was: from pbhgr.pm3d import ...

Same names, arguments and results as the reference (pm3d.py:15-159).  Differences a caller sees:

* particle state stays in HBM; ``pm.x`` / ``pm.p`` return host copies (assign to them to upload,
  e.g. ``pm.p = pm.p + dp``; an in-place ``pm.p += dp`` on the returned copy is not sent back);
* ``step`` returns a :class:`DeltaMesh`: ``delta.max()`` is read from the device, ``np.asarray(delta)``
  copies the mesh, valid until the next call that deposits;
* ``forces`` returns host arrays, as the reference does.

Every compute entry point goes through `_pm3d_hip.so`; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np

from . import _lib
from ._lib import as_f64, check, dptr, iptr

# pm3d.py:15-17 -- G = c = 1, k = 1, a_i = 1
H_I = 5.0
T_I = 2.0 / (3.0 * H_I)
RHO_A3 = 3.0 * H_I ** 2 / (8.0 * np.pi)


def zeldovich_axes(e: float, p: float) -> np.ndarray:
    """Axis ratios (product 1) from the Zel'dovich eigenvalues (1+3e+p, 1-2p, 1-3e+p)/3 (pm3d.py:20-23)."""
    lam = np.array([1 + 3 * e + p, 1 - 2 * p, 1 - 3 * e + p]) / 3.0
    ax = lam ** -0.5
    return ax / np.prod(ax) ** (1.0 / 3.0)


def step_scalars(a: float, da: float, smooth: float, dx: float, need_forces: bool) -> _lib.Pm3dStepScalars:
    """Host scalars of one KDK step, evaluated with Python floats in the order of pm3d.py:118-127
    (`**` is libm pow; the device only multiplies, divides and adds them)."""
    H = H_I * a ** -1.5
    a_mid = a + 0.5 * da
    H_mid = H_I * a_mid ** -1.5
    a_new = a + da
    H_new = H_I * a_new ** -1.5
    sc = _lib.Pm3dStepScalars()
    sc.a = a
    sc.da = da
    sc.c1 = 0.5 * da
    sc.den1 = a * H
    sc.den_d = a_mid ** 3 * H_mid
    sc.den2 = a_new * H_new
    sc.s_old = -1.5 * H_I ** 2 / a
    sc.s_new = -1.5 * H_I ** 2 / a_new
    sc.sm2 = (smooth * dx) ** 2 if smooth > 0 else 0.0
    sc.need_forces = 1 if need_forces else 0
    return sc


@functools.lru_cache(maxsize=256)
def _schedule_table(a_start, da_frac, a_end, max_steps, c1p, denp, smooth, dx):
    """The schedule table of PM3D.schedule_entries: host scalars of every step (step_scalars, in
    the order of pm3d.py:118-127) until a_end or max_steps."""
    rows, avals = [], []
    a = a_start
    while len(avals) < max_steps and a < a_end:
        da = min(da_frac * a, a_end - a)
        sc = step_scalars(a, da, smooth, dx, False)
        rows.append((c1p, denp, sc.c1, sc.den1, sc.da, sc.den_d, sc.den2, sc.s_new, sc.sm2))
        c1p, denp = sc.c1, sc.den2
        a = a + da
        avals.append(a)
    ents = (_lib.Pm3dSchedEntry * max(len(rows), 1))(*rows)
    return ents, tuple(avals)


class DeltaMesh:
    """The density contrast returned by :meth:`PM3D.step` (pm3d.py:104,129), kept on the device.

    ``max()`` reads the deposit's maximum from the device (valid for the last 4096 deposits);
    ``np.asarray(delta)`` copies the (N,N,N) mesh to the host, and is only possible until the next
    deposit of the same PM3D.
    """

    def __init__(self, pm: "PM3D", serial: int):
        self._pm = pm
        self._serial = serial
        self._max = None
        self._host = None
        self.shape = (pm.N, pm.N, pm.N)

    def max(self) -> float:
        if self._max is None:
            v = ctypes.c_double()
            check(self._pm._L.pm3d_delta_max_at(self._pm._h, self._serial, ctypes.byref(v)), self._pm._h)
            self._max = float(v.value)
        return self._max

    def _fetch(self) -> np.ndarray:
        if self._host is None:
            if _lib.load().pm3d_deposit_serial(self._pm._h) != self._serial:
                raise RuntimeError("delta mesh was overwritten by a later deposit; materialise it right after step()")
            out = np.empty(self.shape)
            check(self._pm._L.pm3d_get_delta(self._pm._h, dptr(out)), self._pm._h)
            self._host = out
        return self._host

    def __array__(self, dtype=None, copy=None):
        a = self._fetch()
        return a.astype(dtype) if dtype is not None else a

    def __getitem__(self, idx):
        return self._fetch()[idx]

    def __float__(self):
        return self.max()


class PM3D:
    """Newtonian PM for one periodic box in EdS on one MI355X (reference class pm3d.py:26-129).

    ``precision``: "fp64" (default; parity mode) or "mixed" (int32 cell + FP32 offset, FP32
    momenta/FFT/forces, int64 fixed-point deposit, FP64 scalars).
    """

    def __init__(self, N: int, L: float, Np: int, seed: int = 0, smooth: float = 0.0,
                 precision: str = "fp64", frac_bits: int = 0):
        self.N, self.L, self.Np = N, L, Np
        self.smooth = smooth
        self.dx = L / N
        self.rng = np.random.default_rng(seed)
        self.m_part = RHO_A3 * L ** 3 / Np ** 3
        self.precision = precision
        self._L = _lib.load()
        cfg = _lib.Pm3dCfg(N=N, Np=Np, L=L, smooth=smooth,
                           precision=_lib.PM3D_FP64 if precision == "fp64" else _lib.PM3D_MIXED,
                           frac_bits=frac_bits)
        h = ctypes.c_void_p()
        check(self._L.pm3d_create(ctypes.byref(cfg), ctypes.byref(h)), None)
        self._h = h
        self._acc_a = None
        self._kgrid = None
        self.n = Np ** 3
        self.frac_bits = int(self._L.pm3d_frac_bits(h))  # the deposit's grid is 2^-frac_bits of a particle

    def close(self):
        """Release the handle's device memory now (otherwise when the object is collected)."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.pm3d_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state (pm3d.py:67-68,124)
    @property
    def x(self) -> np.ndarray:
        out = np.empty((self.n, 3))
        check(self._L.pm3d_get_state(self._h, dptr(out), None), self._h)
        return out

    @x.setter
    def x(self, v):
        v = as_f64(v).reshape(self.n, 3)
        check(self._L.pm3d_set_state(self._h, dptr(v), None, self.n), self._h)

    @property
    def p(self) -> np.ndarray:
        out = np.empty((self.n, 3))
        check(self._L.pm3d_get_state(self._h, None, dptr(out)), self._h)
        return out

    @p.setter
    def p(self, v):
        v = as_f64(v).reshape(self.n, 3)
        check(self._L.pm3d_set_state(self._h, None, dptr(v), self.n), self._h)

    def set_state(self, x, p):
        x = as_f64(x).reshape(self.n, 3)
        p = as_f64(p).reshape(self.n, 3)
        check(self._L.pm3d_set_state(self._h, dptr(x), dptr(p), self.n), self._h)

    # -- k-grids of the reference (pm3d.py:31-35), built on the host only for make_ic
    def _kgrids(self):
        if self._kgrid is None:
            k1 = 2 * np.pi * np.fft.fftfreq(self.N, d=self.dx)
            kz = 2 * np.pi * np.fft.rfftfreq(self.N, d=self.dx)
            kx, ky, kzz = np.meshgrid(k1, k1, kz, indexing="ij")
            k2 = kx ** 2 + ky ** 2 + kzz ** 2
            k2[0, 0, 0] = 1.0
            self._kgrid = (kx, ky, kzz, k2)
        return self._kgrid

    # -- initial data (pm3d.py:40-73): transforms and the lattice interpolation run on the device,
    # in FP64 (a mixed-precision PM3D builds its initial data with an FP64 helper handle)
    def make_ic(self, mu, profile_spline, e=0.0, p=0.0, tail_rms=0.0, tail_k=(2.0, 10.0)):
        if self.precision != "fp64":
            eng = PM3D(self.N, self.L, self.Np, smooth=self.smooth, precision="fp64")
            eng.rng = self.rng
            eng.make_ic(mu, profile_spline, e, p, tail_rms, tail_k)
            self.set_state(eng.x, eng.p)
            for k in ("q", "delta_lin", "r_ell_q", "A", "tail_rms_grid"):
                setattr(self, k, getattr(eng, k))
            self._acc_a = None
            return self
        N, L = self.N, self.L
        kx, ky, kz, k2 = self._kgrids()
        xg = (np.arange(N) + 0.5) * self.dx - L / 2
        X, Y, Z = np.meshgrid(xg, xg, xg, indexing="ij")
        A = zeldovich_axes(e, p)
        r_ell = np.sqrt((X / A[0]) ** 2 + (Y / A[1]) ** 2 + (Z / A[2]) ** 2)
        zeta = mu * profile_spline(np.minimum(r_ell, profile_spline.x[-1]))
        zk = self.rfftn(zeta)
        if tail_rms > 0:
            kk = np.sqrt(k2)
            kk[0, 0, 0] = 0.0
            k_lo, k_hi = tail_k
            band = (kk >= k_lo) & (kk <= k_hi)
            Pk = np.where(band, 2 * np.pi ** 2 * tail_rms ** 2
                          / (np.maximum(kk, 1e-30) ** 3 * np.log(k_hi / k_lo)), 0.0)
            noise = self.rfftn(self.rng.standard_normal((N, N, N)))
            tail = noise * np.sqrt(Pk * N ** 3 / L ** 3)
            self.tail_rms_grid = float(np.std(self.irfftn(tail)))
            zk = zk + tail
        else:
            self.tail_rms_grid = 0.0
        dk = 0.4 * k2 / H_I ** 2 * zk
        dk[0, 0, 0] = 0.0
        phik = -dk / k2
        psi = [self.irfftn(1j * kc * phik) * -1.0 for kc in (kx, ky, kz)]
        q1 = (np.arange(self.Np) + 0.5) * (L / self.Np) - L / 2
        Q = np.stack(np.meshgrid(q1, q1, q1, indexing="ij"), axis=-1).reshape(-1, 3)
        Qw = (Q + L / 2) % L
        disp = np.stack([self.interp(f, Qw) for f in psi], axis=-1)
        self.set_state((Q + disp + L / 2) % L, H_I * disp)
        self.q = Q
        self.delta_lin = self.irfftn(dk)
        self.r_ell_q = np.sqrt(np.sum((Q / A) ** 2, axis=1))
        self.A = A
        self._acc_a = None
        return self

    # -- mesh operations (pm3d.py:76-99)
    def _cic_index(self, X1):
        """One axis of the CIC stencil (pm3d.py:76-79), computed on the device."""
        X1 = as_f64(X1).reshape(-1)
        Xs = np.zeros((len(X1), 3))
        Xs[:, 0] = X1
        key = np.empty(len(X1), np.int64)
        frac = np.empty((len(X1), 3))
        check(self._L.pm3d_cic_index(self._h, dptr(Xs), len(X1), iptr(key), dptr(frac)), self._h)
        i0 = key // (self.N * self.N)
        return i0, (i0 + 1) % self.N, frac[:, 0]

    def cic_keys(self, X):
        """(key (n,) int64, frac (n,3)) of the flat CIC base cell (pm3d.py:76-79,88) for positions X."""
        X = as_f64(X).reshape(-1, 3)
        key = np.empty(len(X), np.int64)
        frac = np.empty((len(X), 3))
        check(self._L.pm3d_cic_index(self._h, dptr(X), len(X), iptr(key), dptr(frac)), self._h)
        return key, frac

    def deposit(self, X):
        X = as_f64(X).reshape(-1, 3)
        rho = np.empty((self.N, self.N, self.N))
        check(self._L.pm3d_deposit(self._h, dptr(X), len(X), dptr(rho)), self._h)
        return rho

    def interp(self, F, X):
        F = as_f64(F).reshape(self.N, self.N, self.N)
        X = as_f64(X).reshape(-1, 3)
        out = np.empty(len(X))
        check(self._L.pm3d_interp(self._h, dptr(F), dptr(X), len(X), dptr(out)), self._h)
        return out

    def rfftn(self, a) -> np.ndarray:
        """numpy.fft.rfftn of an (N,N,N) real mesh on the device FFT."""
        a = as_f64(a).reshape(self.N, self.N, self.N)
        out = np.empty((self.N, self.N, self.N // 2 + 1), np.complex128)
        check(self._L.pm3d_rfftn(self._h, dptr(a), out.ctypes.data_as(_lib.PD)), self._h)
        return out

    def irfftn(self, c) -> np.ndarray:
        """numpy.fft.irfftn(c, s=(N,N,N)) on the device FFT."""
        c = np.ascontiguousarray(c, dtype=np.complex128).reshape(self.N, self.N, self.N // 2 + 1)
        out = np.empty((self.N, self.N, self.N))
        check(self._L.pm3d_irfftn(self._h, c.ctypes.data_as(_lib.PD), dptr(out)), self._h)
        return out

    def poisson_gradient(self, delta, a):
        """g_c = irfftn(-1j*k_c*phik) for a given delta mesh (pm3d.py:105-112); (N,N,N,3)."""
        delta = as_f64(delta).reshape(self.N, self.N, self.N)
        g = np.empty((self.N, self.N, self.N, 3))
        sm2 = (self.smooth * self.dx) ** 2 if self.smooth > 0 else 0.0
        check(self._L.pm3d_poisson_gradient(self._h, dptr(delta), -1.5 * H_I ** 2 / a, sm2, dptr(g)), self._h)
        return g

    def forces(self, a):
        """(acc (n,3), delta (N,N,N)) at the current positions (pm3d.py:101-114)."""
        acc = np.empty((self.n, 3))
        delta = np.empty((self.N, self.N, self.N))
        sm2 = (self.smooth * self.dx) ** 2 if self.smooth > 0 else 0.0
        bad = check(self._L.pm3d_forces(self._h, a, -1.5 * H_I ** 2 / a, sm2, dptr(acc), dptr(delta)), self._h)
        if bad:
            raise FloatingPointError(f"{bad} particles with non-finite state")
        return acc, delta

    # -- leapfrog in a (pm3d.py:117-129)
    def step(self, a, da):
        need = self._acc_a is None or self._acc_a != a
        sc = step_scalars(a, da, self.smooth, self.dx, need)
        check(self._L.pm3d_step(self._h, ctypes.byref(sc), None, None), self._h)
        a_new = a + da
        self._acc_a = a_new
        self.last_kick = (sc.c1, sc.den2)
        return a_new, DeltaMesh(self, self._L.pm3d_deposit_serial(self._h))

    def synchronize(self) -> int:
        """Wait for the device; returns the count of particles flagged non-finite since the last call."""
        return check(self._L.pm3d_synchronize(self._h), self._h)

    # -- device-resident schedules (graph replay; see pbh_amd.batch)
    def schedule_entries(self, a_start: float, da_frac: float, a_end: float, max_steps: int, prev: tuple):
        """Schedule table continuing from a_start with da = min(da_frac*a, a_end-a) (pm3d_peak.py:27);
        prev = (c1, den2) of the last step taken.  Returns (entries, a_values) with a_values[k] the
        scale factor after entry k.  Tables depend only on their arguments, the smoothing and dx, so
        boxes started alike share one (the table is read, never written, by set_schedule)."""
        ents, avals = _schedule_table(float(a_start), float(da_frac), float(a_end), int(max_steps),
                                      float(prev[0]), float(prev[1]), float(self.smooth), float(self.dx))
        return ents, list(avals)

    def set_schedule(self, entries, n: int):
        check(self._L.pm3d_set_schedule(self._h, entries, n), self._h)

    def schedule_step(self):
        check(self._L.pm3d_schedule_step(self._h), self._h)

    def schedule_status(self, with_history: bool = True):
        done = ctypes.c_int()
        check(self._L.pm3d_schedule_status(self._h, ctypes.byref(done), None), self._h)
        if not with_history:
            return done.value, None
        hist = np.empty(max(done.value, 1))
        check(self._L.pm3d_schedule_status(self._h, ctypes.byref(done), dptr(hist)), self._h)
        return done.value, hist[: done.value]

    def schedule_end(self, a_now: float) -> int:
        """Hand the state back to step(): returns the number of scheduled steps taken."""
        done = ctypes.c_int()
        check(self._L.pm3d_schedule_end(self._h, ctypes.byref(done)), self._h)
        self._acc_a = a_now
        return done.value

    def min_jacobian(self) -> float:
        """min det(dx/dq) over the Lagrangian lattice (first shell crossing when it reaches 0)."""
        v = ctypes.c_double()
        check(self._L.pm3d_min_jacobian(self._h, ctypes.byref(v)), self._h)
        return float(v.value)

    def totals(self) -> dict:
        """Global sums (FP64 / exact in every precision mode): mass of the most recent deposit in
        particle units and its exact deviation from n, and the total momentum sum_i p_i."""
        m, dev = ctypes.c_double(), ctypes.c_double()
        mom = np.zeros(3)
        check(self._L.pm3d_totals(self._h, ctypes.byref(m), ctypes.byref(dev), dptr(mom)), self._h)
        return {"mass": m.value, "mass_dev": dev.value, "momentum": mom}

    def device_bytes(self) -> int:
        return int(self._L.pm3d_device_bytes(self._h))

    def device_info(self) -> dict:
        """Device name, architecture and total memory of the handle's GPU."""
        name = ctypes.create_string_buffer(256)
        arch = ctypes.create_string_buffer(256)
        tot = ctypes.c_int64()
        check(self._L.pm3d_device_info(self._h, name, 256, arch, 256, ctypes.byref(tot)), self._h)
        return {"name": name.value.decode(), "arch": arch.value.decode(), "total_bytes": int(tot.value)}

    # -- per-kernel timing (hipEvents on the handle's stream)
    def set_timing(self, on: bool):
        check(self._L.pm3d_set_timing(self._h, 1 if on else 0), self._h)

    def analysis_sets(self, fracs=(0.1, 0.2, 0.3)) -> list:
        """Index sets of analyse (pm3d.py:136-146), uploaded to the device once per fracs: the 27
        innermost particles by r_ell_q, then per fraction r_ell_q < f*sqrt(6) (the 30 innermost
        when fewer than 30), each in ascending particle order as the reference's boolean masks."""
        key = tuple(float(f) for f in fracs)
        if getattr(self, "_an_key", None) == key:
            return self._an_sets
        if not hasattr(self, "r_ell_q"):
            raise RuntimeError("analyse needs the Lagrangian radii r_ell_q of make_ic")
        rm = np.sqrt(6.0)
        order = np.argsort(self.r_ell_q)
        sets = [np.sort(order[:27])]
        for f in key:
            sel = np.flatnonzero(self.r_ell_q < f * rm)
            if len(sel) < 30:
                sel = np.sort(order[:30])
            sets.append(sel)
        idx = np.ascontiguousarray(np.concatenate(sets), dtype=np.int32)
        offs = np.zeros(len(sets) + 1, np.int64)
        offs[1:] = np.cumsum([len(t) for t in sets])
        check(self._L.pm3d_set_analysis_sets(self._h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             iptr(offs), len(sets)), self._h)
        self._an_key, self._an_sets = key, sets
        return sets

    def device_analyse(self, radii):
        """Centre, per-set mean and outer-product sums, counts within radii and the density
        maximum of the current state (pm3d_analyse)."""
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        ns = len(self._an_sets) - 1
        centre = np.zeros(3)
        mom = np.zeros((max(ns, 1), 9))
        cnt = np.zeros(max(len(radii), 1), np.int64)
        rmax = ctypes.c_double()
        check(self._L.pm3d_analyse(self._h, dptr(radii), len(radii), dptr(centre), dptr(mom), iptr(cnt),
                                   ctypes.byref(rmax)), self._h)
        return centre, mom[:ns], cnt[: len(radii)], rmax.value

    def fft_tiles(self) -> dict:
        """FFT tile variant per pass chosen at creation (0: 16 KB buffers, radix <= 4; 1: 32 KB, radix 8/16)."""
        v = np.zeros(5, np.int32)
        check(self._L.pm3d_fft_tiles(self._h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 5), self._h)
        return dict(zip(("r2c_z", "y_fwd", "x", "y_inv", "c2r_z"), (int(t) for t in v)))

    def deposit_tiles(self) -> tuple:
        """Current deposit column shape (BX, BY, N); (0, 0, 0) for the atomic deposit.  Chosen at
        creation, then re-chosen by timed trials every 64 steps (every shape gives the same bits)."""
        v = np.zeros(3, np.int32)
        check(self._L.pm3d_deposit_tiles(self._h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), self._h)
        return tuple(int(t) for t in v)

    def deposit_trials(self) -> dict:
        """Shape trials of the step deposits so far: finished trials, shape changes, candidate mask."""
        v = np.zeros(3, np.int64)
        check(self._L.pm3d_deposit_trials(self._h, iptr(v)), self._h)
        return {"trials": int(v[0]), "changes": int(v[1]), "candidates": int(v[2])}

    def timings(self) -> dict:
        names = self._L.pm3d_kernel_names().decode().split(",")
        cnt = np.zeros(len(names), np.int64)
        ms = np.zeros(len(names))
        check(self._L.pm3d_get_timings(self._h, iptr(cnt), dptr(ms), len(names)), self._h)
        return {nm: (int(c), float(t)) for nm, c, t in zip(names, cnt, ms)}


def stream_probe(nbytes: int = 1 << 30, reps: int = 10) -> dict:
    """Attainable HBM bandwidth of the current GPU (GB/s): a hand-written 16-byte-vector streaming
    copy (read + write bytes) and a streaming read over `nbytes` (pm3d_stream_probe)."""
    L = _lib.load()
    cp, rd = ctypes.c_double(), ctypes.c_double()
    check(L.pm3d_stream_probe(int(nbytes), int(reps), ctypes.byref(cp), ctypes.byref(rd)), None)
    return {"copy_gbs": cp.value, "read_gbs": rd.value, "bytes": int(nbytes), "reps": int(reps)}


def analyse(pm: PM3D, a: float, fracs=(0.1, 0.2, 0.3), radii_cells=(1.0, 2.0, 4.0, 8.0)):
    """Inner Lagrangian-set axes and enclosed-mass compaction (pm3d.py:132-159), on the device.

    The index sets (the 27 innermost particles by r_ell_q, and r_ell_q < f*sqrt(6) per fraction,
    or the 30 innermost when a set has fewer than 30) are computed on the host from r_ell_q once
    and kept on the device; the centre, set means and second moments, enclosed counts and the
    density maximum are reductions over the particle state in HBM.  Only the 3 x 3 eigenvalue
    problems run on the host.
    """
    out = {"a": a}
    sets = pm.analysis_sets(tuple(fracs))
    radii = np.array([nc * pm.dx for nc in radii_cells], dtype=np.float64)
    _, mom, cnt, rho_max = pm.device_analyse(radii)
    for k, f in enumerate(fracs):
        m = len(sets[k + 1])
        xx, xy, xz, yy, yz, zz = mom[k][3:]
        inertia = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]) / m
        ev = np.sqrt(np.maximum(np.linalg.eigvalsh(inertia), 0.0)) * np.sqrt(5.0)
        M = pm.m_part * np.int64(m)
        axes_phys = a * ev[::-1]
        R_hoop = np.sqrt(0.5 * (axes_phys[0] ** 2 + axes_phys[1] ** 2))
        out[f"axes_{f}"] = axes_phys.tolist()
        out[f"axes_com_{f}"] = ev[::-1].tolist()
        out[f"M_{f}"] = M
        out[f"hoop_{f}"] = 2 * M / max(R_hoop, 1e-30)
    for nc, c in zip(radii_cells, cnt):
        R = a * nc * pm.dx
        M = pm.m_part * int(c)
        out[f"m_in_{nc:g}c"] = M
        out[f"C_{nc:g}c"] = 2 * M / R
    out["rho_max_over_mean"] = float(rho_max / (pm.Np ** 3 / pm.N ** 3))
    return out
