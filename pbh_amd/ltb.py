"""Host helpers the pm3d driver needs once per run: the 'yoo' curvature profile spline and the exact
LTB scalars that set the box size, the time unit t_H and the collapse-time normaliser t_C(0).

Restated from pbhgr/pbhgr/profiles.py:50-62 ('yoo' only) and pbhgr/pbhgr/ltb.py:16-35 (the parts the
driver reads: r_m, t_H, tC).  Units G = c = 1, k = 1, a_i = 1, H_i = 5k, t_i = 2/(3 H_i).
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import CubicSpline


def profile_table(name: str = "yoo") -> dict:
    """Psi(r) = exp(-r^2/6) on r in [0, 40] (4001 points), clamped cubic spline (profiles.py:50-62)."""
    if name != "yoo":
        raise ValueError(f"profile {name!r} is not provided (only 'yoo')")
    r = np.linspace(0.0, 40.0, 4001)
    d = {"r": r, "Psi": np.exp(-r ** 2 / 6)}
    d["spline"] = CubicSpline(d["r"], d["Psi"], bc_type=((1, 0.0), (1, 0.0)))
    d["name"] = name
    return d


class LTBGaussian:
    """Growing-mode LTB scalars for zeta = mu exp(-k^2 r^2/6) (ltb.py:16-35)."""

    def __init__(self, mu: float, k: float = 1.0, H_i: float = 5.0):
        self.mu, self.k, self.H_i = mu, k, H_i
        self.t_i = 2.0 / (3.0 * H_i)
        self.eps_i = k / H_i
        self.t_H = self.t_i * (np.sqrt(6.0) / self.eps_i) ** 3
        self.r_m = np.sqrt(6.0) / k

    def zeta(self, r):
        return self.mu * np.exp(-self.k ** 2 * r ** 2 / 6)

    def dzeta(self, r):
        return -(self.k ** 2 * r / 3) * self.zeta(r)

    def m(self, r):
        """Misner-Sharp mass (H_i^2/2)(e^zeta r)^3."""
        return 0.5 * self.H_i ** 2 * (np.exp(self.zeta(r)) * r) ** 3

    def E(self, r):
        rz = r * self.dzeta(r)
        return rz * (1 + 0.5 * rz)

    def T(self, r):
        return self.m(r) / (-2 * self.E(r)) ** 1.5

    def tC(self, r):
        """Cycloid collapse time 2 pi T(r)."""
        return 2 * np.pi * self.T(r)
