"""Run a PM3D box to its first shell crossing and report the collapse time.

pm3d has no shell-crossing diagnostic (its driver reports t/t_C0 and delta_max only,
pbhgr/scripts/pm3d_peak.py:29-33).  This module defines one: the minimum over Lagrangian lattice
cells of det(dx/dq) (forward differences along the q-lattice, minimum-image unwrapped; 1 for the
undisplaced lattice).  The first shell crossing is the first step at which that minimum is <= 0;
its scale factor a_sc is interpolated linearly in a between the last two steps, and
t_c = T_I * a_sc**1.5 (EdS, pm3d.py:16; t = T_I a^{3/2} as in pm3d_peak.py:28).
"""
from __future__ import annotations

from dataclasses import dataclass

from .pm3d import T_I


@dataclass
class CollapseResult:
    crossed: bool
    a_sc: float        # interpolated scale factor of the first shell crossing (nan if none)
    t_c: float         # T_I * a_sc**1.5
    steps: int
    a_last: float
    jmin_last: float
    history: list      # (a, jmin) per step


def crossing_a(a0: float, j0: float, a1: float, j1: float) -> float:
    """Zero of the linear interpolant of min det(dx/dq) between (a0, j0 > 0) and (a1, j1 <= 0)."""
    return a0 + (a1 - a0) * j0 / (j0 - j1)


def run_to_crossing(pm, a_start: float, da_frac: float, a_max: float, max_steps: int = 100000) -> CollapseResult:
    """Step pm (PM3D, or anything with step(a, da) and min_jacobian()) with da = da_frac*a until
    the first shell crossing or a_max."""
    a = a_start
    j_prev = pm.min_jacobian()
    hist = [(a, j_prev)]
    k = 0
    while a < a_max and k < max_steps:
        da = min(da_frac * a, a_max - a)
        a_new, _ = pm.step(a, da)
        j = pm.min_jacobian()
        hist.append((a_new, j))
        k += 1
        if j <= 0.0:
            a_sc = crossing_a(a, j_prev, a_new, j)
            return CollapseResult(True, a_sc, T_I * a_sc ** 1.5, k, a_new, j, hist)
        a, j_prev = a_new, j
    return CollapseResult(False, float("nan"), float("nan"), k, a, j_prev, hist)

