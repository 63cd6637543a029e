"""Write small golden vectors of the reference pm3d path (run in the build container only).

The reference repository is imported from /root/reference/pbhgr (read-only) to record input/output
vectors for the functions that closed forms and invariants do not pin on their own:
``PM3D.make_ic`` (profile clamp, tail band, rng order), ``PM3D.forces``/``step`` trajectories and
``analyse`` (pm3d.py:40-73,101-159), plus the driver's LTB scalars (ltb.py:16-35, profiles.py:50-62).

Output: tests/golden/pm3d_golden.npz (data only: inputs and expected outputs) and
tests/golden/pm3d_golden_meta.json.  Nothing here is imported by the product or the GPU tests' code
paths; the GPU tests never read /root/reference.

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/pbhgr")
OUT = Path(__file__).resolve().parent

CASES = [
    # name, N, Np, L_over_rm, smooth, mu, e, p, tail, seed, nsteps
    ("n8_plain", 8, 8, 8.0, 0.0, 0.1, 0.0, 0.0, 0.0, 0, 3),
    ("n8_tri_smooth", 8, 8, 8.0, 2.0, 0.3, 0.2, 0.05, 0.0, 0, 3),
    ("n8_tail", 8, 8, 8.0, 2.0, 0.2, 0.1, 0.0, 0.05, 3, 3),
    ("n8_np16", 8, 16, 8.0, 0.0, 0.3, 0.15, -0.05, 0.0, 0, 3),
    ("n16_tri", 16, 16, 8.0, 0.0, 0.5, 0.25, 0.1, 0.0, 0, 3),
    ("n16_box1", 16, 8, 1.0 / np.sqrt(6.0), 1.0, 0.4, 0.1, 0.05, 0.02, 7, 2),
]


def main():
    sys.path.insert(0, str(REF))
    from pbhgr.pm3d import PM3D, analyse          # noqa: E402  (reference, this container only)
    from pbhgr.profiles import profile_table     # noqa: E402
    from pbhgr.ltb import LTBGaussian             # noqa: E402

    table = profile_table("yoo")
    arrays = {}
    meta = {"cases": [], "ltb": []}
    for (name, N, Np, Lr, smooth, mu, e, p, tail, seed, nsteps) in CASES:
        rm = LTBGaussian(mu).r_m
        L = Lr * rm
        pm = PM3D(N, L, Np, seed=seed, smooth=smooth).make_ic(mu, table["spline"], e, p, tail, (2.0, 10.0))
        arrays[f"{name}/x0"] = pm.x.copy()
        arrays[f"{name}/p0"] = pm.p.copy()
        arrays[f"{name}/delta_lin"] = pm.delta_lin.copy()
        arrays[f"{name}/r_ell_q"] = pm.r_ell_q.copy()
        arrays[f"{name}/A"] = np.asarray(pm.A)
        acc0, delta0 = pm.forces(1.0)
        arrays[f"{name}/acc0"] = acc0
        arrays[f"{name}/delta0"] = delta0
        arrays[f"{name}/rho0"] = pm.deposit(pm.x)
        a = 1.0
        rows = []
        for k in range(nsteps):
            da = 0.02 * a * (1 + k)
            a, delta = pm.step(a, da)
            if k + 1 in (1, nsteps):
                arrays[f"{name}/x{k + 1}"] = pm.x.copy()
                arrays[f"{name}/p{k + 1}"] = pm.p.copy()
                arrays[f"{name}/delta{k + 1}"] = delta.copy()
            rows.append({"a": a, "da": da})
        ana = analyse(pm, a)
        meta["cases"].append(dict(name=name, N=N, Np=Np, L=L, smooth=smooth, mu=mu, e=e, p=p, tail=tail,
                                  seed=seed, nsteps=nsteps, steps=rows, tail_rms_grid=pm.tail_rms_grid,
                                  m_part=pm.m_part, analyse=ana))
    for mu in (0.1, 0.2, 0.3):
        g = LTBGaussian(mu)
        meta["ltb"].append(dict(mu=mu, r_m=float(g.r_m), t_H=float(g.t_H), tC0=float(g.tC(1e-6)),
                                tC_rm=float(g.tC(g.r_m))))
    r = np.linspace(0.0, 40.0, 13)
    meta["yoo_spline_r"] = (r * 0.37).tolist()
    meta["yoo_spline_v"] = table["spline"](r * 0.37).tolist()
    np.savez_compressed(OUT / "pm3d_golden.npz", **arrays)
    (OUT / "pm3d_golden_meta.json").write_text(json.dumps(meta, indent=1))
    print("wrote", len(arrays), "arrays")


if __name__ == "__main__":
    main()
