"""Newtonian 3D PM run of one triaxial peak on an MI355X: the drop-in for the reference driver
pbhgr/scripts/pm3d_peak.py, with the same flags, defaults, diagnostics cadence, printed rows and
JSON output ({out}/{tag}.json with args, rows, rm, t_H, tC0, dx, m_part).

    python -m pbh_amd.pm3d_peak --mu 0.1 --e 0 --p 0 --N 128 --Np 64 --L 8 --tend 40 [--out runs/pm3d]

Additional flags (not in the reference): --precision {fp64,mixed}; --collapse stops at the first
shell crossing (min det(dx/dq) <= 0, Np lattice) and records t_c in the JSON.  Every run also
writes {out}/{tag}_provenance.json (pbh_amd.provenance: git commit, compiler and flags, host, GPU,
parameters, seed, times, peak RSS, steps, particle-steps/s, t_c).
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

from .ltb import LTBGaussian, profile_table
from .pm3d import PM3D, T_I, analyse
from .provenance import RunRecord


def parser():
    # flags and defaults of pm3d_peak.py:8-13
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--mu", type=float, default=0.1)
    ap.add_argument("--e", type=float, default=0.0)
    ap.add_argument("--p", type=float, default=0.0)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--Np", type=int, default=64)
    ap.add_argument("--L", type=float, default=8.0)
    ap.add_argument("--tend", type=float, default=40.0)
    ap.add_argument("--tail", type=float, default=0.0)
    ap.add_argument("--k1", type=float, default=2.0)
    ap.add_argument("--k2", type=float, default=10.0)
    ap.add_argument("--profile", default="yoo")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--smooth", type=float, default=2.0)
    ap.add_argument("--da_frac", type=float, default=0.02)
    ap.add_argument("--out", default="runs/pm3d")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--collapse", action="store_true")
    return ap


def run_tag(args) -> str:
    """Output tag of pm3d_peak.py:21."""
    tag = f"mu{args.mu:g}_e{args.e:g}_p{args.p:g}_N{args.N}_Np{args.Np}_L{args.L:g}_sm{args.smooth:g}"
    if args.tail > 0:
        tag += f"_tail{args.tail:g}_k{args.k1:g}-{args.k2:g}_s{args.seed}"
    if args.profile != "yoo":
        tag += f"_{args.profile}"
    return tag


def main(argv=None):
    args = parser().parse_args(argv)
    Path(args.out).mkdir(parents=True, exist_ok=True)
    record = RunRecord(args.out, vars(args), args.seed, name=f"{run_tag(args)}_provenance.json")
    table = profile_table(args.profile)
    ref = LTBGaussian(args.mu)
    rm, t_H = ref.r_m, ref.t_H
    tC0 = ref.tC(1e-6)
    if not np.isfinite(tC0):
        tC0 = ref.tC(rm)
    pm = PM3D(args.N, args.L * rm, args.Np, seed=args.seed, smooth=args.smooth, precision=args.precision).make_ic(
        args.mu, table["spline"], args.e, args.p, args.tail, (args.k1, args.k2))
    tag = run_tag(args)
    print(f"PM3D {tag}: box {args.L} r_m = {args.L * rm:.2f}/k, dx = {pm.dx:.4f}/k ({pm.dx / rm:.4f} r_m), "
          f"{args.Np ** 3} particles, m_part = {pm.m_part:.3e}; axes A = {np.round(pm.A, 3)}; "
          f"t_C(0) = {tC0 / t_H:.2f} t_H, run to {args.tend} t_H; central delta_lin(a=1) = "
          f"{pm.delta_lin[args.N // 2, args.N // 2, args.N // 2]:.4f} (expected 0.4 mu/25 = {0.4 * args.mu / 25:.4f}); "
          f"tail rms on grid = {pm.tail_rms_grid:.4f}", flush=True)
    a = 1.0
    t = T_I
    step = 0
    rows = []
    t0 = time.time()
    next_diag = 0.0
    a_end = (args.tend * t_H / T_I) ** (2.0 / 3.0)
    collapse = None
    j_prev = pm.min_jacobian() if args.collapse else None
    while a < a_end:
        da = min(args.da_frac * a, a_end - a)
        a_prev = a
        a, delta = pm.step(a, da)
        t = T_I * a ** 1.5
        step += 1
        crossed = False
        if args.collapse:
            j = pm.min_jacobian()
            if j <= 0.0:
                a_sc = a_prev + (a - a_prev) * j_prev / (j_prev - j)
                collapse = {"a_sc": a_sc, "t_c": T_I * a_sc ** 1.5, "t_c_over_tH": T_I * a_sc ** 1.5 / t_H, "step": step}
                crossed = True
            j_prev = j
        if t / t_H >= next_diag or a >= a_end or crossed:
            d = analyse(pm, a)
            d.update(t=t, t_over_tH=t / t_H, t_over_tC0=t / tC0, step=step, delta_max=float(delta.max()))
            rows.append(d)
            print(f"step {step:5d} t/t_H={t / t_H:7.3f} (t/tC0={t / tC0:.3f}) a={a:8.2f} delta_max={delta.max():8.1f} | "
                  f"comoving axes(0.1/0.2/0.3 r_m)={np.round(d['axes_com_0.1'], 3)} {np.round(d['axes_com_0.2'], 3)} "
                  f"{np.round(d['axes_com_0.3'], 3)} hoop(0.2)={d['hoop_0.2']:.4f} | m(<1,2,4,8 cells)="
                  f"{d['m_in_1c']:.2f} {d['m_in_2c']:.2f} {d['m_in_4c']:.2f} {d['m_in_8c']:.2f} C(<2c)={d['C_2c']:.4f} | "
                  f"{time.time() - t0:.0f}s", flush=True)
            next_diag = t / t_H + max(0.5, 0.02 * tC0 / t_H)
        if crossed:
            break
    out = dict(args=vars(args), rows=rows, rm=rm, t_H=t_H, tC0=tC0, dx=pm.dx, m_part=pm.m_part)
    if args.collapse:
        out["collapse"] = collapse
    with open(f"{args.out}/{tag}.json", "w") as fh:
        json.dump(out, fh)
    loop_s = time.time() - t0
    record.finish(pm, extra={"steps": step, "a_end": a, "loop_wall_s": loop_s,
                             "particle_steps_per_s": args.Np ** 3 * step / max(loop_s, 1e-9),
                             "collapse": collapse, "output": f"{args.out}/{tag}.json"})
    print("done", f"{time.time() - t0:.0f}s")
    return out


if __name__ == "__main__":
    main()
