"""Benchmark of the pm3d particle-mesh step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload peak128] [--no-cpu]

Workload (default "peak128"): periodic box L=1, 128^3 mesh, 128^3 = 2,097,152 particles, one
triaxial Gaussian peak mu=0.2, sigma=(0.10,0.09,0.08) L, Zel'dovich growing mode at a_i=0.01 in EdS,
seed 1, FP64 (positions, momenta, FFT, forces; 128-bit fixed-point deposit at 2^-52).  A step is one KDK
leapfrog step of PM3D.step (deposit, FFT Poisson, gather, kick, drift) with inputs resident in HBM.
The step's gather runs at the start of the next step, fused with its kicks and drift, so each of
the K timed steps runs one gather (the last warm-up step's at the start of the first).

Prints one JSON line: value = particle-steps/s = n*K/wall over K timed steps; `roofline` for the
kernel with the largest share of step time (algorithmic bytes per launch / average launch time
from the start/stop hipEvents that each launch's dispatch stamps on the library's stream); `cpu_baseline` = the CPU FP64 oracle on the same workload,
3 steps, single-threaded numpy (1 core).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s + HBM GB/s frac; batched solves/s, p50/p99 latency vs plain FP64"
PEAK_HBM_GBS = 8000.0


def algorithmic_bytes(N: int, n: int, prec: str) -> dict:
    """Bytes each kernel must move once per step (one read of every input, one write of every
    output).  FP64: positions/momenta/accelerations 3 x 8 B SoA, 16 B fixed-point cells, complex128
    half spectra (Mh = N*N*(N/2+1) points), interleaved 3 x 8 B gradient mesh.  Mixed: positions
    3 x (4 B cell + 4 B offset), FP32 momenta/accelerations/spectra/gradients, 8 B cells.  The mixed
    drift rewrites a particle's cell only when it moves to another cell, so kick_drift and
    gather_kick count the offsets (3 x 4 B) as the positions they write."""
    M = N ** 3
    Mh = N * N * (N // 2 + 1)
    r = 8 if prec == "fp64" else 4
    c = 2 * r
    q = 16 if prec == "fp64" else 8                    # fixed-point cell: 128-bit (FP64) / 64-bit (mixed)
    return {
        "kick_drift": 5 * 3 * r * n,                   # read x, p, acc; write x, p
        "deposit": 24 * n + q * M,                     # read x (24 B both modes); write the mesh
        "r2c_z": q * M + c * Mh,                       # read mesh; write spectrum
        "y_fwd": 2 * c * Mh,
        "x_poisson": c * Mh + 3 * c * Mh,              # read 1, write 3 spectra
        "y_inv": 2 * 3 * c * Mh,
        "c2r_z": 3 * c * Mh + 3 * r * M,               # read 3 spectra; write the gradient meshes
        "gather": 24 * n + 3 * r * M + 3 * r * n,      # read x (24 B both modes), gradient; write acc
        # z and y fused per x plane (N <= 128 FP64, N <= 256 FP32)
        "fwd_zy": q * M + c * Mh,                      # read mesh; write spectrum
        "inv_yz": 3 * c * Mh + 3 * r * M,              # read 3 spectra; write the gradient meshes
        # the previous step's gather + kicks + drift: read x (24 B both modes), p, gradient; write x, p
        "gather_kick": 24 * n + 3 * r * n + 3 * r * M + 2 * 3 * r * n,
    }


def load_traffic(kernel: str, workload: str):
    f = ROOT / "profiles" / "step_kernels_traffic_peak128_fp64_peak256_mixed.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        return d.get(workload, {}).get(kernel)
    except Exception:
        return None


def cpu_baseline(st: dict, spec, nsteps: int, da_frac: float) -> dict:
    from oracle.pm3d_oracle import OraclePM3D
    ref = OraclePM3D(spec.N, spec.L, spec.n_side)
    ref.x, ref.p = st["x"].copy(), st["p"].copy()
    a = st["a_i"]
    a, _ = ref.step(a, da_frac * a)          # first step also computes the opening forces
    t0 = time.perf_counter()
    for _ in range(nsteps):
        a, _ = ref.step(a, da_frac * a)
    dt = time.perf_counter() - t0
    n = spec.n_side ** 3
    return {"value": n * nsteps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{spec.name}: {nsteps} steady-state KDK steps of the numpy FP64 oracle "
                      f"({dt / nsteps:.2f} s/step, single-threaded numpy; host reports "
                      f"{os.cpu_count()} CPUs, affinity {len(os.sched_getaffinity(0))})"}


def copy_probe(mib: int = 512, reps: int = 20) -> dict:
    """Plain device-to-device copy of `mib` MiB (torch's copy kernel), timed with events: the
    attainable HBM rate on this box, and the kernel whose FETCH_SIZE / WRITE_SIZE counters scale
    the traffic figures."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * n * 8
    return {"copy_probe": {"read_bytes": n * 8, "write_bytes": n * 8, "avg_us": ms * 1e3,
                           "gbs": nbytes / (ms * 1e-3) / 1e9}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="peak128")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--da-frac", type=float, default=0.0125)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true", help="time a plain 512 MiB device copy and exit")
    args = ap.parse_args()
    if args.copy_probe:
        print(json.dumps(copy_probe()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    from pbh_amd import ic
    from pbh_amd.pm3d import PM3D

    spec = ic.named(args.workload)
    st = ic.make(spec)
    n = spec.n_side ** 3
    pm = PM3D(spec.N, spec.L, spec.n_side, precision=args.precision)
    pm.set_state(st["x"], st["p"])
    a = st["a_i"]
    for _ in range(args.warmup):
        a, _ = pm.step(a, args.da_frac * a)
    pm.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, dm = pm.step(a, args.da_frac * a)
    bad = pm.synchronize()
    wall = time.perf_counter() - t0
    if bad:
        raise SystemExit(f"{bad} particles went non-finite")

    # per-kernel times from dispatch-stamped hipEvents on the library's stream, same number of steps
    pm.set_timing(True)
    for _ in range(args.steps):
        a, dm = pm.step(a, args.da_frac * a)
    pm.synchronize()
    tim = pm.timings()
    pm.set_timing(False)
    nbytes = algorithmic_bytes(spec.N, n, args.precision)
    kernels = {}
    for name, (cnt, ms) in tim.items():
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        entry = {"launches": cnt, "avg_us": avg_ms * 1e3, "share": 0.0}
        if name in nbytes:
            gbs = nbytes[name] / (avg_ms * 1e-3) / 1e9
            entry.update(bytes=nbytes[name], gbs=gbs, frac=gbs / PEAK_HBM_GBS)
        kernels[name] = entry
    tot = sum(tim[k][1] for k in kernels)
    for k in kernels:
        kernels[k]["share"] = tim[k][1] / tot
    dom = max((k for k in kernels if k in nbytes), key=lambda k: tim[k][1])
    kd = kernels[dom]
    roofline = {"bound": "hbm", "achieved": kd["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": kd["frac"],
                "traffic": load_traffic(dom, args.workload), "kernel": dom,
                "bytes_per_launch": kd["bytes"], "avg_launch_us": kd["avg_us"]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(st, spec, args.cpu_steps, args.da_frac)

    ms = wall / args.steps * 1e3
    out = {
        "metric": METRIC, "value": n * args.steps / wall, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (Zel'dovich initial conditions generated on the host)",
        "config": {"workload": f"{spec.name}: {spec.N}^3 mesh, {spec.n_side}^3 particles, mu={spec.mu}, "
                               f"sigma={tuple(round(s, 4) for s in spec.sigma)} L, a_i={spec.a_i}, seed {spec.seed}, "
                               f"{args.precision}, da/a={args.da_frac}",
                   "N": spec.N, "Np": spec.n_side, "precision": args.precision, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "step_kernel_sum_us": tot / args.steps * 1e3,
        "fft_tiles": pm.fft_tiles(),
        "deposit_tiles": list(pm.deposit_tiles()),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
