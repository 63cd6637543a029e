"""Benchmark of the pm3d particle-mesh step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload peak128] [--no-cpu] [--no-extras]

Headline workload "peak128": periodic box L=1, 128^3 mesh, 128^3 = 2,097,152 particles, one
triaxial Gaussian peak mu=0.2, sigma=(0.10,0.09,0.08) L, Zel'dovich growing mode at a_i=0.01 in EdS,
seed 1, FP64 (positions, momenta, FFT, forces; 128-bit fixed-point deposit at 2^-52), stepped with
da = 0.0125 a to its first shell crossing (171 steps).  A step is one KDK leapfrog step of
PM3D.step (deposit, FFT Poisson, gather, kick, drift) with the state resident in HBM.

The run is first stepped to its first shell crossing (step S) to find S; the state is then reset
to the initial conditions, W warm-up steps run, untimed steps bring it to step S - K, and the K
timed steps are the ones that END at the first shell crossing (the collapse phase, where particles
crowd into few cells): value = particle-steps/s = n*K/wall.  A second pass over the same K steps
times every step kernel with the start/stop events its dispatch stamps (`kernels`, `roofline` for
the kernel with the largest share).  `cpu_baseline` = the numpy FP64 oracle on the same workload,
20 steps from the initial conditions (about 11 s of CPU work), single-threaded (1 core).

Extra keys of the same line (skipped with --no-extras):
  early_window     K steps after the warm-up (from a_i), the same kernel table;
  run_to_crossing  all S steps with min det(dx/dq) read after each, as a collapse run does;
  mixed256         peak256 (256^3, 16.8M particles, random field) in the mixed precision mode, the
                   same late window (K steps ending at its crossing) with its own roofline;
  batched          64 peaks (seeds 100-163, meshes 32-128, mixed) to first crossing: one graph
                   replay per step for all, against one at a time (mixed and FP64);
  serving          --serve-n 256^3 requests of the serving workload at --serve-rate req/s
                   (Poisson arrivals): p50/p99 time-to-collapse latency and solves/s, against plain
                   FP64 one-at-a-time runs of a sample of the same requests;
  totals           FP64 global mass (exact fixed-point total) and momentum sums at the crossing.
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


def say(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(N: int, n: int, prec: str) -> dict:
    """Bytes each kernel must move once per step (one read of every input, one write of every
    output).  FP64: positions/momenta/accelerations 3 x 8 B SoA, 16 B fixed-point cells, complex128
    half spectra (Mh = N*N*(N/2+1) points), 3 x 8 B gradient meshes.  Mixed: positions 3 x (4 B cell
    + 4 B offset), FP32 momenta/accelerations/spectra/gradients, 8 B cells.  The mixed drift
    rewrites a particle's cell only when it moves to another cell, so kick_drift and gather_kick
    count the offsets (3 x 4 B) as the positions they write."""
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
    """The library's hand-written streaming copy (16-byte vector loads and stores) of `mib` MiB and
    streaming read: the attainable HBM rates of this box, and a kernel of known size whose
    FETCH_SIZE / WRITE_SIZE counters scale the traffic figures."""
    from pbh_amd.pm3d import stream_probe
    r = stream_probe(mib << 20, reps)
    return {"copy_probe": {"read_bytes": mib << 20, "write_bytes": mib << 20, "copy_gbs": r["copy_gbs"],
                           "read_gbs": r["read_gbs"], "reps": reps}}


# ----------------------------------------------------------------------------- one workload
class Run:
    """One box of a workload, resettable to its initial state: `reset()` puts the initial state
    back, `step()` advances with da = da_frac * a."""

    def __init__(self, spec, precision: str, da_frac: float, host_ic: bool, frac_bits: int = 0):
        from pbh_amd import ic
        from pbh_amd.pm3d import PM3D
        self.spec, self.precision, self.da_frac = spec, precision, da_frac
        self.n = spec.n_side ** 3
        self.pm = PM3D(spec.N, spec.L, spec.n_side, precision=precision, frac_bits=frac_bits)
        if host_ic:
            self.st = ic.make(spec)
            self.eng = None
        else:
            self.st = None
            self.eng = PM3D(spec.N, spec.L, spec.N, precision="fp64")
            ic.device_ic(self.eng, spec)
        self.reset()

    def reset(self):
        if self.st is not None:
            self.pm.set_state(self.st["x"], self.st["p"])
        elif self.precision == "fp64":
            self.pm.set_state(self.eng.x, self.eng.p)
        else:
            from pbh_amd._lib import check
            check(self.pm._L.pm3d_copy_state(self.pm._h, self.eng._h), self.pm._h)
        self.pm._acc_a = None
        self.a = self.spec.a_i
        self.k = 0

    def step(self):
        self.a, dm = self.pm.step(self.a, self.da_frac * self.a)
        self.k += 1
        return dm

    def to_crossing(self, max_steps: int = 5000):
        """Steps from the initial state to the first step with min det(dx/dq) <= 0 (each step's
        minimum read back, as a collapse run does).  Returns (steps, wall seconds)."""
        self.reset()
        self.pm.synchronize()
        t0 = time.perf_counter()
        while self.k < max_steps:
            self.step()
            if self.pm.min_jacobian() <= 0.0:
                break
        wall = time.perf_counter() - t0
        return self.k, wall

    def timed(self, nsteps: int) -> float:
        self.pm.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            self.step()
        bad = self.pm.synchronize()
        if bad:
            raise SystemExit(f"{bad} particles went non-finite")
        return time.perf_counter() - t0

    def kernel_table(self, nsteps: int, workload: str) -> dict:
        """Per-kernel average duration and algorithmic GB/s over the next nsteps steps."""
        pm = self.pm
        pm.set_timing(True)
        pm.timings()
        self.timed(nsteps)
        tim = pm.timings()
        pm.set_timing(False)
        nbytes = algorithmic_bytes(self.spec.N, self.n, self.precision)
        kernels = {}
        for name, (cnt, ms) in tim.items():
            if cnt == 0:
                continue
            entry = {"launches": cnt, "avg_us": ms / cnt * 1e3, "share": 0.0}
            if name in nbytes:
                gbs = nbytes[name] / (ms / cnt * 1e-3) / 1e9
                entry.update(bytes=nbytes[name], gbs=gbs, frac=gbs / PEAK_HBM_GBS)
            kernels[name] = entry
        tot = sum(tim[k][1] for k in kernels)
        for k in kernels:
            kernels[k]["share"] = tim[k][1] / tot
        dom = max((k for k in kernels if k in nbytes), key=lambda k: tim[k][1])
        kd = kernels[dom]
        roof = {"bound": "hbm", "achieved": kd["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": kd["frac"],
                "traffic": load_traffic(dom, workload), "kernel": dom,
                "bytes_per_launch": kd["bytes"], "avg_launch_us": kd["avg_us"]}
        return {"kernels": kernels, "roofline": roof, "step_kernel_sum_us": tot / nsteps * 1e3}


def late_window(run: Run, K: int, W: int, workload: str) -> dict:
    """Crossing step S; then W warm-up steps, the K steps after them (early window), untimed steps
    to S - K, the K steps ending at S (late window), and the kernel table of the same K steps."""
    S, wall_run = run.to_crossing()
    run.reset()
    for _ in range(W):
        run.step()
    early_wall = run.timed(K)
    run.reset()
    for _ in range(W):
        run.step()
    early_tab = run.kernel_table(K, workload)
    run.reset()
    start = max(W, S - K)
    while run.k < start:
        run.step()
    late_wall = run.timed(K)
    run.reset()
    while run.k < start:
        run.step()
    late_tab = run.kernel_table(K, workload)
    return {"S": S, "late_first_step": start + 1, "late_wall": late_wall, "late_tab": late_tab,
            "early_wall": early_wall, "early_tab": early_tab, "run_wall": wall_run}


def far_share(tab: dict) -> float:
    k = tab["kernels"]
    return k.get("deposit_far", {}).get("share", 0.0)


# ----------------------------------------------------------------------------------- extras
def bench_mixed256(K: int, W: int) -> dict:
    from pbh_amd import ic
    spec = ic.named("peak256")
    run = Run(spec, "mixed", 0.0125, host_ic=False)
    r = late_window(run, K, W, "peak256")
    n = run.n
    out = {"workload": "peak256: 256^3 mesh, 256^3 particles, mu=0.1, e=0.2, p=0, axes 0.1 L, random field "
                       "rms 0.01 mu, seed 2, mixed precision, da/a=0.0125",
           "crossing_step": r["S"], "late_first_step": r["late_first_step"],
           "ms_per_step": r["late_wall"] / K * 1e3, "particle_steps_per_s": n * K / r["late_wall"],
           "roofline": r["late_tab"]["roofline"], "kernels": r["late_tab"]["kernels"],
           "far_pass_share": far_share(r["late_tab"]),
           "early_window": {"ms_per_step": r["early_wall"] / K * 1e3, "particle_steps_per_s": n * K / r["early_wall"],
                            "roofline_frac": r["early_tab"]["roofline"]["frac"],
                            "far_pass_share": far_share(r["early_tab"])},
           "run_to_crossing": {"steps": r["S"], "wall_s": r["run_wall"],
                               "particle_steps_per_s": n * r["S"] / r["run_wall"]}}
    run.pm.close()
    run.eng.close()
    return out


def bench_fp64_frac40(K: int, W: int) -> dict:
    """The headline workload with the FP64 deposit on 64-bit cells at 2^-40 (frac_bits = 40, opt-in)."""
    from pbh_amd import ic
    spec = ic.named("peak128")
    run = Run(spec, "fp64", 0.0125, host_ic=True, frac_bits=40)
    r = late_window(run, K, W, "peak128")
    n = run.n
    out = {"workload": "peak128 as the headline, FP64 with the deposit on 64-bit cells at 2^-40 "
                       "(frac_bits=40: half the mesh bytes, about 1e-12 of a particle per cell)",
           "crossing_step": r["S"], "ms_per_step": r["late_wall"] / K * 1e3,
           "particle_steps_per_s": n * K / r["late_wall"],
           "kernels_avg_us": {k: v["avg_us"] for k, v in r["late_tab"]["kernels"].items()},
           "early_window": {"ms_per_step": r["early_wall"] / K * 1e3, "particle_steps_per_s": n * K / r["early_wall"]}}
    run.pm.close()
    return out


def bench_batched() -> dict:
    from pbh_amd.bench_batch import batched_vs_plain
    r = batched_vs_plain()
    return {"workload": "64 peaks, seeds 100-163, N from {32,64,96,128}, mu~U[0.05,0.5], e~U[0,0.3], "
                        "p~U[-e,e], axes 0.1 L, mixed, da/a=0.02, to first shell crossing",
            "solves_per_s": r["batched_solves_per_s"], "plain_mixed_solves_per_s": r["plain_solves_per_s"],
            "plain_fp64_solves_per_s": r["plain_fp64_solves_per_s"],
            "speedup_vs_plain_fp64": r["speedup_vs_plain_fp64"], "speedup_vs_plain_mixed": r["speedup"],
            "crossed": r["crossed"], "identical_to_plain": r["identical_results"],
            "max_rel_t_c_mixed_vs_fp64": r["max_rel_t_c_mixed_vs_fp64"], "box_steps": r["steps_total"]}


def bench_serving(rate: float, n: int, plain_sample: int) -> dict:
    from pbh_amd.bench_serve import plain_fp64
    from pbh_amd.serve import CollapseServer, Request, config_requests, poisson_arrivals, summarize
    specs = config_requests(n, N=256, seed0=1000)
    times = poisson_arrivals(rate, n, seed=int(rate))
    srv = CollapseServer(precision="mixed", poll=8)
    done = srv.run([(float(t), Request(rid=k, spec=s, arrival=float(t))) for k, (t, s) in enumerate(zip(times, specs))])
    row = summarize(done)
    plain = plain_fp64(specs[:plain_sample])
    row.update(workload=f"{n} requests of 256^3 peaks (mu~U[0.05,0.5], e~U[0,0.3], p~U[-e,e], random field "
                        f"rms 0.01 mu), mixed, da/a=0.02, to first shell crossing, Poisson arrivals",
               rate_req_per_s=rate, max_active=srv.max_active, graph_builds=srv.graph_builds,
               held_GB=srv.held / 1e9, plain_fp64_solves_per_s=plain["solves_per_s"],
               plain_fp64_sample=plain_sample, speedup_vs_plain_fp64=row["solves_per_s"] / plain["solves_per_s"])
    return row


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="peak128")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--da-frac", type=float, default=0.0125)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-n", type=int, default=200, help="serving: number of requests")
    ap.add_argument("--serve-plain", type=int, default=8, help="serving: requests run one at a time in FP64")
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

    spec = ic.named(args.workload)
    K, W = args.steps, args.warmup
    say(f"{spec.name}: crossing step, early window, late window")
    run = Run(spec, args.precision, args.da_frac, host_ic=(spec.N <= 128))
    r = late_window(run, K, W, args.workload)
    n = run.n
    tot = run.pm.totals()
    cpu = None
    if not args.no_cpu:
        say("cpu baseline")
        st = run.st if run.st is not None else {"x": run.eng.x, "p": run.eng.p, "a_i": spec.a_i}
        cpu = cpu_baseline(st, spec, args.cpu_steps, args.da_frac)
    late = r["late_tab"]
    out = {
        "metric": METRIC, "value": n * K / r["late_wall"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": r["late_wall"] / K * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (Zel'dovich initial conditions of a triaxial peak)",
        "config": {"workload": f"{spec.name}: {spec.N}^3 mesh, {spec.n_side}^3 particles, mu={spec.mu}, "
                               f"sigma={tuple(round(s, 4) for s in spec.sigma)} L, a_i={spec.a_i}, seed {spec.seed}, "
                               f"{args.precision}, da/a={args.da_frac}; timed: the {K} steps ending at the first "
                               f"shell crossing (steps {r['late_first_step']}-{r['late_first_step'] + K - 1} of "
                               f"{r['S']}), after {W} warm-up steps from a_i",
                   "N": spec.N, "Np": spec.n_side, "precision": args.precision, "parallelism": "single GPU"},
        "roofline": late["roofline"],
        "cpu_baseline": cpu,
        "kernels": late["kernels"],
        "step_kernel_sum_us": late["step_kernel_sum_us"],
        "far_pass_share": far_share(late),
        "early_window": {"steps": f"{W + 1}-{W + K}", "ms_per_step": r["early_wall"] / K * 1e3,
                         "particle_steps_per_s": n * K / r["early_wall"], "roofline": r["early_tab"]["roofline"],
                         "far_pass_share": far_share(r["early_tab"]),
                         "kernels_avg_us": {k: v["avg_us"] for k, v in r["early_tab"]["kernels"].items()}},
        "run_to_crossing": {"steps": r["S"], "wall_s": r["run_wall"], "ms_per_step": r["run_wall"] / r["S"] * 1e3,
                            "particle_steps_per_s": n * r["S"] / r["run_wall"],
                            "note": "every step followed by a min det(dx/dq) read-back, as a collapse run does"},
        "totals": {"mass_minus_n": tot["mass_dev"], "momentum": tot["momentum"].tolist(),
                   "note": "after the crossing step; exact fixed-point mass total and FP64 momentum sum"},
        "fft_tiles": run.pm.fft_tiles(),
        "deposit_tiles": list(run.pm.deposit_tiles()),
        "deposit_trials": run.pm.deposit_trials(),
    }
    run.pm.close()
    from pbh_amd.pm3d import stream_probe
    att = stream_probe(1 << 30, 10)
    out["attainable"] = {**att, "note": "hand-written streaming copy / read of 1 GiB (pm3d_stream_probe)"}
    out["roofline"]["frac_of_attainable_copy"] = out["roofline"]["achieved"] / att["copy_gbs"]
    if not args.no_extras:
        say("FP64 on 64-bit cells")
        out["fp64_frac40"] = bench_fp64_frac40(K, W)
        say("mixed 256^3")
        out["mixed256"] = bench_mixed256(K, W)
        say("batched 64 peaks")
        out["batched"] = bench_batched()
        say(f"serving {args.serve_n} requests at {args.serve_rate} req/s")
        out["serving"] = bench_serving(args.serve_rate, args.serve_n, args.serve_plain)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
