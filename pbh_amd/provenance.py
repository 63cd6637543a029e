"""Run record of a pm3d run: the fields of the reference's provenance record
(pbhgr/pbhgr/provenance.py:48-76: git commit, compiler and flags, host, CPU, MPI, parameters, seed,
start/end time, wall time, peak RSS, solver-supplied extras), with the GPU side filled in (device,
architecture, HIP compiler and the library's build flags, HBM held by the run).

    rec = RunRecord(out_dir, params, seed, name="mu0.1_..._provenance.json")
    ...
    rec.finish(pm, extra={"collapse": ..., "steps": ...})
"""
from __future__ import annotations

import json
import os
import platform
import resource
import subprocess
import sys
import time
from datetime import datetime, timezone
from pathlib import Path

import numpy as np

from . import build


def _run(cmd):
    try:
        return subprocess.check_output(cmd, stderr=subprocess.STDOUT, text=True, timeout=10).strip()
    except Exception as e:  # noqa: BLE001
        return f"unavailable ({e.__class__.__name__})"


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _git():
    root = Path(__file__).resolve().parent.parent
    out = _run(["git", "-C", str(root), "rev-parse", "HEAD"])
    dirty = _run(["git", "-C", str(root), "status", "--porcelain"])
    return {"commit": out, "dirty": bool(dirty) and not dirty.startswith("unavailable")}


def _default(o):
    if isinstance(o, (np.floating, np.integer)):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError(f"not JSON serialisable: {type(o)}")


class RunRecord:
    def __init__(self, run_dir, params: dict, seed, name: str = "provenance.json"):
        self.path = Path(run_dir) / name
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.t0 = time.time()
        hipcc = _run([build.HIPCC, "--version"]).splitlines()
        self.record = {
            "git": _git(),
            "python": sys.version,
            "numpy": np.__version__,
            "compiler": next((ln for ln in hipcc if "clang version" in ln), hipcc[0] if hipcc else "unavailable"),
            "compile_flags": " ".join(build.FLAGS) + " (pbh_amd/csrc/pm3d.hip)",
            "host": platform.node(),
            "cpu": _cpu_model(),
            "n_cpu": os.cpu_count(),
            "mpi": "none (one process per GPU; single-GPU run)",
            "params": params,
            "seed": seed,
            "start_utc": datetime.now(timezone.utc).isoformat(),
        }

    def finish(self, pm=None, extra: dict | None = None) -> dict:
        self.record["end_utc"] = datetime.now(timezone.utc).isoformat()
        self.record["wall_time_s"] = time.time() - self.t0
        self.record["peak_rss_MB"] = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024.0
        if pm is not None:
            self.record["gpu"] = {**pm.device_info(), "hbm_held_bytes": pm.device_bytes(),
                                  "precision": pm.precision, "fft_tiles": pm.fft_tiles(),
                                  "deposit_tiles": list(pm.deposit_tiles())}
        if extra:
            self.record.update(extra)
        with open(self.path, "w") as fh:
            json.dump(self.record, fh, indent=2, default=_default)
        return self.record
