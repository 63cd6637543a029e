"""Run record of a pm3d run on the GPU.

The record keeps the top-level field names of the reference's provenance file
(pbhgr/pbhgr/provenance.py:52-75: git, python, numpy, compiler, compile_flags, host, cpu, n_cpu,
mpi, params, seed, start_utc, end_utc, wall_time_s, peak_rss_MB, plus solver extras), so tools that
read reference runs read these too.  What a GPU run adds is grouped under "gpu": the device, its
architecture and memory, the ROCm release, the HBM the run held, the precision mode and the tile
shapes the library chose at creation.

    rec = RunRecord(out_dir, params, seed, name="mu0.1_..._provenance.json")
    ...
    rec.finish(pm, extra={"collapse": ..., "steps": ...})
"""
from __future__ import annotations

import json
import os
import platform
import re
import resource
import shutil
import subprocess
import sys
import time
from datetime import datetime, timezone
from pathlib import Path

import numpy as np

from . import build

REPO = Path(__file__).resolve().parent.parent


def _utc_now() -> str:
    return datetime.now(timezone.utc).isoformat()


def _read(path: Path) -> str | None:
    try:
        return path.read_text().strip()
    except OSError:
        return None


def git_state(root: Path = REPO) -> dict:
    """Commit of the working tree, read from .git directly (HEAD -> ref file or packed-refs); the
    dirty flag asks git itself and is None when git cannot be run."""
    gdir = root / ".git"
    head = _read(gdir / "HEAD")
    commit = None
    if head and head.startswith("ref:"):
        ref = head.split(None, 1)[1]
        commit = _read(gdir / ref)
        if commit is None:
            packed = _read(gdir / "packed-refs") or ""
            hit = [ln.split()[0] for ln in packed.splitlines() if ln.endswith(" " + ref)]
            commit = hit[0] if hit else None
    elif head:
        commit = head
    dirty = None
    if shutil.which("git") and gdir.exists():
        try:
            res = subprocess.run(["git", "-C", str(root), "status", "--porcelain", "--untracked-files=no"],
                                 capture_output=True, text=True, timeout=10)
            dirty = bool(res.stdout.strip()) if res.returncode == 0 else None
        except (OSError, subprocess.SubprocessError):
            dirty = None
    return {"commit": commit or "unknown", "dirty": dirty}


def toolchain() -> dict:
    """hipcc's clang version line and the ROCm release of the install the library was built with."""
    rocm = _read(Path(build.HIPCC).resolve().parent.parent / ".info" / "version")
    clang = None
    try:
        res = subprocess.run([build.HIPCC, "--version"], capture_output=True, text=True, timeout=30)
        clang = next((ln.strip() for ln in res.stdout.splitlines() if "clang version" in ln), None)
    except (OSError, subprocess.SubprocessError):
        pass
    return {"compiler": clang or "hipcc (version unavailable)", "rocm": rocm or "unknown"}


def cpu_name() -> str:
    text = _read(Path("/proc/cpuinfo")) or ""
    m = re.search(r"^model name\s*:\s*(.+)$", text, re.MULTILINE)
    return m.group(1).strip() if m else (platform.processor() or platform.machine())


def _jsonable(o):
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    if isinstance(o, Path):
        return str(o)
    raise TypeError(f"{type(o).__name__} is not JSON serialisable")


class RunRecord:
    """Collects the record at construction (code, toolchain, host, parameters) and completes it in
    `finish` (times, memory, the GPU side from a PM3D handle, solver extras)."""

    def __init__(self, run_dir, params: dict, seed, name: str = "provenance.json"):
        self.path = Path(run_dir) / name
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self._t_start = time.time()
        tools = toolchain()
        self.record = {
            "git": git_state(),
            "python": sys.version,
            "numpy": np.__version__,
            "compiler": tools["compiler"],
            "compile_flags": f"{' '.join(build.FLAGS)} ({build.SRC.relative_to(REPO)})",
            "host": platform.node(),
            "cpu": cpu_name(),
            "n_cpu": os.cpu_count(),
            "mpi": "none: one GPU per process, no collectives",
            "params": dict(params),
            "seed": seed,
            "start_utc": _utc_now(),
            "rocm": tools["rocm"],
        }

    def finish(self, pm=None, extra: dict | None = None) -> dict:
        rec = self.record
        rec["end_utc"] = _utc_now()
        rec["wall_time_s"] = time.time() - self._t_start
        rec["peak_rss_MB"] = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024.0  # Linux: KiB
        if pm is not None:
            info = pm.device_info()
            rec["gpu"] = {
                "name": info["name"], "arch": info["arch"], "hbm_total_bytes": info["total_bytes"],
                "hbm_held_bytes": pm.device_bytes(), "precision": pm.precision,
                "fft_tiles": pm.fft_tiles(), "deposit_tiles": list(pm.deposit_tiles()),
            }
        rec.update(extra or {})
        self.path.write_text(json.dumps(rec, indent=2, default=_jsonable))
        return rec
