"""Per-stage and per-kernel timing of one build+solve (used by bench.py).

Kernel timings come from the library's own HIP-event instrumentation
(``hps_prof_*``): every launch is bracketed by events on its stream and
carries its algorithmic flop/byte count, so the roofline fraction of the
dominant kernel is measured live, not estimated.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib

# MI355X published dense peaks
PEAK_TFLOPS = {"f64": 78.6, "c128": 78.6, "f32": 157.3, "c64": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0


_stage_events = None


class stage:
    """Context manager recording a named stage with CUDA events while
    instrumented_step is active (a no-op otherwise)."""

    def __init__(self, name):
        self.name = name

    def __enter__(self):
        if _stage_events is not None:
            self.e0 = torch.cuda.Event(enable_timing=True)
            self.e0.record()
        return self

    def __exit__(self, *exc):
        if _stage_events is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            _stage_events.append((self.name, self.e0, e1))
        return False


def kernel_table():
    lib = _lib.load_library()
    out = []
    for i in range(lib.hps_prof_count()):
        name = ctypes.create_string_buffer(128)
        launches = ctypes.c_long()
        ms = ctypes.c_double()
        fl = ctypes.c_double()
        by = ctypes.c_double()
        lib.hps_prof_query(i, name, 128, ctypes.byref(launches), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by))
        out.append(
            dict(name=name.value.decode(), launches=launches.value, ms=ms.value, flops=fl.value, bytes=by.value)
        )
    return sorted(out, key=lambda r: -r["ms"])


def roofline_of(rec):
    """Roofline entry of one kernel class: flop-bound if it does arithmetic."""
    name = rec["name"]
    dt = name[name.find("<") + 1 : name.find(">")] if "<" in name else "f64"
    per_launch_ms = rec["ms"] / max(rec["launches"], 1)
    if rec["flops"] > 0:
        achieved = rec["flops"] / max(rec["launches"], 1) / (per_launch_ms * 1e-3) / 1e12
        peak = PEAK_TFLOPS.get(dt, 78.6)
        return dict(bound="mfma", achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak, traffic=None,
                    kernel=name, launches=rec["launches"], avg_launch_ms=per_launch_ms,
                    algorithmic_flops_per_launch=rec["flops"] / max(rec["launches"], 1))
    achieved = rec["bytes"] / max(rec["launches"], 1) / (per_launch_ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s", frac=achieved / PEAK_HBM_GBS,
                traffic=None, kernel=name, launches=rec["launches"], avg_launch_ms=per_launch_ms,
                algorithmic_bytes_per_launch=rec["bytes"] / max(rec["launches"], 1))


def instrumented_step(pde, g_dev, build=None, solve=None):
    """Run one build+solve with stage events and kernel instrumentation."""
    import jaxhps_amd as jh

    build = build or jh.build_solver
    solve = solve or jh.solve
    global _stage_events
    lib = _lib.load_library()
    torch.cuda.synchronize()
    lib.hps_prof_reset()
    lib.hps_prof_enable(1)
    _stage_events = []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    try:
        ev[0].record()
        build(pde)
        ev[1].record()
        solve(pde, g_dev)
        ev[2].record()
        torch.cuda.synchronize()
    finally:
        lib.hps_prof_enable(0)
        recorded, _stage_events = _stage_events, None
    table = kernel_table()
    lib.hps_prof_reset()
    stages = {"build_ms": ev[0].elapsed_time(ev[1]), "solve_ms": ev[1].elapsed_time(ev[2])}
    for name, e0, e1 in recorded:
        stages[name] = stages.get(name, 0.0) + e0.elapsed_time(e1)
    dominant = table[0] if table else None
    return {
        "stages_ms": stages,
        "roofline": roofline_of(dominant) if dominant else None,
        "kernels": [
            dict(name=r["name"], launches=r["launches"], ms=round(r["ms"], 3),
                 tflops=(r["flops"] / (r["ms"] * 1e-3) / 1e12) if r["ms"] > 0 and r["flops"] else None,
                 gbs=(r["bytes"] / (r["ms"] * 1e-3) / 1e9) if r["ms"] > 0 else None)
            for r in table[:24]
        ],
    }


def instrumented(fn):
    """Run fn() once with per-kernel instrumentation.  Returns (result,
    kernels, roofline of the dominant kernel)."""
    lib = _lib.load_library()
    torch.cuda.synchronize()
    lib.hps_prof_reset()
    lib.hps_prof_enable(1)
    try:
        out = fn()
        torch.cuda.synchronize()
    finally:
        lib.hps_prof_enable(0)
    table = kernel_table()
    lib.hps_prof_reset()
    kernels = [
        dict(name=r["name"], launches=r["launches"], ms=round(r["ms"], 3),
             tflops=(r["flops"] / (r["ms"] * 1e-3) / 1e12) if r["ms"] > 0 and r["flops"] else None,
             gbs=(r["bytes"] / (r["ms"] * 1e-3) / 1e9) if r["ms"] > 0 else None)
        for r in table[:12]
    ]
    return out, kernels, (roofline_of(table[0]) if table else None)
