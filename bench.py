"""Benchmark of the film step on one MI355X.

Workload: the plain single-canvas film, 2048x2048 cells, FP32, initial height
a sum of 64 seeded Gaussian blobs, per-cell yield stress and viscosity,
Dirichlet(0.3) pigments, 15 deg tilt (paint-sim-3d_amd/scenes.py).  A step is
one frame of the reference's encode_sim/encode_frame (src/main.rs:1875-1909,
src/film/mod.rs:1322-1462) at frame_dt = 1/60 s with the substep count chosen
from the previous frame's stats.  value = canvas cells x frames / wall time
of the timed frames (state resident in HBM).

roofline: the kernel family with the largest device time; achieved =
algorithmic bytes (per active cell, below, times the active cells each launch
processes) / HIP-event time of its launches in the timed region.
cpu_baseline: the CPU restatement (oracle/, one core) on a 512x512 window of
the same canvas, second frame, cells / second.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "paint-sim-3d_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

# Compulsory HBM bytes per active cell of one launch of each kernel family.
ALGO_BYTES = {
    "lap_plug": 4 + 4 + 4 + 8 + 8 + 4 + 8 + 8 + 4,  # R h m delta fnorm rheo, W delta fnorm pres tmp
    "flux": 4 + 8 + 8 + 4 + 4 + 8 + 16,             # R h rheo pres tmp delta fnorm, W flux
    "apply": 16 + 4 + 16 + 4 + 16,                  # R flux h acc, W h acc
}
SUBSTEP_FAMILIES = tuple(ALGO_BYTES)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(state_fn, settings_fn):
    import numpy as np
    import canvaz_film as czf
    from helpers import make_pair
    win = 512
    full = state_fn()
    st = {k: np.ascontiguousarray(v[:win, :win]) for k, v in full.items()}
    s = settings_fn(win)
    _, o = make_pair(win, win, s, st, pigtab=czf.default_pigment_table(), cmf=czf.cmf_table(), gpu=False)
    o.step()
    t0 = time.perf_counter()
    o.step()
    dt = time.perf_counter() - t0
    return {"value": win * win / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"CPU restatement, 512x512 window of the 2048^2 blob canvas, frame 2 "
                      f"({o.active_count()} active tiles, {dt:.1f} s)"}


def bench_km(args):
    """Spectral Kubelka-Munk mix: 4096x4096 cells x 12 pigments x 36 bands,
    BF16 concentrations ~ Dirichlet(0.3), FP32 K and S tables log-uniform in
    [0.01, 10]; value = cells / s; the kernel reads 24 B and writes 144 B per
    cell and does 2 x 12 x 72 x 2 MFMA flops per cell (hi and lo parts)."""
    import numpy as np
    import torch
    import km
    cells, bands = 4096 * 4096, 36
    rng = np.random.default_rng(36)
    log("bench km: generating inputs")
    conc = torch.from_numpy(rng.dirichlet(np.full(12, 0.3), cells).astype(np.float32)).to(torch.bfloat16).cuda()
    kt = np.exp(rng.uniform(np.log(0.01), np.log(10), (12, bands))).astype(np.float32)
    st = np.exp(rng.uniform(np.log(0.01), np.log(10), (12, bands))).astype(np.float32)
    ops = torch.from_numpy(km.prepare(kt, st).view(np.int16)).cuda()
    out = torch.empty((cells, bands), dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        km.mix(conc, ops, bands, out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        km.mix(conc, ops, bands, out)
    ev1.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    avg_us = ev0.elapsed_time(ev1) * 1e3 / args.steps
    bytes_per = (24 + 4 * bands) * cells
    achieved = bytes_per / (avg_us * 1e-6) / 1e9
    flops = 2 * 12 * 2 * (2 * bands) * cells
    # check a sample against the FP64 restatement
    import ctypes as C
    import oracle as orc
    idx = rng.integers(0, cells, 4096)
    cs = conc[idx].cpu().view(torch.int16).numpy().view(np.uint16).copy()
    ref = np.zeros((len(idx), bands), np.float64)
    orc.lib().of_km_mix_f64(cs.ctypes.data_as(C.POINTER(C.c_uint16)), kt.ctypes.data_as(C.POINTER(C.c_float)),
                            st.ctypes.data_as(C.POINTER(C.c_float)), len(idx), bands,
                            ref.ctypes.data_as(C.POINTER(C.c_double)))
    err = float(np.abs(out[idx].cpu().numpy() - ref).max())
    line = {"metric": "km mix cells/s", "value": cells * args.steps / elapsed, "unit": "cells/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic",
            "config": {"workload": "KM mix 4096x4096 cells x 12 pigments x 36 bands", "max_abs_err_vs_fp64": err},
            "roofline": {"bound": "hbm", "kernel": "k_km_mix", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "avg_launch_us": avg_us, "mfma_tflops": flops / (avg_us * 1e-6) / 1e12},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--workload", default="film", choices=("film", "km"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (canvases are independent "
              f"and would be placed one GPU each; nothing here runs on more than one GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    import canvaz_film as czf
    import scenes
    if args.workload == "km":
        return bench_km(args)

    n = args.size
    log(f"bench: building the {n}x{n} blob canvas")
    state_fn = lambda: scenes.blobs_state(n, n, seed=2048)  # noqa: E731
    settings_fn = lambda w: scenes.blobs_settings(w, tilt_deg=15.0)  # noqa: E731
    film = czf.Film(n, n, settings_fn(n))
    czf.check(czf.lib().cz_film_use_graphs(film._p, 0 if args.no_graphs else 1))
    assert film.write_state_arrays(state_fn())
    log(f"bench: warmup {args.warmup} frames")
    for _ in range(args.warmup):
        film.step()
    film.sync()
    # pick the dominant family on one profiled frame
    film.profile(True)
    film.step()
    film.sync()
    fam_time = {}
    breakdown = {}
    for fam in SUBSTEP_FAMILIES + ("lap", "split", "reduce_pool", "strata_limit", "strata_update", "copy_back",
                                   "strata_fill", "cure", "mark", "compact", "deposit", "delta_sync",
                                   "frame_stats", "cg_setup", "cg_lap", "cg_init", "cg_iter_a", "cg_iter_b",
                                   "cg_iter_c", "cg_reduce", "cg_finish", "cg_rescale"):
        avg, cnt = film.profile_read(fam)
        if cnt:
            breakdown[fam] = (avg, cnt)
        if fam in SUBSTEP_FAMILIES:
            fam_time[fam] = avg * cnt
    log("bench: one-frame breakdown (us avg x launches): " +
        ", ".join(f"{k} {v[0]:.1f}x{v[1]}" for k, v in sorted(breakdown.items(), key=lambda kv: -kv[1][0] * kv[1][1])))
    dom = max(fam_time, key=fam_time.get)
    film.profile(False)
    film.reset_counters()
    log(f"bench: timing {args.steps} frames (dominant kernel {dom})")
    torch.cuda.synchronize()
    film.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        film.step()
    film.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = film.stats()
    frames = st.frames_counted
    # event timing of the dominant kernel over a second pass of the same length
    film.profile(True)
    film.profile_only(dom)
    film.reset_counters()
    for k in range(args.steps):
        film.step()
    film.sync()
    st2 = film.stats()
    avg_us, launches = film.profile_read(dom)
    film.profile(False)
    cells_per_launch = st2.act_cell_substeps / max(1, launches)
    bytes_per_launch = ALGO_BYTES[dom] * cells_per_launch
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0
    ms_per_step = elapsed * 1e3 / args.steps
    value = n * n * args.steps / elapsed
    log(f"bench: {ms_per_step:.3f} ms/frame, substeps {st.substeps_last}, active cells/frame "
        f"{st.act_cells / max(1, frames):.0f}, {dom}: {avg_us:.2f} us x {launches}; timed pass "
        f"{st.act_cells}/{st.act_cell_substeps}/{st.frames_counted}, profiled pass "
        f"{st2.act_cells}/{st2.act_cell_substeps}/{st2.frames_counted}")
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU restatement baseline")
        cpu = cpu_baseline(state_fn, settings_fn)
    line = {
        "metric": "film cell-steps/s", "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain film {n}x{n} FP32, 64 seeded Gaussian blobs, tilt 15 deg, frame_dt 1/60",
                   "cells": n * n, "substeps_last": st.substeps_last,
                   "active_cells_per_frame": st.act_cells / max(1, frames),
                   "implicit_steps": st.implicit_steps},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algo_bytes_per_cell": ALGO_BYTES[dom], "avg_launch_us": avg_us,
                     "cells_per_launch": cells_per_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
