"""Benchmark of the film step on one MI355X.

Workload: the plain single-canvas film, 2048x2048 cells, FP32, initial height
a sum of 64 seeded Gaussian blobs, per-cell yield stress and viscosity,
Dirichlet(0.3) pigments, 15 deg tilt (paint-sim-3d_amd/scenes.py), and a
power-law index per wet cell uniform in [0.4, 1] (settings.n_per_cell; the
batched and served canvases are seeded the same way); records beside it at
one film index, the reference's n = 1 and the shear-thinning n = 0.7
(--index 1 or 0.7 puts every workload on a uniform index).  A step is one frame of the reference's
encode_sim/encode_frame (src/main.rs:1875-1909, src/film/mod.rs:1322-1462) at
frame_dt = 1/60 s with the substep count chosen from the previous frame's
stats; the display shade pass (shade_full, src/film/mod.rs:1449-1456) is off,
as its output only feeds the renderer.  value = canvas cells x frames / wall
time of the timed frames (state resident in HBM).

Phases, each on a fresh film from the same initial state: W warmup frames
and the K timed frames as the library runs them by default (substep shapes
tuned online by the device's tuner, which the warmup frames train, graph
replay of small frames, no events), and the same K frames replayed with
events around every launch (and the tuner's shapes) for the roofline,
whose unit is the substep (the dominant work: about 85 % of the frame's
device time), whatever launches the tuner split it into:
achieved = 76 B (+ 4 with an index per cell) per active cell and substep /
the substep's device time (its launches' event times summed).  Between the
phases a one-element fill kernel marks the boundary in a kernel trace.
roofline.traffic: HBM bytes per substep: the HBM bytes per active cell and
substep of each substep shape, from rocprof FETCH_SIZE / WRITE_SIZE runs of
the same command with that shape pinned (recorded under profiles/),
weighted by the cell-substeps this run's replay gave each shape, times its
cells per substep (null when the records cover less than 99 % of them).  cpu_baseline: the CPU
restatement (oracle/, its per-tile passes on every host thread) stepping the
same canvas from the same initial state, frame by frame, for up to 25 s.
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
    "lap_plug": 4 + 4 + 4 + 8 + 8 + 4 + 4 + 8 + 8 + 4 + 4,  # R h m delta fnorm rheo lmu, W delta fnorm pres tmp h_sub
    "flux": 4 + 8 + 4 + 8 + 4 + 4 + 8 + 16,         # R h rheo lmu pres tmp delta fnorm, W flux
    "apply": 16 + 4 + 16 + 4 + 16,                  # R flux h acc, W h acc
}
SUBSTEP_FAMILIES = tuple(ALGO_BYTES)
SUBSTEP_BYTES = sum(ALGO_BYTES.values())            # per active cell and substep
ALGO_BYTES["lap"] = 4 + 4 + 4 + 8 + 4 + 8 + 8       # last nose: R h m delta fnorm, W delta fnorm pres
# flux + apply in one launch: the flux pass's reads and the apply's, the
# tile's fluxes written once and never read back
ALGO_BYTES["flux_apply"] = (4 + 8 + 4 + 8 + 4 + 4 + 8) + (4 + 16) + (16 + 4 + 16)
# the 1024-thread shapes of the same passes move the same bytes
for _fam in ("lap_plug", "flux", "apply", "flux_apply"):
    ALGO_BYTES[_fam + "_w"] = ALGO_BYTES[_fam]
# the whole substep in one launch: R h m delta rheo lmu fnorm, RW acc, W h delta fnorm tmp
ALGO_BYTES["substep_fused"] = (4 + 4 + 4 + 8 + 4 + 8) + (16 + 16) + (4 + 4 + 8 + 4)
# with an index per cell: the plug window and the face pass read it
CELL_N_BYTES = {"lap_plug": 4, "flux": 4, "flux_apply": 4, "lap_plug_w": 4, "flux_w": 4, "flux_apply_w": 4,
                "substep_fused": 4}


def algo_bytes(fam, cell_n):
    return ALGO_BYTES[fam] + (CELL_N_BYTES.get(fam, 0) if cell_n else 0)


# the canvases of every workload: an index per wet cell, uniform in [0.4, 1]
# (slot 15 of its stratum), or the reference's single film index n = 1
INDEX_RANGE = (0.4, 1.0)


def canvas_settings(w, tilt_deg, cell_n):
    import scenes
    s = scenes.blobs_settings(w, tilt_deg=tilt_deg)
    s.n_per_cell = 1 if cell_n else 0
    return s


def index_words(index):
    if index == "cell":
        return f"power-law index per wet cell U[{INDEX_RANGE[0]}, {INDEX_RANGE[1]}]"
    return f"uniform power-law index n = {index}"


def canvas_scene(seed, cell_n):
    import canvaz_film as czf
    if cell_n:
        return czf.scene(seed=seed, n_lo=INDEX_RANGE[0], n_hi=INDEX_RANGE[1])
    return czf.scene(seed=seed)
# families the roofline is reported for: the substep passes, each over both
# block sizes (the 1024-thread shapes move the same bytes), the whole
# substep in one launch (k_substep_fused) and all substeps in one launch
# (k_substeps); profiled launch families of each
ROOFLINE_FAMILIES = {"lap_plug": ("lap_plug", "lap_plug_w"), "flux": ("flux", "flux_w"), "apply": ("apply", "apply_w"),
                     "flux_apply": ("flux_apply", "flux_apply_w"), "substep_fused": ("substep_fused",),
                     "substeps": ("substeps",)}
SUBSTEP_LAUNCH_FAMILIES = tuple(f for fams in ROOFLINE_FAMILIES.values() for f in fams)
# Compulsory bytes of one substep per active cell (a fused explicit substep:
# R h, m, rheo, delta, fnorm, acc; W h, delta, fnorm, acc), + 4 with an index
# per cell (R n)
SUBSTEP_ALGO_BYTES = 76


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(state, settings, frames, budget_s=25.0):
    """The CPU restatement on all host threads (its per-tile passes run in
    parallel; the FP64 sums stay in a fixed order) stepping the same canvas
    from the same initial state for the same frames, stopping after the first
    frame that ends past budget_s seconds: canvas cells x frames / time."""
    import canvaz_film as czf
    import oracle as orc
    from helpers import make_pair
    n = state["h"].shape[1]
    _, o = make_pair(n, state["h"].shape[0], settings, state, pigtab=czf.default_pigment_table(),
                     cmf=czf.cmf_table(), gpu=False)
    done, dt = 0, 0.0
    while done < frames and dt < budget_s:
        t0 = time.perf_counter()
        o.step()
        dt += time.perf_counter() - t0
        done += 1
    cells = state["h"].size
    threads = int(orc.lib().of_omp_threads())
    return {"value": cells * done / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU restatement (oracle/, not the reference), {threads} threads, the same "
                      f"{state['h'].shape[1]}x{state['h'].shape[0]} canvas and settings, frames 1-{done} of the "
                      f"timed {frames} from the same initial state ({dt:.1f} s)"}


def km_record(steps, warmup, cells=4096 * 4096, bands=36):
    """Spectral Kubelka-Munk mix: 4096x4096 cells x 12 pigments x 36 bands,
    BF16 concentrations ~ Dirichlet(0.3) (drawn on the device), FP32 K and S
    tables log-uniform in [0.01, 10]; the kernel reads 24 B and writes 4 B
    per band per cell and does 2 x 12 x (2 x bands) x 2 MFMA flops per cell
    (K and S, hi and lo table parts).  Average launch time from HIP events on
    torch's current stream (the kernel's stream); error against the FP64
    restatement on 4096 sampled cells."""
    import ctypes as C
    import numpy as np
    import torch
    import km
    import oracle as orc
    rng = np.random.default_rng(36)
    g = torch.Generator(device="cuda").manual_seed(36)
    gam = torch._standard_gamma(torch.full((cells, 12), 0.3, device="cuda"), generator=g)
    conc = (gam / gam.sum(dim=1, keepdim=True)).to(torch.bfloat16).contiguous()
    del gam
    kt = np.exp(rng.uniform(np.log(0.01), np.log(10), (12, bands))).astype(np.float32)
    st = np.exp(rng.uniform(np.log(0.01), np.log(10), (12, bands))).astype(np.float32)
    ops = torch.from_numpy(km.prepare(kt, st).view(np.int16)).cuda()
    out = torch.empty((cells, bands), dtype=torch.float32, device="cuda")
    for _ in range(warmup):
        km.mix(conc, ops, bands, out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        km.mix(conc, ops, bands, out)
    ev1.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    avg_us = ev0.elapsed_time(ev1) * 1e3 / steps
    bytes_per = (24 + 4 * bands) * cells
    achieved = bytes_per / (avg_us * 1e-6) / 1e9
    flops = 2 * 12 * 2 * (2 * bands) * cells
    idx = torch.from_numpy(rng.integers(0, cells, 4096)).cuda()
    cs = conc[idx].cpu().view(torch.int16).numpy().view(np.uint16).copy()
    ref = np.zeros((len(idx), bands), np.float64)
    orc.lib().of_km_mix_f64(cs.ctypes.data_as(C.POINTER(C.c_uint16)), kt.ctypes.data_as(C.POINTER(C.c_float)),
                            st.ctypes.data_as(C.POINTER(C.c_float)), len(idx), bands,
                            ref.ctypes.data_as(C.POINTER(C.c_double)))
    err = float(np.abs(out[idx].cpu().numpy() - ref).max())
    return {"value": cells * steps / elapsed, "unit": "cells/s", "ms_per_step": elapsed * 1e3 / steps,
            "steps": steps, "dtype": "bf16", "max_abs_err_vs_fp64": err,
            "workload": f"KM mix {int(round(cells ** 0.5))}x{int(round(cells ** 0.5))} cells x 12 pigments x "
                        f"{bands} bands, BF16 concentrations, FP32 accumulate",
            "roofline": {"bound": "hbm", "kernel": "k_km_mix", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": recorded_traffic("km_mix_4096sq_36band_kernel_stats.md", kernel="k_km_mix",
                                                     size=cells, bands=bands),
                         "avg_launch_us": avg_us, "mfma_tflops": flops / (avg_us * 1e-6) / 1e12,
                         "mfma_peak_tflops": 2500.0}}


def bench_km(args):
    """The KM mix alone (--workload km): one JSON line."""
    r = km_record(args.steps, args.warmup)
    line = {"metric": "km mix cells/s", "value": r["value"], "unit": "cells/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic",
            "config": {"workload": r["workload"], "max_abs_err_vs_fp64": r["max_abs_err_vs_fp64"]},
            "roofline": r["roofline"], "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def bench_next(args):
    """The passes beside the film step, each on the 2048^2 blob canvas (or
    the empty canvas where the pass wants bare paint), device time from HIP
    events around its launches (average over the frames it ran in) and its
    compulsory HBM bytes:
    - deposit binning on the device (k_dep_rects/scan/fill) for 16384
      capsules in one frame, beside the host binning's wall time;
    - the undo journal saving every tile of the canvas in one frame (53
      words read and written per cell) and the restore of all of them;
    - the shade pass over every cell (212 B read: the state parts it uses,
      16 B written);
    - the brush scrape of 4096 bristle contacts and the knife across the
      canvas (their cells read, the changed cells' 212 B state read and
      written);
    - the .canvaz save and open of the canvas (host LZ4; wall time).
    One JSON line with a record per pass."""
    import math
    import tempfile
    import numpy as np
    import canvaz_film as czf
    import cz_types
    import scenes
    n = 2048
    state = scenes.blobs_state(n, n, seed=2048)
    s = scenes.blobs_settings(n, tilt_deg=15.0)
    dx = s.width_m / n
    pig = scenes.jet_paint()
    rng = np.random.default_rng(77)
    out = {}

    def film_with(st=True):
        f = czf.Film(n, n, s)
        if st:
            assert f.write_state_arrays(state)
        return f

    def rec(name, label_us, launches, bytes_per_frame, note):
        ach = bytes_per_frame / (label_us * 1e-6) / 1e9 if label_us > 0 else 0.0
        out[name] = {"device_us_per_frame": label_us, "launch_groups": launches, "algo_bytes": bytes_per_frame,
                     "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS, "note": note}
        log(f"bench next: {name}: {label_us:.1f} us/frame, {bytes_per_frame / 1e6:.1f} MB, {ach:.0f} GB/s")

    # deposit binning
    deps = []
    for _ in range(16384):
        x, y = rng.uniform(0.05, 0.95, 2) * n * dx
        a = rng.uniform(0, 2 * math.pi)
        ln = rng.uniform(1, 6) * dx
        deps.append(czf.deposit_capsule((x, y), (x + ln * math.cos(a), y + ln * math.sin(a)), 2.5 * dx, 2e-10,
                                        pig, dx))
    f = film_with()
    for mode, name in ((0, "binning_host"), (1, "binning_device")):
        f.bin_on_device(mode)
        f.profile(True)
        host_s = 0.0
        frames = 6
        for _ in range(frames):
            t0 = time.perf_counter()
            f.upload_deposits(deps)
            host_s += time.perf_counter() - t0
            f.encode_frame(4)
            f.sync()
        us, cnt = f.profile_read("dep_bin")
        dep_us, _ = f.profile_read("deposit")
        if mode == 1:
            # three launches; 96 B per deposit read twice (rectangles, fill),
            # 16 B rectangle written and read, 4 B per item written
            nb = len(deps) * (96 * 2 + 16 * 2) + 4 * 16 * len(deps)
            rec(name, us * cnt / frames, cnt // frames, nb,
                f"{len(deps)} capsules in one frame; host upload_deposits {host_s / frames * 1e3:.2f} ms/frame, "
                f"k_deposit {dep_us:.0f} us")
        else:
            out[name] = {"host_ms_per_frame": host_s / frames * 1e3, "k_deposit_us": dep_us,
                         "note": f"{len(deps)} capsules binned on the host (upload_deposits wall time)"}
            log(f"bench next: {name}: {host_s / frames * 1e3:.2f} ms/frame on the host")
        f.profile(False)
    f.close()

    # journal: every tile saved in one frame, then restored
    f = film_with()
    f.journal_budget(20000 * 53 * 256 * 4)
    f.journal_gen(1)  # a first journal frame and restore: the kernels' first launches out of the timing
    f.step()
    f.restore_journal(1)
    f.sync()
    head0, _ = f.journal_head()
    f.journal_gen(2)
    f.profile(True)
    f.touch_rect(-1.0, -1.0, 10.0, 10.0)
    f.step()
    f.sync()
    us, cnt = f.profile_read("journal")
    head, cap = f.journal_head()
    saved = head - head0
    rec("journal_save", us, cnt, saved * 256 * 53 * 4 * 2,
        f"{saved} tiles saved in one frame (ring of {cap})")
    f.profile(False)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f.restore_journal(2)
    f.sync()
    dt = time.perf_counter() - t0
    rec("journal_restore", dt * 1e6, 3, saved * 256 * 53 * 4 * 2,
        "wall time of cz_film_restore_journal + sync (three launches)")
    f.close()

    # shade over every cell
    f = film_with()
    f.shade_output(True)
    f.profile(True)
    for _ in range(5):
        f.step()
    f.sync()
    us, cnt = f.profile_read("shade")
    rec("shade", us, cnt, n * n * (4 + 4 + 8 + 16 + 128 + 16 + 32 + 4 + 16), "2048^2 cells, every cell shaded")
    f.close()

    # brush: 4096 bristle contacts over a 512-cell square of the painted canvas
    f = film_with()
    cs = []
    for b in range(4096):
        c = cz_types.Contact()
        x, y = (rng.uniform(768, 1280, 2)) * dx
        c.p0[:] = (x, y)
        c.p1[:] = (x + 3 * dx, y)
        c.r, c.zb, c.bristle, c.pad = 1.5 * dx, 2e-5, b % 512, 0
        cs.append(c)
    f.upload_contacts(cs[:1])  # the first launch out of the timing
    f.step()
    f.sync()
    f.profile(True)
    frames = 5
    for _ in range(frames):
        f.upload_contacts(cs)
        f.step()
    f.sync()
    us, cnt = f.profile_read("brush")
    visited = 9 * 64 * 64  # the 64-tiles of the 512-cell square (and its edge tiles)
    rec("brush", us, cnt, visited * 8 + 512 * 512 * 212 * 2 // 4,
        "4096 bristle contacts in a 512x512 square, bytes assume a quarter of its cells scraped")
    f.close()

    # knife: a 0.3 m blade swept 8 cells per frame across the canvas
    f = film_with()
    frames = 5
    for k in range(-1, frames):  # frame -1: the first launches, out of the timing
        if k == 0:
            f.sync()
            f.take_scraps()
            f.profile(True)
        kn = cz_types.Knife()
        y0 = (600 + 8 * max(k, 0)) * dx
        kn.a[:] = (300 * dx, y0)
        kn.b[:] = (300 * dx + 0.3, y0)
        kn.dir[:] = (0.0, 1.0)
        kn.sweep, kn.height, kn.half_width, kn.active, kn.frame, kn.bands = 8 * dx, 5e-5, 0.15, 1, max(k, 0), 8
        f.set_knife(kn)
        f.step()
    f.set_knife(None)
    f.sync()
    us, cnt = f.profile_read("knife")
    box_cells = int((0.3 / dx + 17) * ((0.07 + dx) / dx + 17))
    cut = int(0.3 / dx) * 8
    rec("knife", us, cnt, box_cells * 8 * 2 + cut * (212 * 2 + 96),
        f"{box_cells} cells visited per frame, about {cut} cut")
    sc = f.take_scraps()
    out["knife"]["scraps"] = int(len(sc))
    f.close()

    # .canvaz save / open (host LZ4 over read_state / write_state)
    f = film_with()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "bench.canvaz")
        t0 = time.perf_counter()
        f.save_canvaz(path)
        t_save = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        f.load_canvaz(path)
        t_load = time.perf_counter() - t0
    raw = sum(f.state_sizes())
    out["canvaz"] = {"state_bytes": raw, "file_bytes": size, "save_s": t_save, "open_s": t_load,
                     "save_MBps": raw / t_save / 1e6, "open_MBps": raw / t_load / 1e6,
                     "note": "2048^2 blob canvas; PCIe copies and host LZ4 included"}
    log(f"bench next: canvaz: {raw / 1e6:.0f} MB state, {size / 1e6:.0f} MB file, save {t_save:.2f} s, "
        f"open {t_load:.2f} s")
    f.close()
    print(json.dumps({"metric": "beside-the-step passes", "unit": "us per frame", "n_gpus": 1,
                      "config": {"workload": "2048x2048 blob canvas"}, "passes": out}), flush=True)


def batch_canvases(count=64, seed=4):
    """The batched workload: `count` square canvases, side log-uniform in
    [256, 2048], seeded blob states as the plain canvas, tilt U[0, 60] deg."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = np.round(np.exp(rng.uniform(np.log(256.0), np.log(2048.0), count))).astype(int)
    tilts = rng.uniform(0.0, 60.0, count)
    return [(int(w), float(t), 5000 + k) for k, (w, t) in enumerate(zip(sides, tilts))]


def bench_batch(args, plain_value):
    """All canvases, seeded on the device, stepped together through one
    cz_batch for `frames` frames from their initial states; then each canvas stepped alone through
    the plain path for the same frames, timed, and compared with its batched
    twin (largest relative height difference, floor 1e-9 m)."""
    import numpy as np
    import torch
    import canvaz_film as czf
    import scenes
    specs = batch_canvases(args.batch_canvases)
    frames = args.batch_frames
    cells = sum(w * w for (w, _, _) in specs)
    cell_n = args.index == "cell"
    films = []
    for (w, tilt, sd) in specs:
        f = czf.Film(w, w, canvas_settings(w, tilt, cell_n))
        f.seed_blobs(canvas_scene(sd, cell_n))
        films.append(f)
    b = czf.Batch(max(len(films), 1))
    b.use_graphs(not args.no_graphs)
    b.flux_shape(args.flux_shape if args.flux_shape < 5 else -1)
    for f in films:
        b.add(f)
    log(f"bench batch: {len(films)} canvases, {cells / 1e6:.1f} M cells, timing {frames} frames")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        b.step()
    b.sync()
    t_batch = time.perf_counter() - t0
    batched_value = cells * frames / t_batch
    subs = [f.stats().act_cell_substeps / max(1, f.stats().act_cells) for f in films]
    act = sum(f.stats().act_cells for f in films) / max(1, frames)
    log(f"bench batch: {t_batch * 1e3 / frames:.2f} ms/frame, {batched_value:.3e} cell-steps/s; "
        f"plain comparison")
    t_plain, worst = 0.0, 0.0
    for k, (w, tilt, sd) in enumerate(specs):
        g = czf.Film(w, w, canvas_settings(w, tilt, cell_n))
        g.seed_blobs(canvas_scene(sd, cell_n))
        g.sync()
        t0 = time.perf_counter()
        for _ in range(frames):
            g.step()
        g.sync()
        t_plain += time.perf_counter() - t0
        hb, hp = films[k].buffer("h").astype(np.float64), g.buffer("h").astype(np.float64)
        worst = max(worst, float(np.max(np.abs(hb - hp) / np.maximum(np.maximum(np.abs(hb), np.abs(hp)), 1e-9))))
        g.close()
    b.close()
    for f in films:
        f.close()
    plain_same = cells * frames / t_plain
    return {"value": batched_value, "unit": "cell-steps/s", "ms_per_frame": t_batch * 1e3 / frames,
            "canvases": len(specs), "cells": cells, "frames": frames,
            "vs_plain_2048": batched_value / plain_value if plain_value else None,
            "plain_same_canvases": plain_same, "vs_plain_same_canvases": batched_value / plain_same,
            "max_rel_err_vs_plain": worst, "active_cells_per_frame": act,
            "mean_substeps": float(np.mean(subs)), "graphs": not args.no_graphs,
            "workload": f"{len(specs)} square canvases, side log-uniform [256, 2048], seeded blob canvases "
                        f"(64 blobs, device generator, {index_words(args.index)}), tilt U[0, 60] deg, "
                        f"{frames} frames from the initial states"}


def serve_trace(seconds, rate=200.0, seed=5, cell_n=True):
    """Poisson arrivals at `rate` per second for `seconds`; canvas side
    log-uniform in [512, 4096]; steps Pareto(1.5, min 10) truncated at 2000;
    each canvas a seeded blob scene with the plain workload's cell size and
    a 15 deg tilt (with an index per cell when cell_n)."""
    import numpy as np
    import canvaz_film as czf
    import scenes
    rng = np.random.default_rng(seed)
    t, k, reqs = 0.0, 0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= seconds:
            break
        w = int(round(float(np.exp(rng.uniform(np.log(512.0), np.log(4096.0))))))
        steps = int(min(2000, np.floor(10.0 / rng.uniform(0.0, 1.0) ** (1.0 / 1.5))))
        reqs.append(czf.request(k, w, w, steps, t, canvas_settings(w, 15.0, cell_n), canvas_scene(7000 + k, cell_n)))
        k += 1
    return reqs


def serve_once(args, rate, max_wait):
    """One run of the serving loop on a Poisson trace at `rate` requests per
    second: latency = completion - arrival."""
    import numpy as np
    import canvaz_film as czf
    reqs = serve_trace(args.serve_seconds, rate=rate, cell_n=args.index == "cell")
    offered = sum(r.w * r.h * r.steps for r in reqs) / args.serve_seconds
    policy = "first come first served" if max_wait <= 0 else f"least work first, aged after {max_wait} s (evicting)"
    log(f"bench serve: {len(reqs)} requests at {rate} req/s over {args.serve_seconds} s, offered {offered:.3e} "
        f"cell-steps/s, {policy}")
    srv = czf.Server(hbm_budget=args.serve_budget_gb << 30, max_members=args.serve_members, max_wait_s=max_wait)
    res, st = srv.run(reqs)
    srv.close()
    ok = [r for r in res if r.status == 0]
    lat = np.array([r.done_s - r.arrival_s for r in ok])
    wait = np.array([r.admit_s - r.arrival_s for r in ok])
    cs = sum(r.cell_steps for r in ok)
    out = {"requests": len(reqs), "completed": len(ok), "failed": len(reqs) - len(ok),
           "p50_latency_s": float(np.percentile(lat, 50)) if len(ok) else None,
           "p99_latency_s": float(np.percentile(lat, 99)) if len(ok) else None,
           "mean_queue_wait_s": float(wait.mean()) if len(ok) else None,
           "throughput_cell_steps_per_s": cs / st.wall_s, "wall_s": st.wall_s,
           "offered_cell_steps_per_s": offered, "frames": st.frames, "peak_members": st.peak_members,
           "peak_gb": st.peak_bytes / 2 ** 30, "films_created": st.films_created, "films_reused": st.films_reused,
           "evictions": st.evictions, "resumes": st.resumes, "peak_host_gb": st.peak_host_bytes / 2 ** 30,
           "policy": policy,
           "workload": f"Poisson {rate} req/s for {args.serve_seconds} s, side log-uniform [512, 4096], steps "
                       f"Pareto(1.5, min 10, max 2000), seeded blob canvases ({index_words(args.index)}), "
                       f"budget {args.serve_budget_gb} GB, "
                       f"at most {args.serve_members} canvases per batch"}
    log(f"bench serve: p50 {out['p50_latency_s']:.3f} s, p99 {out['p99_latency_s']:.3f} s, "
        f"{out['throughput_cell_steps_per_s']:.3e} cell-steps/s over {st.wall_s:.1f} s, "
        f"{st.evictions} evictions")
    return out


def bench_serve(args):
    """The serving workload at its stated 200 req/s (several times what one
    GPU serves, so latency is mostly queueing) under the scheduling policy,
    and the same trace shape at a rate one GPU sustains."""
    out = serve_once(args, args.serve_rate, args.serve_max_wait)
    if args.serve_sustainable_rate > 0:
        out["sustainable"] = serve_once(args, args.serve_sustainable_rate, args.serve_max_wait)
    return out


def frame_algo_bytes(n_grid, st, cell_n=False):
    """Compulsory bytes of the frames folded in `st`:
    N_grid*20 + N_act*(632 + 76*S + I*(180 + 68*K)) per frame, K = 40 CG iterations;
    with an index per cell 4 more per grid cell (the cure writes it) and 8
    more per active cell and substep (plug window and faces read it); and
    the log viscosity per cell: 4 per grid cell (written with the rheology)
    and 8 per active cell and substep (read by the plug window and faces)."""
    extra = (4 * n_grid * st.frames_counted + 8 * st.act_cell_substeps) if cell_n else 0
    extra += 4 * n_grid * st.frames_counted + 8 * st.act_cell_substeps
    return (n_grid * 20 * st.frames_counted + 632 * st.act_cells + 76 * st.act_cell_substeps
            + (180 + 68 * 40) * st.act_cell_implicit + extra)


def stream_peak(nbytes=1 << 30, reps=20):
    """What HBM delivers to a kernel on this device: cz_stream_probe's copy
    (read + write bytes counted) and fill of nbytes, GB/s from HIP events on
    torch's current stream."""
    import torch
    import canvaz_film as czf
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    st = torch.cuda.current_stream().cuda_stream
    out = {}
    for mode, name, moved in ((0, "copy_GBps", 2 * nbytes), (1, "fill_GBps", nbytes)):
        for _ in range(3):
            czf.check(czf.lib().cz_stream_probe(b.data_ptr(), a.data_ptr(), nbytes, mode, st))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            czf.check(czf.lib().cz_stream_probe(b.data_ptr(), a.data_ptr(), nbytes, mode, st))
        e1.record()
        torch.cuda.synchronize()
        out[name] = moved / (e0.elapsed_time(e1) * 1e-3 / reps) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return out


def phase_marker(tag):
    """One tiny fill kernel between bench phases, so a kernel trace of the run
    can be cut at the phase boundaries (warmup, timed, replay, ...)."""
    import torch
    torch.cuda.synchronize()
    torch.full((1,), float(tag), device="cuda")
    torch.cuda.synchronize()


def recorded_traffic(profile, field="bytes_per_launch", **key):
    """HBM bytes recorded in profiles/<profile> from the FETCH_SIZE /
    WRITE_SIZE profiler runs of a command (lines 'traffic: kernel=K size=N
    ... bytes_per_launch=B' or 'bytes_per_cell=B'); None unless every
    key=value given here matches the record.  The film's substep launches
    cover as many cells as the shape tuner's choices give them, so their
    record is per cell (the family's bytes over its launches / the cells
    they covered), scaled by this run's cells per launch."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if not line.startswith("traffic: "):
            continue
        rec = dict(kv.split("=", 1) for kv in line.split()[1:])
        if all(rec.get(k) == str(v) for k, v in key.items()):
            return float(rec[field]) if field in rec else None
    return None


TRAFFIC_PROFILE = "film_plain_2048_hbm_traffic.md"


def substep_traffic(n, steps, warmup, shape_cells):
    """HBM bytes per active cell and substep of the events replay: the
    per-shape records (each measured with the substep shape pinned) weighted
    by the cells x substeps the replay ran in each shape.  Returns (bytes,
    covered fraction); bytes is None when the recorded shapes cover less
    than 99 % of the replay's cell-substeps."""
    total = float(sum(shape_cells))
    covered = weighted = 0.0
    for k, c in enumerate(shape_cells):
        if not c:
            continue
        b = recorded_traffic(TRAFFIC_PROFILE, "bytes_per_cell", kernel="substep", shape=k, size=n, steps=steps,
                             warmup=warmup)
        if b is not None:
            covered += c
            weighted += b * c
    frac = covered / total if total else 0.0
    return (weighted / covered if covered and frac >= 0.99 else None), frac


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="direct launches instead of captured hipGraph replays")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device to device 3 times (a known-size dispatch that "
                         "calibrates the HBM byte counters of a profiler run)")
    ap.add_argument("--flux-shape", type=int, default=-1, choices=(-1, 0, 1, 2, 3, 4, 5, 6),
                    help="substep shape: -1 tuned online, 0 window, 1 face-parallel, 2 flux + apply fused, "
                         "3 and 4 as 0 and 2 with 1024 threads per tile, 5 the whole substep in one launch, "
                         "6 the flux of 1 between the lap/plug and apply of 3 "
                         "(single canvases; the batch keeps its tuned shapes)")
    ap.add_argument("--substep-launch", default="auto", choices=("auto", "per-pass", "one"),
                    help="substeps as one launch per frame on small active sets (auto), always, or never")
    ap.add_argument("--events-in-timed", action="store_true",
                    help="record the roofline's HIP events inside the timed frames (default: a separate replay)")
    ap.add_argument("--workload", default="film", choices=("film", "km", "batch", "serve", "next", "stream"))
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload in the film run")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-budget-gb", type=int, default=200)
    ap.add_argument("--serve-members", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="arrivals per second of the serving workload")
    ap.add_argument("--serve-max-wait", type=float, default=0.0,
                    help="serving policy: 0 first come first served; > 0 least work first, requests waiting "
                         "longer than this (s) first, evicting long-running canvases to host memory")
    ap.add_argument("--serve-sustainable-rate", type=float, default=30.0,
                    help="a second serving run at this arrival rate (0: none)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload in the film run")
    ap.add_argument("--batch-canvases", type=int, default=64)
    ap.add_argument("--batch-frames", type=int, default=200)
    ap.add_argument("--no-km", action="store_true", help="skip the KM mix record in the film run")
    ap.add_argument("--index", default="cell", choices=("cell", "1", "0.7"),
                    help="power-law index of every workload's canvases: one per wet cell in [0.4, 1] "
                         "(default), or uniform")
    ap.add_argument("--no-uniform", action="store_true",
                    help="skip the uniform-index (n = 1 and 0.7) plain records in the film run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (canvases are independent "
              f"and would be placed one GPU each; nothing here runs on more than one GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    import canvaz_film as czf
    import scenes
    if args.copy_probe_mib > 0:
        nb = args.copy_probe_mib << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copy probe, 3 device copies of {args.copy_probe_mib} MiB")
        del src, dst
    if args.workload == "stream":
        print(json.dumps({"metric": "HBM stream GB/s", "unit": "GB/s", "n_gpus": 1, "peak": HBM_PEAK_GBS,
                          **stream_peak()}), flush=True)
        return
    if args.workload == "km":
        return bench_km(args)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args, None)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.workload == "next":
        return bench_next(args)

    n = args.size
    cell_n = args.index == "cell"
    log(f"bench: building the {n}x{n} blob canvas ({index_words(args.index)})")
    state = scenes.blobs_state(n, n, seed=2048, n_range=INDEX_RANGE if cell_n else None)

    def settings_fn(w):
        s = canvas_settings(w, 15.0, cell_n)
        if not cell_n:
            s.n_index = float(args.index)
        return s

    def fresh():
        f = czf.Film(n, n, settings_fn(n))
        czf.check(czf.lib().cz_film_use_graphs(f._p, 0 if args.no_graphs else 1))
        f.substep_launch({"auto": -1, "per-pass": 0, "one": 1}[args.substep_launch])
        f.flux_shape(args.flux_shape)
        assert f.write_state_arrays(state)
        return f

    # warmup film: the first frames of the same run, as the library runs them
    # (the substep tuner of the device times its launch shapes here; later
    # films on the device start from its measurements)
    phase_marker(1)
    film = fresh()
    for _ in range(max(1, args.warmup)):
        film.step()
    film.sync()
    log(f"bench: warmup {args.warmup} frames")
    film.close()
    del film

    # timed pass: a fresh film from the initial state, K frames
    phase_marker(2)
    film = fresh()
    if args.events_in_timed:
        film.profile(True)
    film.reset_counters()
    log(f"bench: timing {args.steps} frames (events in the timed frames: {args.events_in_timed})")
    torch.cuda.synchronize()
    film.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        film.step()
    film.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = film.stats()
    timed_st = st
    if not args.events_in_timed:
        # the same frames again with events around every launch
        film.close()
        del film
        phase_marker(3)
        film = fresh()
        film.profile(True)
        film.reset_counters()
        for _ in range(args.steps):
            film.step()
        film.sync()
        st = film.stats()
    prof_l = {f: film.profile_read(f) for f in SUBSTEP_LAUNCH_FAMILIES}
    film.profile(False)
    film.close()
    del film
    phase_marker(4)
    # The roofline's unit is the substep.  The tuner runs a substep as three
    # launches (lap_plug, flux, apply), two (lap_plug, flux_apply) or one
    # (the whole substep), in 256- or 1024-thread blocks, and which it picks
    # varies from run to run; a substep's device time (the sum of its
    # launches) and its compulsory bytes compare whatever the mix.
    sub_us = sum(prof_l[f][0] * prof_l[f][1] for f in SUBSTEP_LAUNCH_FAMILIES)
    one_sub = st.one_launch_cell_substeps / max(1, st.one_launch_cells)   # substeps of a one-launch frame
    n_sub = prof_l["lap_plug"][1] + prof_l["lap_plug_w"][1] + prof_l["substep_fused"][1] + \
        round(prof_l["substeps"][1] * one_sub)
    algo_per_cell = SUBSTEP_ALGO_BYTES + (4 if cell_n else 0)
    cells_per_launch = st.act_cell_substeps / max(1, n_sub)
    bytes_per_launch = algo_per_cell * cells_per_launch
    avg_us = sub_us / max(1, n_sub)
    launches = n_sub
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0
    ms_per_step = elapsed * 1e3 / args.steps
    value = n * n * args.steps / elapsed
    frames = max(1, st.frames_counted)
    frame_gbs = frame_algo_bytes(n * n, st, cell_n) / elapsed / 1e9
    log(f"bench: replay, per-family device time (us avg x launches): " +
        ", ".join(f"{k} {v[0]:.2f}x{v[1]}" for k, v in prof_l.items() if v[1]))
    log(f"bench: replay cell-substeps per substep shape {list(st.flux_shape_cell_substeps)}")
    log(f"bench: {ms_per_step:.3f} ms/frame, active cells/frame {st.act_cells / frames:.0f}, "
        f"substeps/frame {st.act_cell_substeps / max(1, st.act_cells):.1f}, substep: {avg_us:.2f} us x {launches}, "
        f"{cells_per_launch:.0f} cells/substep")
    uniform = None
    if not args.no_uniform:
        # the same canvas and frames at one film index: the reference's
        # Newtonian n = 1 and the shear-thinning n = 0.7 (the power-law
        # branches of hb_flux and band_flux, film.wgsl:94-126)
        uniform = []
        for nu in (1.0, 0.7):
            su = canvas_settings(n, 15.0, False)
            su.n_index = nu
            fu = czf.Film(n, n, su)
            czf.check(czf.lib().cz_film_use_graphs(fu._p, 0 if args.no_graphs else 1))
            assert fu.write_state_arrays(state)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fu.step()
            fu.sync()
            elu = time.perf_counter() - t0
            su_st = fu.stats()
            uniform.append({"value": n * n * args.steps / elu, "unit": "cell-steps/s",
                            "ms_per_step": elu * 1e3 / args.steps, "n_index": nu,
                            "active_cells_per_frame": su_st.act_cells / max(1, su_st.frames_counted),
                            "substeps_per_frame": su_st.act_cell_substeps / max(1, su_st.act_cells)})
            fu.close()
            log(f"bench: uniform n = {nu}: {elu * 1e3 / args.steps:.3f} ms/frame")
        phase_marker(5)
    measured = stream_peak()
    log(f"bench: measured HBM stream: copy {measured['copy_GBps']:.0f} GB/s, fill {measured['fill_GBps']:.0f} GB/s")
    phase_marker(6)
    km_rec = None
    if not args.no_km:
        km_rec = km_record(max(5, min(args.steps, 50)), 3)
        km_rec["roofline"]["measured_copy_GBps"] = measured["copy_GBps"]
        km_rec["roofline"]["frac_of_measured_copy"] = km_rec["roofline"]["achieved"] / measured["copy_GBps"]
        log(f"bench km: {km_rec['roofline']['avg_launch_us']:.1f} us, {km_rec['roofline']['achieved']:.0f} GB/s, "
            f"max err {km_rec['max_abs_err_vs_fp64']:.2e}")
        phase_marker(7)
    batched = None if args.no_batch else bench_batch(args, value)
    served = None if args.no_serve else bench_serve(args)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU restatement baseline")
        cpu = cpu_baseline(state, settings_fn(n), args.steps)
    traffic_pc, traffic_cov = substep_traffic(n, args.steps, args.warmup, list(st.flux_shape_cell_substeps))
    line = {
        "metric": "film cell-steps/s", "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain film {n}x{n} FP32, 64 seeded Gaussian blobs (radius 2-40 cells, peak "
                               f"20 um-2 mm), {index_words(args.index)}, yield stress log-U[0.1, 50] Pa, viscosity "
                               f"log-U[0.5, 50] Pa s, Dirichlet(0.3) pigments, tilt 15 deg, frame_dt 1/60, "
                               f"{args.steps} frames from the initial state, display shade pass off",
                   "index": args.index,
                   "cells": n * n, "graphs": not args.no_graphs, "substep_launch": args.substep_launch,
                   "one_launch_frames": st.one_launch_frames / max(1, st.frames_counted),
                   "flux_shape": args.flux_shape,
                   "flux_shape_cell_substeps": list(timed_st.flux_shape_cell_substeps),
                   "active_cells_per_frame": st.act_cells / frames,
                   "substeps_per_frame": st.act_cell_substeps / max(1, st.act_cells),
                   "implicit_cell_steps": st.act_cell_implicit,
                   "frame_algo_GBps": frame_gbs},
        "roofline": {"bound": "hbm", "kernel": "substep (k_lap_plug[_w], k_fluxes[_w|_faces], k_limit_apply[_w], "
                                                 "k_flux_apply[_w], k_substep_fused)",
                     "unit_of_work": "one substep over the active tiles (the sum of its launches)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic_pc is None else traffic_pc * cells_per_launch,
                     "traffic_bytes_per_cell": traffic_pc, "traffic_coverage": traffic_cov,
                     "measured_copy_GBps": measured["copy_GBps"], "frac_of_measured_copy": achieved / measured["copy_GBps"],
                     "algo_bytes_per_substep": bytes_per_launch, "algo_bytes_per_cell": algo_per_cell,
                     "avg_substep_us": avg_us, "substeps": launches, "cells_per_substep": cells_per_launch,
                     "launches": {f: [round(prof_l[f][0], 3), prof_l[f][1]] for f in SUBSTEP_LAUNCH_FAMILIES
                                  if prof_l[f][1]}},
        "cpu_baseline": cpu,
        "plain_uniform_index": uniform,
        "km": km_rec,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
