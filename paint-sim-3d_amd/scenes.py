"""Seeded synthetic workloads of the film step (numpy PCG64, seeds explicit).

- blobs_state: the plain single-canvas film (2048x2048 by default): h as a sum
  of 64 Gaussian blobs (sigma 2-40 cells, peak log-uniform 20 um - 2 mm,
  truncated at 3 sigma, zero below DRAINED_FILM), one stratum per wet cell
  with Dirichlet(0.3) pigments, yield stress log-uniform [0.1, 50] Pa and
  viscosity log-uniform [0.5, 50] Pa s per cell, stored as the reference
  stores them: pigment slots 13 (tau_y) and 14 (ln mu + 10) in f16
  (shaders/cure.wgsl:62-117).
- jet_deposits: the headless jet scene's source (src/main.rs:1410-1432):
  a nozzle on the Lissajous path x = 0.6 + 0.36 sin(1.667 t),
  y = 0.4 + 0.24 sin(3.333 t) laying pure Ivory Black (thinner 0.3,
  ln mu = ln 4) as one capsule per frame for 5 < frame < 399.  The reference
  lays the rod's landed ribbon (an unseeded CPU model); this replays a fixed
  stream of that path and flow instead.
- scripted_deposits: the deterministic no-script scene of the reference
  (src/main.rs:386-419), four capsule pours.
"""
import math

import numpy as np

import canvaz_film as czf
from cz_types import default_settings

FRAME_DT = 1.0 / 60.0


def f16_bits(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)


def pack_pig(frac):
    """[..., 16] float -> [..., 8] u32 (pack2x16float pairs, round to nearest even)."""
    b = f16_bits(frac)
    return (b[..., 0::2] | (b[..., 1::2] << 16)).astype(np.uint32)


def blobs_settings(w, tilt_deg=15.0, dx=1.2 / 2048):
    """Square canvas at the plain film's cell size (1.2 m over 2048 cells)."""
    s = default_settings()
    s.width_m = w * dx
    s.height_m = w * dx
    s.tilt_deg = tilt_deg
    s.tilt_dir_deg = -90.0
    return s


def blobs_state(w, h, seed=1, nblobs=64, sigma_cells=(2.0, 40.0), peak=(20e-6, 2e-3), tau=(0.1, 50.0),
                mu=(0.5, 50.0), dirichlet=0.3, weave=None, cell_m=None, n_range=None):
    """State parts (dict of numpy arrays) of the plain blob canvas.

    n_range=(lo, hi) gives every wet cell its own power-law index, uniform in
    [lo, hi], in pigment slot 15; a film reads it when n_per_cell is set."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hgt = np.zeros((h, w), np.float64)
    cx = rng.uniform(0, w, nblobs)
    cy = rng.uniform(0, h, nblobs)
    sg = rng.uniform(sigma_cells[0], sigma_cells[1], nblobs)
    pk = np.exp(rng.uniform(math.log(peak[0]), math.log(peak[1]), nblobs))
    for k in range(nblobs):
        r = 3.0 * sg[k]
        x0, x1 = max(0, int(cx[k] - r)), min(w, int(cx[k] + r) + 2)
        y0, y1 = max(0, int(cy[k] - r)), min(h, int(cy[k] + r) + 2)
        if x0 >= x1 or y0 >= y1:
            continue
        xs = np.arange(x0, x1) + 0.5
        ys = np.arange(y0, y1) + 0.5
        d2 = (xs[None, :] - cx[k]) ** 2 + (ys[:, None] - cy[k]) ** 2
        g = pk[k] * np.exp(-0.5 * d2 / sg[k] ** 2)
        g[d2 > r * r] = 0.0
        hgt[y0:y1, x0:x1] += g
    hgt = hgt.astype(np.float32)
    hgt[hgt < 1e-9] = 0.0
    n = w * h
    wet = hgt > 0
    frac = np.zeros((h, w, 16), np.float32)
    nw = int(wet.sum())
    conc = rng.dirichlet(np.full(12, dirichlet), nw).astype(np.float32)
    ty = np.exp(rng.uniform(math.log(tau[0]), math.log(tau[1]), nw)).astype(np.float32)
    mu_v = np.exp(rng.uniform(math.log(mu[0]), math.log(mu[1]), nw)).astype(np.float32)
    fw = np.zeros((nw, 16), np.float32)
    fw[:, :12] = conc
    fw[:, 13] = ty
    fw[:, 14] = np.log(mu_v) + 10.0
    if n_range is not None:
        fw[:, 15] = rng.uniform(n_range[0], n_range[1], nw).astype(np.float32)
    frac[wet] = fw
    pig = np.zeros((h, w, 4, 8), np.uint32)
    pig[:, :, 0, :] = pack_pig(frac)
    strata_h = np.zeros((h, w, 4), np.float32)
    strata_h[:, :, 0] = hgt
    rheo = np.zeros((h, w, 2), np.float32)
    rheo[:, :, 0] = 1.0
    q = frac.astype(np.float16).astype(np.float32)
    rheo[wet, 0] = np.exp(q[wet, 14] - 10.0)
    rheo[wet, 1] = q[wet, 13]
    dx = 1.2 / 2048 if cell_m is None else cell_m
    if weave is None:
        m = np.zeros(n, np.float32)
        czf.lib().cz_weave(w, h, dx, 25e-6, czf._fp(m))
        m = m.reshape(h, w)
    else:
        m = weave
    sub = np.zeros(n * 8, np.uint32)
    czf.lib().cz_substrate_refl(n, 1, czf._up(sub))
    return {
        "h": hgt, "m": m, "rheo": rheo, "strata_h": strata_h, "strata_pig": pig,
        "strata_cure": np.zeros((h, w, 4), np.float32), "sub_refl": sub.reshape(h, w, 8),
        "lay": np.zeros((h, w), np.uint32),
    }


def jet_settings():
    s = default_settings()
    s.sigma = 0.04  # tool sigma written every frame (src/main.rs:2186, src/tools/mod.rs:207)
    return s


def jet_paint():
    """PaintMix::pure(1) with thinner 0.3, slot 14 = ln 4 + 10 (src/main.rs:264-265,2185)."""
    f = np.zeros(16, np.float32)
    f[1] = 1.0
    f[12] = 0.3
    f[14] = math.log(4.0) + 10.0
    return czf.paint_pack(f, normalize=True)


def jet_flow_volume(frame_dt=FRAME_DT):
    """q = 0.016 d^2 (g nu u)^(1/4) with d = 1 cm, nu = 4*0.7^3/1000, u = 0.5
    (src/tools/mod.rs:222-230) -> volume per frame."""
    d, nu, u = 0.01, 4.0 * 0.7 ** 3 / 1000.0, 0.5
    q = 0.016 * d * d * (9.81 * nu * u) ** 0.25
    return q * frame_dt


def jet_path(t):
    return (0.6 + 0.36 * math.sin(1.667 * t), 0.4 + 0.24 * math.sin(3.333 * t))


def jet_deposits(frame, dx, make_capsule=None, frame_dt=FRAME_DT, radius=1.5e-3):
    """Deposits of frame `frame` of the jet stream (empty outside 5 < f < 399)."""
    if not (5 < frame < 399):
        return []
    cap = make_capsule or czf.deposit_capsule
    p0 = jet_path((frame - 1) * frame_dt)
    p1 = jet_path(frame * frame_dt)
    d = cap(p0, p1, radius, jet_flow_volume(frame_dt), jet_paint(), dx)
    d.kind |= 1 << 8  # laid by thread (Method::Thread)
    return [d]


def scripted_deposits(frame, w_m, h_m, dx, make_capsule=None, frame_dt=FRAME_DT):
    """App::scripted_deposits (src/main.rs:386-419)."""
    cap = make_capsule or czf.deposit_capsule
    out = []
    pour = 3 << 8
    if 10 <= frame < 130:
        t0 = np.float32(frame - 10) / np.float32(120.0)
        t1 = np.float32(frame - 9) / np.float32(120.0)

        def curve(t):
            t = np.float32(t)
            return (np.float32(w_m) * (np.float32(0.15) + np.float32(0.7) * t),
                    np.float32(h_m) * (np.float32(0.3) + np.float32(0.35) * np.float32(math.sin(np.float32(t * np.float32(6.0)))) * np.float32(0.5) + np.float32(0.2) * t))
        f = np.zeros(16, np.float32)
        f[1] = 1.0
        f[12] = 0.3
        d = cap(curve(t0), curve(t1), 0.0012, 1.5e-7 * frame_dt, czf.paint_pack(f, True), dx)
        d.kind |= pour
        out.append(d)
    if 40 <= frame < 100:
        t0 = np.float32(frame - 40) / np.float32(60.0)
        t1 = np.float32(frame - 39) / np.float32(60.0)

        def line(t):
            return (np.float32(w_m) * (np.float32(0.2) + np.float32(0.6) * t),
                    np.float32(h_m) * (np.float32(0.7) - np.float32(0.3) * t))
        f = np.zeros(16, np.float32)
        f[3] = 1.0
        d = cap(line(t0), line(t1), 0.0025, 1.2e-6 * frame_dt, czf.paint_pack(f, True), dx)
        d.kind |= pour
        out.append(d)
    if 60 <= frame < 75:
        f = np.zeros(16, np.float32)
        f[2] = 1.0
        f[0] = 0.3
        f[12] = 0.3
        c = (w_m * 0.5, h_m * 0.5)
        d = cap(c, c, 0.02, 6.0e-6 * frame_dt * 4.0, czf.paint_pack(f, True), dx)
        d.kind |= pour
        out.append(d)
    if 150 <= frame < 170:
        f = np.zeros(16, np.float32)
        f[5] = 1.0
        f[12] = 0.1
        d = cap((w_m * 0.5, h_m * 0.5), (w_m * 0.52, h_m * 0.51), 0.008, 4.0e-6 * frame_dt,
                czf.paint_pack(f, True), dx)
        d.kind |= pour
        out.append(d)
    return out
