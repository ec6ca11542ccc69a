"""Per-cell power-law index in the restated frame (settings.n_per_cell; the
reference's index is the film's, film.wgsl:95).  A stratum's pigment slot 15
holds its index (0 = the film's), a cell's index is the thickness-weighted
mean over its strata and a face's the height-weighted mean of its two cells'.

Pins: the cell index equals a numpy weighted mean over hand-built strata;
dry cells take the film's index; a canvas whose wet cells all carry n0 runs
as a film of index n0 without the flag; two blobs that never meet, one at
n = 0.375 and one at n = 1, each run as a film of their own index."""
import numpy as np

import scenes
from helpers import make_pair

W = 64
FLOW = dict(seed=13, nblobs=6, sigma_cells=(3.0, 7.0), peak=(2e-4, 1.5e-3), tau=(0.01, 0.3), mu=(0.5, 5.0))


def settings(n_index, per_cell, tilt=45.0):
    s = scenes.blobs_settings(W, tilt_deg=tilt)
    s.n_index = n_index
    s.n_per_cell = per_cell
    return s


def lockstep(films, frames, frame_dt=1.0 / 60.0):
    """Frames of src/film/mod.rs:1330-1400 on every film with one substep
    count, the largest any of them would choose, so that films whose flows
    differ elsewhere on the canvas integrate with the same time step."""
    for f in range(frames):
        if f > 0:
            for o in films:
                o.poll_stats()
        sub = max(o.choose_substeps(frame_dt) for o in films)
        for o in films:
            o.update_params(frame_dt, sub)
            o.encode_frame(sub)


def with_n(state, n_cells):
    """The state with stratum slot 15 set to n_cells (an (h, w) array) on every
    occupied stratum."""
    st = {k: v.copy() for k, v in state.items()}
    pig = st["strata_pig"]
    lo = pig[..., 7] & np.uint32(0xFFFF)
    hi = np.float16(0.0) + n_cells.astype(np.float16)
    pig[..., 7] = lo | (hi.view(np.uint16).astype(np.uint32)[:, :, None] << np.uint32(16))
    pig[..., 7] = np.where(st["strata_h"] > 0, pig[..., 7], lo)
    return st


def test_cell_index_is_thickness_weighted_mean():
    st = scenes.blobs_state(W, W, **FLOW)
    h = st["h"]
    wet = h > 0
    # three strata per wet cell with their own indices (0 in stratum 2 = the film's)
    sh = np.zeros((W, W, 4), np.float32)
    sh[..., 0] = 0.5 * h
    sh[..., 1] = 0.3 * h
    sh[..., 2] = h - sh[..., 0] - sh[..., 1]
    st["strata_h"] = sh
    pig = st["strata_pig"]
    pig[:, :, 1, :] = pig[:, :, 0, :]
    pig[:, :, 2, :] = pig[:, :, 0, :]
    rng = np.random.default_rng(3)
    nk = rng.uniform(0.4, 1.0, (W, W, 3)).astype(np.float16)
    nk[..., 2] = 0.0
    lo = pig[..., :3, 7] & np.uint32(0xFFFF)
    pig[..., :3, 7] = lo | (nk.view(np.uint16).astype(np.uint32) << np.uint32(16))
    pig[~wet] = 0
    sh[~wet] = 0
    film_n = 0.8
    _, o = make_pair(W, W, settings(film_n, 1), st, gpu=False)
    o.buf("n_cell")[...] = -1.0
    o.step()
    got = o.buf("n_cell").reshape(W, W)
    t = sh.astype(np.float32)
    n3 = nk.astype(np.float32)
    n3[..., 2] = np.float32(film_n)
    want = (t[..., 0] * n3[..., 0] + t[..., 1] * n3[..., 1] + t[..., 2] * n3[..., 2]) / (t[..., 0] + t[..., 1] + t[..., 2])
    assert wet.any() and (~wet).any()
    assert np.allclose(got[wet], want[wet], rtol=2e-6, atol=0)
    # cells dry all frame carry the film's index, as their rheology the
    # defaults (a cell wetted in the frame inherits its source's)
    dry = ~wet & (o.buf("h").reshape(W, W) <= 0)
    assert dry.any()
    assert (got[dry] == np.float32(film_n)).all()


def test_uniform_cell_index_runs_as_film_index():
    st = scenes.blobs_state(W, W, **FLOW)
    n0 = 0.625  # exact in f16
    _, a = make_pair(W, W, settings(1.0, 1), with_n(st, np.full((W, W), n0)), gpu=False)
    _, b = make_pair(W, W, settings(n0, 0), st, gpu=False)
    for _ in range(6):
        a.step()
        b.step()
    ha, hb = a.buf("h").astype(np.float64), b.buf("h").astype(np.float64)
    assert np.abs(hb - st["h"]).max() > 1e-6  # the paint moved
    assert np.abs(ha - hb).max() <= 1e-4 * hb.max()


def test_separate_blobs_follow_their_own_index():
    # one blob in each half, far from the middle line and from each other
    st_l = scenes.blobs_state(W, W, seed=21, nblobs=1, sigma_cells=(3.0, 3.0), peak=(1e-3, 1e-3), tau=(0.01, 0.01),
                              mu=(0.5, 0.5))
    h = np.zeros((W, W), np.float32)
    yy, xx = np.mgrid[0:W, 0:W] + 0.5
    parts = {}
    for k, cx in ((0, 16.0), (1, 48.0)):
        parts[k] = 1e-3 * np.exp(-0.5 * ((xx - cx) ** 2 + (yy - 20.0) ** 2) / 9.0)
        parts[k][((xx - cx) ** 2 + (yy - 20.0) ** 2) > 81.0] = 0.0
        h += parts[k].astype(np.float32)
    wet = h > 0
    st = {k: v.copy() for k, v in st_l.items()}
    st["h"] = h
    st["strata_h"] = np.zeros((W, W, 4), np.float32)
    st["strata_h"][..., 0] = h
    pig = np.zeros((W, W, 4, 8), np.uint32)
    pig[wet, 0, :] = st_l["strata_pig"][st_l["h"] > 0][0, 0, :]
    st["strata_pig"] = pig
    rheo = np.zeros((W, W, 2), np.float32)
    rheo[..., 0] = 1.0
    rheo[wet] = st_l["rheo"][st_l["h"] > 0][0]
    st["rheo"] = rheo
    left = xx < W / 2
    n_map = np.where(left, 0.375, 1.0)  # exact in f16
    _, mixed = make_pair(W, W, settings(0.7, 1, tilt=30.0), with_n(st, n_map), gpu=False)
    _, lo = make_pair(W, W, settings(0.375, 0, tilt=30.0), st, gpu=False)
    _, hi = make_pair(W, W, settings(1.0, 0, tilt=30.0), st, gpu=False)
    lockstep([mixed, lo, hi], 5)
    hm = mixed.buf("h").astype(np.float64)
    hl, hh = lo.buf("h").astype(np.float64), hi.buf("h").astype(np.float64)
    # the blobs flow along y only (tilt direction -90 degrees) and stay apart
    assert (hm[:, 28:36] == 0).all() and (hl[:, 28:36] == 0).all() and (hh[:, 28:36] == 0).all()
    # the two indices give visibly different flows, and each half follows its own
    assert np.abs(hl[:, 32:] - hh[:, 32:]).max() > 1e-5
    assert np.abs(hm[:, :32] - hl[:, :32]).max() <= 1e-4 * hl.max()
    assert np.abs(hm[:, 32:] - hh[:, 32:]).max() <= 1e-4 * hh.max()


def test_mirror_symmetric_index_keeps_the_flow_symmetric():
    """A blob and an index map both mirror-symmetric about x = W/2, flowing
    along y: the face index, a symmetric mean of its two cells' indices, keeps
    the flow symmetric (an upwind or one-sided face index would not)."""
    st = scenes.blobs_state(W, W, seed=21, nblobs=1, sigma_cells=(3.0, 3.0), peak=(1e-3, 1e-3), tau=(0.01, 0.01),
                            mu=(0.5, 0.5))
    yy, xx = np.mgrid[0:W, 0:W] + 0.5
    h = (1.5e-3 * np.exp(-0.5 * ((xx - W / 2) ** 2 + (yy - 24.0) ** 2) / 36.0)).astype(np.float32)
    h[((xx - W / 2) ** 2 + (yy - 24.0) ** 2) > 18.0 ** 2] = 0.0
    wet = h > 0
    src = st["strata_pig"][st["h"] > 0][0, 0, :]
    rh = st["rheo"][st["h"] > 0][0]
    st["h"] = h
    st["strata_h"] = np.zeros((W, W, 4), np.float32)
    st["strata_h"][..., 0] = h
    st["strata_pig"] = np.zeros((W, W, 4, 8), np.uint32)
    st["strata_pig"][wet, 0, :] = src
    st["rheo"] = np.zeros((W, W, 2), np.float32)
    st["rheo"][..., 0] = 1.0
    st["rheo"][wet] = rh
    st["m"] = np.zeros((W, W), np.float32)
    n_map = 0.375 + 0.625 * np.minimum(np.abs(xx - W / 2) / 12.0, 1.0)
    s = settings(0.7, 1, tilt=35.0)
    s.weave_um = 0.0
    _, o = make_pair(W, W, s, with_n(st, n_map), gpu=False)
    # one frame: later frames mix the indices into the strata, where their f16
    # rounding turns last-bit differences between mirror cells into 2^-11 steps
    o.step()
    hv = o.buf("h").astype(np.float64)
    assert np.abs(hv - h).max() > 1e-4 * h.max()  # the paint moved
    assert np.abs(hv - hv[:, ::-1]).max() <= 2e-6 * hv.max()
