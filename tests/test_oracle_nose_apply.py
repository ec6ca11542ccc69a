"""nose (film.wgsl:680-723), limiter (521-537) and apply (803-863) of the
restatement, pass by pass on hand-built fields.

nose, on a still height field, against brute force over the lattice:
- a dry cell carries no offset; every offset a wet cell carries points at a
  dry cell in the canvas, so it is never shorter than the true distance to
  the nearest dry cell;
- a cell with a dry 8-neighbour points at the nearest one (1 or sqrt 2);
- pass k reaches exactly the cells within Chebyshev distance k of the dry
  set (one ring per pass: each pass reads the offsets of the one before);
- a straight dry edge gives the straight offset (0, -d) exactly, and after
  convergence no offset exceeds the true distance by half a cell;
- the front normal is stored for offsets up to 3 cells, zero beyond.

limiter + apply, against an edge-list restatement of the same face fluxes
(each of the 8-neighbour faces once, with its donor):
- alpha = min(1, h dx / (2 dt out)) from every face the cell drains through;
- the height moves by dt / dx times the net inflow, each face scaled by its
  donor's alpha, so the total height is conserved and stays >= 0;
- a dry cell that takes in paint inherits rheology and top stratum from the
  face with the largest inflow (first in scan order on ties);
- acc gains dt times each owned face's limited flux."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as orc
import scenes
from helpers import make_pair

L = orc.lib()


def film(w, h, tilt=20.0):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    _, o = make_pair(w, h, s, None, gpu=False)
    o.update_params(1.0 / 60.0, 8)
    return o


def activate(o):
    o.buf("tile_moved")[...] = 1
    L.of_k_mark_compact(o._p)


def unpack(d):
    d = np.asarray(d, np.uint32)
    ok = (d & 0x10000) != 0
    ox = ((d >> 8) & 0xFF).astype(np.int64) - 128
    oy = (d & 0xFF).astype(np.int64) - 128
    return ok, ox, oy


def true_dist(dry):
    """Euclidean distance from every cell to the nearest dry cell."""
    hh, ww = dry.shape
    js, is_ = np.nonzero(dry)
    yy, xx = np.mgrid[0:hh, 0:ww]
    d2 = (xx[..., None] - is_[None, None, :]) ** 2 + (yy[..., None] - js[None, None, :]) ** 2
    return np.sqrt(d2.min(-1)), d2


def cheb_dist(dry):
    hh, ww = dry.shape
    js, is_ = np.nonzero(dry)
    yy, xx = np.mgrid[0:hh, 0:ww]
    return np.maximum(np.abs(xx[..., None] - is_[None, None, :]), np.abs(yy[..., None] - js[None, None, :])).min(-1)


def blob_height(w, h, seed, frac_dry=0.5):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w] + 0.5
    f = np.zeros((h, w))
    for _ in range(6):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(4, 12)
        f += np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
    thr = np.quantile(f, frac_dry)
    return np.where(f > thr, 1e-4, 0.0).astype(np.float32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_nose_against_brute_force(seed):
    w, h = 48, 40
    o = film(w, h)
    hv = blob_height(w, h, seed)
    o.buf("h")[...] = hv
    o.buf("delta")[...] = 0
    activate(o)
    dry = hv <= 0
    d_true, _ = true_dist(dry)
    cheb = cheb_dist(dry)
    jj, ii = np.mgrid[0:h, 0:w]
    fnv = (C.c_float * 2)()
    for k in range(1, 31):
        L.of_k_nose(o._p)
        ok, ox, oy = unpack(o.buf("delta"))
        assert not ok[dry].any() and (o.buf("delta")[dry] == 0).all()
        # one ring per pass
        assert np.array_equal(ok, ~dry & (cheb <= k)), k
        # every offset points at a dry cell in the canvas
        tx, ty = ii + ox, jj + oy
        assert ((tx[ok] >= 0) & (tx[ok] < w) & (ty[ok] >= 0) & (ty[ok] < h)).all()
        assert dry[ty[ok], tx[ok]].all()
        length = np.hypot(ox, oy)
        assert (length[ok] >= d_true[ok] - 1e-9).all()
        # a dry 8-neighbour is taken at its own distance
        near = ok & (d_true <= math.sqrt(2) + 1e-9)
        assert np.allclose(length[near], d_true[near])
    # converged: within half a cell of the true distance
    ok, ox, oy = unpack(o.buf("delta"))
    assert ok[~dry].all()
    err = np.hypot(ox, oy)[~dry] - d_true[~dry]
    assert err.max() <= 0.5, err.max()
    # the front normal: for offsets up to 3 cells, that of front_normal
    fn = o.buf("fnorm")
    length = np.hypot(ox, oy)
    for j, i in zip(*np.nonzero(~dry)):
        if length[j, i] <= 3.0:
            L.of_fn_front_normal(o._p, j * w + i, fnv)
            assert (fn[j, i, 0], fn[j, i, 1]) == (fnv[0], fnv[1])
        else:
            assert (fn[j, i] == 0).all()


def test_nose_straight_edge():
    w, h = 32, 48
    o = film(w, h)
    hv = np.full((h, w), 1e-4, np.float32)
    hv[:10, :] = 0.0
    o.buf("h")[...] = hv
    o.buf("delta")[...] = 0
    activate(o)
    for _ in range(40):
        L.of_k_nose(o._p)
    ok, ox, oy = unpack(o.buf("delta"))
    j = np.arange(h)[:, None] * np.ones((1, w), int)
    wet = j >= 10
    assert ok[wet].all()
    assert (ox[wet] == 0).all() and np.array_equal(oy[wet], -(j[wet] - 9))


FACES = ((1, 0), (0, 1), (1, 1), (1, -1))   # owned faces of a cell: +x, +y, +x+y, +x-y


def edge_restatement(h, flux, dx, dt):
    """(alpha, new heights, per-owned-face limited flux, inflow source) from
    the owned-face fluxes, visiting each face once with its donor."""
    H_, W_ = h.shape
    out = np.zeros((H_, W_))
    faces = []
    for j in range(H_):
        for i in range(W_):
            for k, (di, dj) in enumerate(FACES):
                q = float(flux[j, i, k])
                ib, jb = i + di, j + dj
                if not (0 <= ib < W_ and 0 <= jb < H_):
                    assert q == 0.0
                    continue
                faces.append((i, j, ib, jb, k, q))
                if q > 0:
                    out[j, i] += q
                elif q < 0:
                    out[jb, ib] += -q
    alpha = np.ones((H_, W_))
    pos = out > 0
    alpha[pos] = np.minimum(1.0, 0.5 * h[pos].astype(np.float64) * dx / (dt * out[pos]))
    net = np.zeros((H_, W_))
    lim = np.zeros(flux.shape)
    for (i, j, ib, jb, k, q) in faces:
        a = alpha[j, i] if q >= 0 else alpha[jb, ib]
        lim[j, i, k] = q * a
        net[j, i] -= q * a
        net[jb, ib] += q * a
    hn = np.maximum(h + dt * net / dx, 0.0)
    return alpha, hn, lim, faces


@pytest.mark.parametrize("seed", [5, 6])
def test_limiter_apply_against_edge_list(seed):
    w, h = 40, 36
    o = film(w, h)
    rng = np.random.default_rng(seed)
    hv = rng.uniform(2e-4, 1e-3, (h, w)).astype(np.float32)
    dry = rng.random((h, w)) < 0.08
    hv[dry] = 0.0
    P = o.params
    dx, dt = P.dx, P.dt
    # fluxes [m^2/s] sized so that about a third of the cells hit the limiter
    scale = 0.5 * 6e-4 * dx / dt / 2.5
    fl = (rng.normal(0.0, scale, (h, w, 4))).astype(np.float32)
    fl[:, -1, 0] = 0.0; fl[-1, :, 1] = 0.0; fl[:, -1, 2] = 0.0; fl[-1, :, 2] = 0.0
    fl[:, -1, 3] = 0.0; fl[0, :, 3] = 0.0
    # a dry cell only takes in paint: every face turned so that its donor is wet
    for j in range(h):
        for i in range(w):
            for k, (di, dj) in enumerate(FACES):
                ib, jb = i + di, j + dj
                if not (0 <= ib < w and 0 <= jb < h):
                    continue
                if dry[j, i] and dry[jb, ib]:
                    fl[j, i, k] = 0.0
                elif dry[j, i]:
                    fl[j, i, k] = -abs(fl[j, i, k])
                elif dry[jb, ib]:
                    fl[j, i, k] = abs(fl[j, i, k])
    o.buf("h")[...] = hv
    # every wet cell one stratum with its own pigment word; rheology by cell
    o.buf("strata_h")[...] = 0
    o.buf("strata_h")[..., 0] = hv
    pig = o.buf("strata_pig")
    pig[...] = 0
    pig[:, :, 0, 0] = np.arange(w * h, dtype=np.uint32).reshape(h, w) + 1
    o.buf("rheo")[..., 0] = rng.uniform(0.5, 5.0, (h, w)).astype(np.float32)
    o.buf("rheo")[..., 1] = rng.uniform(0.0, 2.0, (h, w)).astype(np.float32)
    rheo0 = o.buf("rheo").copy()
    activate(o)
    o.buf("flux")[...] = fl
    o.buf("acc")[...] = 0.0
    alpha, hn, lim, faces = edge_restatement(hv.astype(np.float64), fl.astype(np.float64), dx, dt)
    assert (alpha < 1).mean() > 0.1
    L.of_k_limiter(o._p)
    assert np.allclose(o.buf("alpha"), alpha, rtol=2e-6, atol=0)
    L.of_k_apply(o._p)
    got = o.buf("h").astype(np.float64)
    assert np.allclose(got, hn, rtol=0, atol=2e-6 * hv.max())
    assert (got >= 0).all()
    assert got.sum() == pytest.approx(hv.astype(np.float64).sum(), rel=2e-6)
    assert np.allclose(o.buf("acc"), lim * dt, rtol=0, atol=2e-6 * np.abs(lim * dt).max())
    # inheritance: a dry cell that took in paint copies its largest donor
    took = dry & (got > 0)
    assert took.sum() >= 5
    nb = [(1, 0, 0, +1), (0, 1, 1, +1), (1, 1, 2, +1), (1, -1, 3, +1),
          (-1, 0, 0, -1), (0, -1, 1, -1), (-1, -1, 2, -1), (-1, 1, 3, -1)]
    for j, i in zip(*np.nonzero(took)):
        best, src = 0.0, None
        # the scan order of apply: own faces +x, +y, +x+y, +x-y, then the
        # neighbours' faces from -x, -y, -x-y, -x+y
        for (di, dj, k, own) in nb:
            ib, jb = i + di, j + dj
            if not (0 <= ib < w and 0 <= jb < h):
                continue
            inflow = -lim[j, i, k] if own > 0 else lim[jb, ib, k]
            if inflow > best:
                best, src = inflow, (ib, jb)
        assert src is not None
        si, sj = src
        assert tuple(o.buf("rheo")[j, i]) == tuple(rheo0[sj, si])
        assert o.buf("strata_pig")[j, i, 0, 0] == sj * w + si + 1
        assert o.buf("strata_h")[j, i, 0] == np.float32(got[j, i])


def test_apply_inflow_tie_keeps_the_first_face():
    """Two donors pouring exactly the same into a dry cell: the first face in
    apply's scan order (its own +x face before its left neighbour's) wins."""
    w, h = 32, 32
    o = film(w, h)
    hv = np.full((h, w), 5e-4, np.float32)
    hv[10, 10] = 0.0
    o.buf("h")[...] = hv
    o.buf("strata_h")[...] = 0
    o.buf("strata_h")[..., 0] = hv
    pig = o.buf("strata_pig")
    pig[...] = 0
    pig[:, :, 0, 0] = np.arange(w * h, dtype=np.uint32).reshape(h, w) + 1
    activate(o)
    fl = np.zeros((h, w, 4), np.float32)
    q = np.float32(1e-9)
    fl[10, 10, 0] = -q   # from (11, 10) through the cell's own +x face
    fl[10, 9, 0] = q     # from (9, 10) through its +x face
    o.buf("flux")[...] = fl
    L.of_k_limiter(o._p)
    assert o.buf("alpha")[10, 11] == 1.0 and o.buf("alpha")[10, 9] == 1.0
    L.of_k_apply(o._p)
    assert o.buf("h")[10, 10] > 0
    assert o.buf("strata_pig")[10, 10, 0, 0] == 10 * w + 11 + 1
