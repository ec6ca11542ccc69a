"""The face flux of the restatement (face and face_pair, film.wgsl:297-446)
on a flat film covering the canvas on a tilted plane, where the pressure
and plug gradients vanish and no front is near, so every face carries the
tilt-driven flux of its own direction:

  q_x = W_AXIS q(|g|) g_x / |g|,   q_y = W_AXIS q(|g|) g_y / |g|,
  q_d = W_DIAG q(|g|) (g_x + g_y) / (sqrt2 |g|),
  q_e = W_DIAG q(|g|) (g_x - g_y) / (sqrt2 |g|),

g = rho (g_x, g_y) the in-plane gravity, q = hb_flux (pinned on its own
against the Nusselt, power-law and Bingham closed forms).  With two
viscosities side by side the face between them takes the height-weighted
log mean (equal heights: the geometric mean).  Newtonian, shear-thinning
and yield-stress paint."""
import math

import numpy as np
import pytest

import oracle as orc
import scenes
from cz_types import default_settings
from helpers import make_pair

L = orc.lib()
W_AXIS, W_DIAG_FLUX = 0.6666667, 0.23570226   # shaders/common.wgsl:96-97
S2 = 0.70710678


def flat_film(n_index, mu_left, mu_right, ty, hv=2e-4, w=48, tilt=30.0, tilt_dir=-60.0):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    s.tilt_dir_deg = tilt_dir
    s.n_index = n_index
    s.weave_um = 0.0
    st = scenes.blobs_state(w, w, seed=1, nblobs=1, sigma_cells=(3.0, 3.0))
    src = st["strata_pig"][st["h"] > 0][0]
    h = np.full((w, w), hv, np.float32)
    st["h"] = h
    st["m"] = np.zeros((w, w), np.float32)
    st["strata_h"] = np.zeros((w, w, 4), np.float32)
    st["strata_h"][..., 0] = h
    st["strata_pig"] = np.broadcast_to(src, (w, w, 4, 8)).copy()
    st["strata_pig"][:, :, 1:] = 0
    rheo = np.zeros((w, w, 2), np.float32)
    rheo[:, : w // 2, 0] = mu_left
    rheo[:, w // 2:, 0] = mu_right
    rheo[..., 1] = ty
    st["rheo"] = rheo
    _, o = make_pair(w, w, s, st, gpu=False)
    o.update_params(1.0 / 60.0, 8)
    o.buf("tile_moved")[...] = 1
    L.of_k_mark_compact(o._p)
    L.of_k_laplacian(o._p)
    L.of_k_plug(o._p)
    L.of_k_fluxes(o._p)
    return o


def hb(n, g, h, mu, ty):
    y, up = orc.C.c_float(), orc.C.c_float()
    return L.of_hb_flux(n, g, h, mu, ty, orc.C.byref(y), orc.C.byref(up))


@pytest.mark.parametrize("n_index,ty", [(1.0, 0.0), (0.7, 0.0), (0.45, 0.0), (1.0, 0.02)])
def test_flat_tilted_film_face_fluxes(n_index, ty):
    w = 48
    mu = 1.5
    hv = 2e-4
    o = flat_film(n_index, mu, mu, ty, hv=hv, w=w)
    P = o.params
    gx, gy = P.phys[0] * P.g[0], P.phys[0] * P.g[1]
    gm = math.hypot(gx, gy)
    assert abs(gx) > 0.1 * gm and abs(gy) > 0.1 * gm       # both components drive
    q = hb(n_index, gm, hv, mu, ty)
    assert q > 0
    want = np.array([W_AXIS * q * gx / gm, W_AXIS * q * gy / gm,
                     W_DIAG_FLUX * q * (gx + gy) * S2 / gm, W_DIAG_FLUX * q * (gx - gy) * S2 / gm])
    fl = o.buf("flux").astype(np.float64)
    inner = fl[4:-4, 4:w // 2 - 4].reshape(-1, 4)
    assert np.allclose(inner, want[None, :], rtol=2e-5, atol=0), (inner[0], want)
    # the tilt drives a uniform film without moving it: every cell's outflow
    # equals its inflow through the same faces of its neighbours
    assert np.allclose(fl[5:-5, 5:-5], fl[4:-6, 4:-6], rtol=1e-6, atol=0)


def test_viscosity_jump_takes_the_log_mean():
    w = 48
    mu_a, mu_b = 0.8, 12.5
    hv = 2e-4
    o = flat_film(1.0, mu_a, mu_b, 0.0, hv=hv, w=w)
    P = o.params
    gx = P.phys[0] * P.g[0]
    gy = P.phys[0] * P.g[1]
    gm = math.hypot(gx, gy)
    fl = o.buf("flux").astype(np.float64)
    j = w // 2
    i = w // 2 - 1                       # the x face between the two paints
    mu_face = math.sqrt(mu_a * mu_b)
    want = W_AXIS * hb(1.0, gm, hv, mu_face, 0.0) * gx / gm
    assert fl[j, i, 0] == pytest.approx(want, rel=2e-5)
    # inside each paint its own viscosity (Newtonian: q proportional to 1/mu)
    assert fl[j, 4, 0] == pytest.approx(W_AXIS * hb(1.0, gm, hv, mu_a, 0.0) * gx / gm, rel=2e-5)
    assert fl[j, w - 6, 0] == pytest.approx(W_AXIS * hb(1.0, gm, hv, mu_b, 0.0) * gx / gm, rel=2e-5)
    assert fl[j, 4, 0] / fl[j, w - 6, 0] == pytest.approx(mu_b / mu_a, rel=1e-5)
