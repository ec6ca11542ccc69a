"""The implicit hydrostatic solve of the restatement (film.wgsl:895-1302) as
a linear system A x = b on the cells of the active tiles, with A read off the
restated operator column by column (cg_lap_p + cg_iter_a on unit vectors):

- A is symmetric and positive definite, I + kappa L_mob + C L_wet^2 (face
  mobilities shared by both sides of a face, the capillary part the squared
  Laplacian of the wet cells), as conjugate gradients needs;
- the Jacobi preconditioner cg_d is exactly A's diagonal;
- the restated iterations follow textbook preconditioned CG from x0 = h
  (float64 here) step for step, and after many iterations reach
  numpy.linalg.solve(A, b);
- a flat film covering the canvas has b = A h: the solve leaves h as is.
Both with the capillary term in the implicit solve (capillary_implicit) and
without it.
"""
import numpy as np
import pytest

import oracle as orc
import scenes
from helpers import make_pair

L = orc.lib()


def fp(a):
    return orc._fp(a)


def thick_film(w=32, seed=7, cap=0):
    s = scenes.blobs_settings(w, tilt_deg=5.0)
    s.capillary_implicit = cap
    st = scenes.blobs_state(w, w, seed=seed, nblobs=2, sigma_cells=(3.0, 5.0), peak=(4e-3, 6e-3), mu=(0.05, 0.1),
                            tau=(1e-3, 2e-3))
    _, o = make_pair(w, w, s, st, gpu=False)
    o.step()
    o.step()
    assert o.implicit_steps() > 0, "the scene does not take the implicit path"
    assert 0.2 < (o.buf("h") > 0).mean() < 0.9, "wet and dry cells both"
    # the implicit part of the next frame: tiles, pressures, face mobilities
    L.of_k_mark_compact(o._p)
    L.of_k_laplacian(o._p)
    L.of_k_cg_setup(o._p)
    return o


def active_cells(o):
    w = o.w
    tw = (w + 15) // 16
    act = np.zeros((o.h, o.w), bool)
    for t in o.buf("tile_list")[:o.active_count()]:
        ty, tx = divmod(int(t), tw)
        act[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16] = True
    return np.flatnonzero(act.ravel())


def operator(o, cells):
    n = o.w * o.h
    A = np.zeros((len(cells), len(cells)))
    e = np.zeros(n, np.float32)
    q = np.zeros(n, np.float32)
    for k, c in enumerate(cells):
        e[c] = 1.0
        L.of_cg_apply(o._p, fp(e), fp(q))
        A[:, k] = q[cells]
        e[c] = 0.0
    return A


def apply(o, v, cells):
    n = o.w * o.h
    full = np.zeros(n, np.float32)
    full[cells] = v
    q = np.zeros(n, np.float32)
    L.of_cg_apply(o._p, fp(full), fp(q))
    return q[cells].astype(np.float64)


def pcg(A, b, x0, d, iters):
    x = x0.copy()
    r = b - A @ x
    z = r / d
    p = z.copy()
    rz = r @ z
    for _ in range(iters):
        q = A @ p
        pq = p @ q
        if pq <= 0:
            break
        a = rz / pq
        x = x + a * p
        r = r - a * q
        z = r / d
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x


@pytest.mark.parametrize("cap", [0, 1])
def test_operator_symmetric_positive_definite_with_exact_diagonal(cap):
    o = thick_film(cap=cap)
    assert (o.params.front3[0] > 0) == bool(cap)
    cells = active_cells(o)
    A = operator(o, cells)
    scale = np.abs(A).max()
    assert np.abs(A - A.T).max() <= 2e-6 * scale
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() >= 1.0 - 1e-4            # I + positive semidefinite parts
    assert ev.max() > 2.0                    # the mobility and capillary parts are there
    L.of_cg_iterate(o._p, 0)
    d = o.buf("cg_d").ravel()[cells].astype(np.float64)
    assert np.allclose(d, np.diag(A), rtol=2e-6, atol=0)


@pytest.mark.parametrize("cap", [0, 1])
def test_iterations_follow_textbook_pcg(cap):
    o = thick_film(cap=cap)
    cells = active_cells(o)
    A = operator(o, cells)
    h = o.buf("h").ravel()[cells].astype(np.float64)
    L.of_cg_iterate(o._p, 0)
    d = o.buf("cg_d").ravel()[cells].astype(np.float64)
    r0 = o.buf("cg_r").ravel()[cells].astype(np.float64)
    assert np.array_equal(o.buf("cg_x").ravel()[cells], o.buf("h").ravel()[cells])
    b = r0 + apply(o, h.astype(np.float32), cells)
    x_star = np.linalg.solve(A, b)
    step = np.abs(x_star - h).max()
    assert step > 1e-8, "the solve has nothing to do"
    for k in (1, 2, 3, 5, 8):
        L.of_cg_iterate(o._p, k)
        got = o.buf("cg_x").ravel()[cells].astype(np.float64)
        want = pcg(A, b, h, d, k)
        assert np.abs(got - want).max() <= 2e-3 * step, (k, np.abs(got - want).max() / step)
    L.of_cg_iterate(o._p, 300)
    got = o.buf("cg_x").ravel()[cells].astype(np.float64)
    assert np.abs(got - x_star).max() <= 2e-3 * step


def test_flat_covering_film_is_a_fixed_point():
    w = 32
    s = scenes.blobs_settings(w, tilt_deg=0.0)
    st = scenes.blobs_state(w, w, seed=3, nblobs=1, sigma_cells=(4.0, 4.0), mu=(0.05, 0.05), tau=(1e-3, 1e-3))
    flat = np.full((w, w), 4e-3, np.float32)
    src_pig = st["strata_pig"][st["h"] > 0][0]
    src_rheo = st["rheo"][st["h"] > 0][0]
    st["h"] = flat
    st["m"] = np.zeros((w, w), np.float32)
    st["strata_h"] = np.zeros((w, w, 4), np.float32)
    st["strata_h"][..., 0] = flat
    st["strata_pig"] = np.broadcast_to(src_pig, (w, w, 4, 8)).copy()
    st["strata_pig"][:, :, 1:] = 0
    st["rheo"] = np.broadcast_to(src_rheo, (w, w, 2)).copy()
    _, o = make_pair(w, w, s, st, gpu=False)
    o.update_params(1.0 / 60.0, 8)
    o.buf("tile_moved")[...] = 1
    L.of_k_mark_compact(o._p)
    L.of_k_laplacian(o._p)
    L.of_k_cg_setup(o._p)
    L.of_cg_iterate(o._p, 20)
    assert np.array_equal(o.buf("cg_x"), flat)
    assert not o.buf("cg_r").any()
