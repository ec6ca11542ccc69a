"""The seeded blob-canvas generator (scene_gen.h) on the host: the
distributions it promises (Dirichlet(alpha) moments, log-uniform ranges,
truncated Gaussian heights), determinism per seed and independence from how
rows are split.  The device generator is checked against this host twin in
test_gpu_scene.py."""
import numpy as np
import pytest

import canvaz_film as czf

DX = 1.2 / 2048


def decode(pig):
    return pig.view(np.uint16).reshape(pig.shape[:-1] + (16,)).view(np.float16).astype(np.float64)


@pytest.fixture(scope="module")
def canvas():
    sc = czf.scene(seed=11)
    return sc, czf.scene_cells_host(320, 288, DX, sc)


def test_dirichlet_moments(canvas):
    sc, (h, m, r, p) = canvas
    wet = h > 0
    f = decode(p)[wet]
    conc = f[:, :12]
    k, a = 12, sc.alpha
    a0 = k * a
    assert np.allclose(conc.sum(1), 1.0, atol=2e-3)          # f16 storage of each share
    assert np.allclose(conc.mean(0), 1.0 / k, atol=4e-3)
    var = a * (a0 - a) / (a0 * a0 * (a0 + 1.0))
    assert np.allclose(conc.var(0), var, rtol=0.08)
    assert (conc >= 0).all()


def test_log_uniform_ranges(canvas):
    sc, (h, m, r, p) = canvas
    wet = h > 0
    f = decode(p)[wet]
    tau, lmu = f[:, 13], f[:, 14] - 10.0
    assert tau.min() >= sc.tau_lo * (1 - 1e-3) and tau.max() <= sc.tau_hi * (1 + 1e-3)
    lt = np.log(tau)
    # uniform in log space: quartiles of the normalised log
    u = (lt - np.log(sc.tau_lo)) / (np.log(sc.tau_hi) - np.log(sc.tau_lo))
    assert np.allclose(np.quantile(u, [0.25, 0.5, 0.75]), [0.25, 0.5, 0.75], atol=0.02)
    v = (lmu - np.log(sc.mu_lo)) / (np.log(sc.mu_hi) - np.log(sc.mu_lo))
    assert np.allclose(np.quantile(v, [0.25, 0.5, 0.75]), [0.25, 0.5, 0.75], atol=0.02)
    # rheology from the stored f16 values; dry cells (1, 0)
    assert np.allclose(r[wet][:, 0], np.exp(f[:, 14] - 10.0), rtol=2e-6)
    assert np.array_equal(r[wet][:, 1], f[:, 13].astype(np.float32))
    assert (r[~wet] == np.array([1.0, 0.0], np.float32)).all()


def test_heights_are_truncated_gaussians():
    """One blob: h = pk exp(-d^2 / (2 sigma^2)) inside 3 sigma, zero outside."""
    sc = czf.scene(seed=3, nblobs=1, sigma_lo=6.0, sigma_hi=6.0, peak_lo=1e-3, peak_hi=1e-3)
    h, m, r, p = czf.scene_cells_host(64, 64, DX, sc)
    j, i = np.nonzero(h == h.max())
    cx, cy = i[0] + 0.5, j[0] + 0.5
    yy, xx = np.mgrid[0:64, 0:64] + 0.5
    inside = h > 0
    # the centre lies within half a cell of the peak cell; fit it from the moments
    w = h.astype(np.float64)
    mx, my = (w * xx).sum() / w.sum(), (w * yy).sum() / w.sum()
    d2 = (xx - mx) ** 2 + (yy - my) ** 2
    assert 18.0 < mx < 46.0 and 18.0 < my < 46.0   # seed 3 puts the blob clear of the edges
    assert np.allclose(h[inside], 1e-3 * np.exp(-d2[inside] / 72.0), rtol=2e-3, atol=2e-9)
    assert (d2[inside] <= 324.0 + 1.5).all() and (d2[~inside] >= 324.0 - 1.5).all()
    assert abs(cx - mx) <= 1.0 and abs(cy - my) <= 1.0


def test_weave_matches_reference_formula(canvas):
    sc, (h, m, r, p) = canvas
    ref = np.zeros(320 * 288, np.float32)
    czf.lib().cz_weave(320, 288, DX, 25e-6, czf._fp(ref))
    assert np.array_equal(m, ref.reshape(288, 320))


def test_deterministic_and_row_split():
    sc = czf.scene(seed=5)
    a = czf.scene_cells_host(96, 80, DX, sc)
    b = czf.scene_cells_host(96, 80, DX, sc, 0, 40)
    c = czf.scene_cells_host(96, 80, DX, sc, 40, 80)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, np.concatenate([y, z], axis=0))
    d = czf.scene_cells_host(96, 80, DX, czf.scene(seed=6))
    assert not np.array_equal(a[0], d[0])


def test_cell_index_range():
    """n_hi > 0: every wet cell draws its power-law index uniform in
    [n_lo, n_hi] into slot 15 (one more draw after the viscosity, so the
    other fields are those of the same scene without it)."""
    sc0 = czf.scene(seed=11)
    sc1 = czf.scene(seed=11, n_lo=0.4, n_hi=1.0)
    h0, m0, r0, p0 = czf.scene_cells_host(160, 144, DX, sc0)
    h1, m1, r1, p1 = czf.scene_cells_host(160, 144, DX, sc1)
    assert np.array_equal(h0, h1) and np.array_equal(r0, r1)
    f0, f1 = decode(p0), decode(p1)
    assert np.array_equal(f0[..., :15], f1[..., :15])
    wet = h1 > 0
    n = f1[wet][:, 15]
    assert (f0[..., 15] == 0).all() and (f1[~wet][:, 15] == 0).all()
    assert n.min() >= 0.4 - 2e-4 and n.max() <= 1.0
    u = (n - 0.4) / 0.6
    assert np.allclose(np.quantile(u, [0.25, 0.5, 0.75]), [0.25, 0.5, 0.75], atol=0.02)


@pytest.mark.parametrize("lo,hi", [(0.0, 1.0), (0.5, 0.4), (0.5, 5.0)])
def test_cell_index_range_rejected(lo, hi):
    with pytest.raises(czf.FilmError):
        czf.scene_cells_host(16, 16, DX, czf.scene(seed=1, n_lo=lo, n_hi=hi))
