"""Serving loop: canvases admitted as they arrive, stepped together with the
canvases already running, retired when their frames are done.  Every
request's final state equals the same canvas stepped alone (bit for bit;
the stated bar is 1e-5 relative), admission respects the memory budget and
the batch limit, and recycled films behave as fresh ones."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu

PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "m", "sub_refl", "lay", "delta", "tile_flags", "fnorm",
         "pres", "tmp")


def trace(n=10, seed=3):
    rng = np.random.default_rng(seed)
    reqs, specs = [], []
    t = 0.0
    for k in range(n):
        w = int(rng.integers(40, 150))
        h = int(rng.integers(40, 150))
        steps = int(rng.integers(2, 9))
        tilt = float(rng.uniform(0.0, 50.0))
        s = scenes.blobs_settings(w, tilt_deg=tilt)
        s.height_m = h * s.width_m / w
        sc = czf.scene(seed=100 + k, nblobs=12, sigma_lo=2.0, sigma_hi=10.0)
        reqs.append(czf.request(k, w, h, steps, t, s, sc))
        specs.append((w, h, steps, s, sc))
        t += float(rng.exponential(0.004))
    return reqs, specs


def alone(w, h, steps, s, sc):
    f = czf.Film(w, h, s)
    f.seed_blobs(sc)
    for _ in range(steps):
        f.step()
    f.sync()
    return f


def assert_same(a, b, tag):
    for name in PARTS:
        x, y = a.buffer(name), b.buffer(name)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{tag}: {name} differs"


def test_each_result_equals_alone(gpu):
    reqs, specs = trace()
    srv = czf.Server(hbm_budget=8 << 30, max_members=4, keep_films=True)
    res, st = srv.run(reqs)
    assert st.admitted == len(reqs)
    assert 1 < st.peak_members <= 4
    for r, (w, h, steps, s, sc) in zip(res, specs):
        assert r.status == 0
        assert r.arrival_s <= r.admit_s <= r.done_s
        assert r.cell_steps == w * h * steps
        g = srv.film(r.id, w, h)
        assert g is not None and g.stats().frame == steps
        assert_same(g, alone(w, h, steps, s, sc), f"request {r.id}")
    assert st.cell_steps == sum(w * h * steps for (w, h, steps, _, _) in specs)
    srv.close()


def test_budget_and_recycling(gpu):
    reqs, specs = trace(12, seed=4)
    one = max(w * h for (w, h, _, _, _) in specs)
    budget = int(2.5 * (one * 600 + 12 * 2 ** 20))   # about two canvases at a time
    srv = czf.Server(hbm_budget=budget, max_members=8)
    res, st = srv.run(reqs)
    assert all(r.status == 0 for r in res)
    assert st.peak_bytes <= budget
    assert st.films_reused > 0
    assert st.films_created + st.films_reused == len(reqs)
    srv.close()


def test_reshaped_film_equals_fresh(gpu):
    big = czf.Film(160, 144, scenes.blobs_settings(160))
    w, h = 96, 72
    s = scenes.blobs_settings(w, tilt_deg=30.0)
    s.height_m = h * s.width_m / w
    sc = czf.scene(seed=9, nblobs=10, sigma_lo=2.0, sigma_hi=8.0)
    big.seed_blobs(czf.scene(seed=1))
    big.step()
    big.sync()
    big.reshape(w, h, s)
    big.seed_blobs(sc)
    for _ in range(5):
        big.step()
    big.sync()
    assert_same(big, alone(w, h, 5, s, sc), "reshaped")
    with pytest.raises(czf.FilmError):
        big.reshape(200, 200, scenes.blobs_settings(200))


def test_least_work_first_with_eviction_equals_alone(gpu):
    """Budget for one large canvas and two small ones: a long large canvas,
    shorter large ones and small ones behind it.  Least work first admits the
    small ones past a large one that does not fit; a large one waiting past
    max_wait evicts the long one (run at least max_wait, more work left) to
    host memory, and the evicted canvas resumes later.  Every result equals
    the canvas stepped alone, bit for bit."""
    specs, reqs = [], []
    plan = [(256, 256, 90, 0.0), (256, 256, 20, 0.001), (60, 50, 6, 0.002), (48, 64, 5, 0.004),
            (56, 56, 8, 0.006), (256, 256, 30, 0.008), (40, 40, 4, 0.010)]
    for k, (w, h, steps, t) in enumerate(plan):
        s = scenes.blobs_settings(w, tilt_deg=20.0 + 5 * k)
        s.height_m = h * s.width_m / w
        sc = czf.scene(seed=300 + k, nblobs=12, sigma_lo=2.0, sigma_hi=10.0)
        reqs.append(czf.request(k, w, h, steps, t, s, sc))
        specs.append((w, h, steps, s, sc))
    # one large canvas and two small ones fit, two large ones do not
    large, small = 256 * 256 * 600 + 12 * 2 ** 20, 64 * 64 * 600 + 12 * 2 ** 20
    budget = large + 2 * small + 4 * 2 ** 20
    assert 2 * large > budget
    srv = czf.Server(hbm_budget=budget, max_members=4, keep_films=True, max_wait_s=0.02)
    res, st = srv.run(reqs)
    print("evictions", st.evictions, "resumes", st.resumes, "peak members", st.peak_members, flush=True)
    assert st.evictions > 0 and st.resumes == st.evictions
    assert st.peak_bytes <= budget
    assert st.peak_host_bytes > 0
    for r, (w, h, steps, s, sc) in zip(res, specs):
        assert r.status == 0, (r.id, r.status)
        assert r.arrival_s <= r.admit_s <= r.done_s
        g = srv.film(r.id, w, h)
        assert g is not None and g.stats().frame == steps
        assert_same(g, alone(w, h, steps, s, sc), f"request {r.id}")
    # the small requests did not wait for the large ones to finish
    small_done = max(res[k].done_s for k in (2, 3, 4, 6))
    assert small_done < res[0].done_s and small_done < res[1].done_s
    srv.close()


@pytest.mark.parametrize("max_wait", [0.0, 0.02])
def test_failed_admission_leaves_a_clean_server(gpu, max_wait):
    """A request whose canvas cannot be seeded (a bad scene) fails the run
    after other canvases were admitted; the run releases them, and the same
    server then serves a new trace whose results equal the canvases alone."""
    reqs, specs = trace(6, seed=8)
    bad = reqs[3]
    bad.scene.n_lo, bad.scene.n_hi = 0.5, 9.0
    srv = czf.Server(hbm_budget=8 << 30, max_members=4, keep_films=True, max_wait_s=max_wait)
    with pytest.raises(czf.FilmError, match="power-law index"):
        srv.run(reqs)
    reqs2, specs2 = trace(5, seed=9)
    for r in reqs2:
        r.id += 100        # the first run's finished canvases are still kept under their ids
    res, st = srv.run(reqs2)
    assert st.admitted == len(reqs2)
    for r, (w, h, steps, s, sc) in zip(res, specs2):
        assert r.status == 0
        assert_same(srv.film(r.id, w, h), alone(w, h, steps, s, sc), f"request {r.id}")
    srv.close()
