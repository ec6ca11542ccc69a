"""1000 frames on the GPU against the restatement's committed trajectories
(tests/golden/long_run_48.npz at n = 1 and long_run_48_n07.npz at the
shear-thinning n = 0.7, written by tests/golden/make_long_run.py):
heights and strata within 1e-3 relative after 1000 FP32 frames (1e-5 after
100), tile flags and nose offsets bit-exact."""
import os

import numpy as np
import pytest

import canvaz_film as czf
from helpers import rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("n_index", [1.0, 0.7])
def test_long_run_matches_golden(gpu, n_index):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mlr", os.path.join(HERE, "golden", "make_long_run.py"))
    mlr = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mlr)
    gold = np.load(os.path.join(HERE, "golden", mlr.FIXTURES[n_index]))
    s, st = mlr.scene(n_index)
    assert mlr.digest(st) == str(gold["init_sha256"]), "the seeded scene changed (numpy generator?)"
    g = czf.Film(mlr.W, mlr.W, s, czf.default_pigment_table(), czf.cmf_table())
    assert g.write_state_arrays(st)
    dx = g.dx
    for f in range(1, mlr.FRAMES[-1] + 1):
        g.step(mlr.deposits(f, dx, czf.deposit_capsule))
        if f in mlr.FRAMES:
            tol = 1e-5 if f <= 100 else 1e-3
            e_h = rel_err(g.buffer("h"), gold[f"f{f}_h"], 1e-9)
            e_s = rel_err(g.buffer("strata_h"), gold[f"f{f}_strata_h"], 1e-9)
            assert e_h <= tol and e_s <= tol, (f, e_h, e_s)
            assert np.array_equal(g.buffer("tile_flags"), gold[f"f{f}_tile_flags"]), f
            assert np.array_equal(g.buffer("delta"), gold[f"f{f}_delta"]), f
            assert abs(g.mass_report().wet_ml - float(gold[f"f{f}_wet_ml"])) <= 1e-6 * float(gold[f"f{f}_wet_ml"])
    g.sync()
