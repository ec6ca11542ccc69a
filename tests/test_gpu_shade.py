"""The device shade pass (k_shade: shade.wgsl:29-83 with km.wgsl:35-62) gives
the restated pass's f16 bits, on a blob canvas after frames of flow, curing
and a deposit, in the normal and the debug view, and running it at the end
of every frame leaves the film's state untouched."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes
from helpers import make_pair, oracle_capsule

pytestmark = pytest.mark.gpu

PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "lay", "m", "sub_refl")


@pytest.mark.parametrize("debug", [0, 1])
def test_shade_equals_restatement(gpu, debug):
    w, h = 128, 96
    s = scenes.blobs_settings(w, tilt_deg=30.0)
    s.height_m = h * s.width_m / w
    s.debug_view = debug
    s.cure_rate = 0.5   # cure within the run, so cured strata and relief are shaded too
    st = scenes.blobs_state(w, h, seed=13, nblobs=10, sigma_cells=(2.0, 10.0))
    g, o = make_pair(w, h, s, st)
    g.shade_output(True)
    plain, _ = make_pair(w, h, s, st, gpu=True)
    pig = scenes.jet_paint()
    for f in range(10):
        deps_g, deps_o = [], []
        if f == 3:
            p0, p1 = (0.3 * w * g.dx, 0.5 * h * g.dx), (0.6 * w * g.dx, 0.5 * h * g.dx)
            deps_g = [czf.deposit_capsule(p0, p1, 3 * g.dx, 3e-9, pig, g.dx)]
            deps_o = [oracle_capsule(p0, p1, 3 * g.dx, 3e-9, pig, g.dx)]
        g.step(deps_g)
        plain.step(deps_g)
        o.step(deps_o)
    g.sync()
    plain.sync()
    color, detail = o.shade()
    cg, dg = g.buffer("color"), g.buffer("detail")
    for name, a, b in (("color", cg, color), ("detail", dg, detail)):
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{name} differs at {bad[:6].tolist()} ({a[tuple(bad[0])]} vs {b[tuple(bad[0])]})"
    assert np.any(detail[..., 1] != 0), "some cells should carry a cure level"
    for name in PARTS:
        assert np.array_equal(g.buffer(name).view(np.uint32), plain.buffer(name).view(np.uint32)), name
