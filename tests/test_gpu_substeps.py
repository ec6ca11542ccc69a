"""The substeps of a frame launched as one kernel with grid-wide barriers
between the passes (k_substeps), the flux and apply passes fused into one
launch, and each whole substep in one launch (k_substep_fused) give the bits
of the three launches per substep: films in each
launch mode step the same scenes (a settling blob
canvas with deposits, a thick canvas with implicit steps, and a canvas whose
active set outgrows the one-launch grid so that blocks take several tiles)."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu

PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "lay", "delta", "fnorm", "tile_flags",
         "tile_moved", "m", "sub_refl", "pres", "tmp", "acc")


def assert_same(a, b, tag):
    for name in PARTS:
        x, y = a.buffer(name), b.buffer(name)
        if not np.array_equal(x.view(np.uint32), y.view(np.uint32)):
            bad = np.flatnonzero(x.view(np.uint32).ravel() != y.view(np.uint32).ravel())
            raise AssertionError(f"{tag}: {name} differs at {bad.size} words, first {bad[:6]}")
    sa, sb = a.stats(), b.stats()
    assert (sa.h_max, sa.active_tiles, sa.substeps_last, sa.implicit_steps) == \
           (sb.h_max, sb.active_tiles, sb.substeps_last, sb.implicit_steps), tag


CASES = {
    "blobs": (96, 96, 25.0, dict(seed=5, nblobs=8, sigma_cells=(2.0, 8.0)), True),
    "implicit": (96, 96, 5.0, dict(seed=5, nblobs=6, sigma_cells=(6.0, 12.0), peak=(4e-3, 6e-3), mu=(0.05, 0.1),
                                   tau=(1e-3, 2e-3)), False),
    "large": (384, 384, 40.0, dict(seed=7, nblobs=160, sigma_cells=(4.0, 16.0)), False),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("graphs", [False, True])
def test_one_launch_equals_per_pass(gpu, case, graphs):
    w, h, tilt, kw, deposits = CASES[case]
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    s.height_m = h * s.width_m / w
    st = scenes.blobs_state(w, h, **kw)
    films = []
    # (substep launch mode, flux shape): per-pass, one launch, automatic,
    # per-pass with flux + apply fused, and the whole substep in one launch
    for mode, shape in ((0, 0), (1, 0), (-1, 0), (0, 2), (0, 5)):
        f = czf.Film(w, h, s)
        f.substep_launch(mode)
        f.use_graphs(graphs)
        f.flux_shape(shape)
        assert f.write_state_arrays(st)
        films.append(f)
    pig = scenes.jet_paint()
    implicit_seen = False
    peak_active = 0
    for fr in range(30):
        deps = []
        if deposits and fr % 7 == 3:
            x = (0.25 + 0.02 * fr) * w * films[0].dx
            deps = [czf.deposit_capsule((x, 0.5 * h * films[0].dx), (x + 4 * films[0].dx, 0.5 * h * films[0].dx),
                                        3 * films[0].dx, 2e-9, pig, films[0].dx)]
        for f in films:
            f.step(deps)
        if fr % 6 == 5 or fr < 3:
            for f in films:
                f.sync()
            implicit_seen = implicit_seen or films[0].stats().implicit_steps > 0
            peak_active = max(peak_active, films[0].stats().active_tiles)
            assert_same(films[0], films[1], f"{case} frame {fr + 1} forced")
            assert_same(films[0], films[2], f"{case} frame {fr + 1} auto")
            assert_same(films[0], films[3], f"{case} frame {fr + 1} fused flux + apply")
            assert_same(films[0], films[4], f"{case} frame {fr + 1} whole substep fused")
    for f in films:
        f.sync()
    assert_same(films[0], films[1], f"{case} end forced")
    assert_same(films[0], films[2], f"{case} end auto")
    assert_same(films[0], films[3], f"{case} end fused flux + apply")
    assert_same(films[0], films[4], f"{case} end whole substep fused")
    used = list(films[4].stats().flux_shape_cell_substeps)
    assert used[5] > 0 and sum(used) == used[4] + used[5], used
    if case == "implicit":
        assert implicit_seen, "the thick canvas should take implicit steps"
    if case == "large":
        assert peak_active > 256, f"the active set should outgrow the one-launch grid ({peak_active} tiles)"
