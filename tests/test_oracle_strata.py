"""strata_limit, strata_update and strata_fill of the restatement
(strata.wgsl:157-323), pass by pass on hand-built stacks.

- Band placement (stack_add_band, strata.wgsl:127-155): paint a donor
  stratum sends across a face lands at the donor's own heights [z0, z1] in
  the receiver's stack; the part overlapping a receiver stratum mixes into
  it in proportion to the overlap, the part above the receiver's top becomes
  a new top stratum of the donor's pigments; the mix weights the viscosity
  slot only over strata that carry one (pig_mix, strata.wgsl:45-59).
- Conservation: on a random canvas with random per-stratum face volumes,
  limited by strata_limit, the update moves pigment between cells without
  creating or losing any (every pigment's volume sum before = after, to
  the f16 storage of the fractions), given heights equal to the transported
  stack sums (so the final rescale to h is the identity).
- strata_fill: a wet cell without strata takes the top stratum of the
  4-neighbour with the largest stack sum (first of -x, -y, +x, +y on ties),
  one ring per pass (each pass reads the sums from before it).
"""
import numpy as np
import pytest

import oracle as orc
import scenes
from helpers import make_pair

L = orc.lib()
FACES = ((1, 0), (0, 1), (1, 1), (1, -1))


def film(w, h):
    s = scenes.blobs_settings(w, tilt_deg=20.0)
    _, o = make_pair(w, h, s, None, gpu=False)
    o.update_params(1.0 / 60.0, 8)
    return o


def activate(o):
    o.buf("tile_moved")[...] = 1
    L.of_k_mark_compact(o._p)


def decode(p8):
    p8 = np.asarray(p8, np.uint32)
    lo = (p8 & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
    hi = (p8 >> 16).astype(np.uint16).view(np.float16).astype(np.float64)
    out = np.empty(p8.shape[:-1] + (16,))
    out[..., 0::2] = lo
    out[..., 1::2] = hi
    return out


def pig16(vals):
    v = np.zeros(16, np.float32)
    for k, x in vals.items():
        v[k] = x
    return scenes.pack_pig(v[None])[0]


P2 = {1: 0.75, 4: 0.25, 13: 2.0}          # pigments 1 and 4, yield 2 Pa, no viscosity slot
P1 = {0: 1.0}
P3 = {7: 0.5, 9: 0.5, 14: 12.0}           # pigments 7 and 9, ln mu + 10 = 12


@pytest.mark.parametrize("tb", [3e-4, 6e-4])
def test_band_lands_at_the_donor_heights(tb):
    w = h = 16
    o = film(w, h)
    dx = o.params.dx
    A, B = (5, 5), (6, 5)
    v = 1e-4                               # thickness the donor's top stratum sends
    for name in ("h", "strata_h", "strata_cure", "acc_s"):
        o.buf(name)[...] = 0
    o.buf("strata_pig")[...] = 0
    sh, pig = o.buf("strata_h"), o.buf("strata_pig")
    sh[A[1], A[0], :2] = (2e-4, 2e-4)
    pig[A[1], A[0], 0] = pig16(P1)
    pig[A[1], A[0], 1] = pig16(P2)
    sh[B[1], B[0], 0] = tb
    pig[B[1], B[0], 0] = pig16(P3)
    o.buf("h")[A[1], A[0]] = 3e-4
    o.buf("h")[B[1], B[0]] = tb + v
    activate(o)
    o.buf("acc_s")[A[1], A[0], 0, 1] = v * dx   # the +x face of A, stratum 1
    for which in (1, 2, 3):                    # limit, update, copy back
        L.of_k_strata(o._p, which)
    assert o.buf("alpha_s")[A[1], A[0], 1] == 1.0
    ta = o.buf("strata_h")[A[1], A[0]]
    assert np.allclose(ta, (2e-4, 1e-4, 0, 0), rtol=1e-6, atol=1e-12)
    t = o.buf("strata_h")[B[1], B[0]].astype(np.float64)
    g = decode(o.buf("strata_pig")[B[1], B[0]])
    p2, p3 = decode(pig16(P2)), decode(pig16(P3))
    over = min(4e-4, tb) - 2e-4            # overlap of [2e-4, 4e-4] with [0, tb]
    part = v * over / 2e-4
    if tb >= 4e-4:
        assert np.allclose(t, (tb + v, 0, 0, 0), rtol=1e-6, atol=1e-12)
    else:
        assert np.allclose(t, (tb + part, v - part, 0, 0), rtol=1e-6, atol=1e-12)
        assert np.allclose(g[1], p2, atol=1e-3)            # the part above: a stratum of its own
    mix = (p3 * tb + p2 * part) / (tb + part)
    mix[14] = p3[14]                                         # only P3 carries a viscosity
    assert np.allclose(g[0], mix, rtol=2e-3, atol=2e-4)


def test_update_conserves_every_pigment():
    w, h = 32, 32
    o = film(w, h)
    dx = o.params.dx
    rng = np.random.default_rng(7)
    nstr = rng.integers(1, 4, (h, w))
    sh = np.zeros((h, w, 4), np.float32)
    fr = np.zeros((h, w, 4, 16), np.float32)
    for k in range(4):
        on = nstr > k
        sh[on, k] = rng.uniform(1e-5, 3e-4, on.sum())
        conc = rng.dirichlet(np.full(12, 0.5), (h, w))
        fr[..., k, :12] = conc
        fr[..., k, 13] = rng.uniform(0, 5, (h, w))
        fr[..., k, 14] = rng.uniform(8, 14, (h, w))
    fr[sh <= 0] = 0
    o.buf("strata_h")[...] = sh
    o.buf("strata_pig")[...] = scenes.pack_pig(fr)
    o.buf("strata_cure")[...] = 0
    o.buf("h")[...] = sh.sum(-1)
    activate(o)
    # random per-stratum volumes through every owned face (m^2), zero through the edge
    acc = rng.normal(0, 6e-5 * dx, (h, w, 4, 4)).astype(np.float32)
    acc[:, -1, 0] = 0; acc[-1, :, 1] = 0; acc[:, -1, 2] = 0; acc[-1, :, 2] = 0; acc[:, -1, 3] = 0; acc[0, :, 3] = 0
    o.buf("acc_s")[...] = acc
    L.of_k_strata(o._p, 1)
    a = o.buf("alpha_s").astype(np.float64)
    assert (a < 1).mean() > 0.05
    # the transported stack sums: each stratum drained by its limited outflow,
    # each receiver gaining the donor's limited band volumes
    th = sh.astype(np.float64)
    rem = th.copy()
    gain = np.zeros((h, w))
    for j in range(h):
        for i in range(w):
            for f, (di, dj) in enumerate(FACES):
                ib, jb = i + di, j + dj
                if not (0 <= ib < w and 0 <= jb < h):
                    continue
                for k in range(4):
                    q = float(acc[j, i, f, k])
                    if q > 0 and th[j, i, k] > 0:
                        vol = np.float32(np.float32(q) * np.float32(a[j, i, k])) * np.float32(1.0 / dx)
                        rem[j, i, k] -= vol
                        gain[jb, ib] += vol
                    elif q < 0 and th[jb, ib, k] > 0:
                        vol = np.float32(np.float32(-q) * np.float32(a[jb, ib, k])) * np.float32(1.0 / dx)
                        rem[jb, ib, k] -= vol
                        gain[j, i] += vol
    assert (rem >= -1e-6 * th - 1e-15).all()
    total = (np.maximum(rem, 0).sum(-1) + gain).astype(np.float32)
    o.buf("h")[...] = total
    before = (th[..., None] * decode(scenes.pack_pig(fr))).sum((0, 1, 2))
    L.of_k_strata(o._p, 2)
    L.of_k_strata(o._p, 3)
    th2 = o.buf("strata_h").astype(np.float64)
    after = (th2[..., None] * decode(o.buf("strata_pig"))).sum((0, 1, 2))
    assert th2.sum() == pytest.approx(float(total.astype(np.float64).sum()), rel=2e-6)
    assert np.allclose(after[:12], before[:12], rtol=1e-3, atol=0), (after[:12] / before[:12])
    # every stack a prefix of non-empty strata summing to the cell height
    assert np.allclose(th2.sum(-1), total, rtol=2e-6, atol=1e-12)
    nz = th2 > 0
    assert (nz[..., 1:] <= nz[..., :-1]).all()


def test_fill_one_ring_per_pass_and_candidate_order():
    w, h = 32, 16
    o = film(w, h)
    for name in ("h", "strata_h", "strata_cure"):
        o.buf(name)[...] = 0
    o.buf("strata_pig")[...] = 0
    # a wet row with a stratum only at its left end
    o.buf("h")[4, 2:20] = 1e-4
    o.buf("strata_h")[4, 2, 0] = 1e-4
    o.buf("strata_pig")[4, 2, 0] = pig16(P1)
    o.buf("strata_cure")[4, 2, 0] = 0.25
    # a cell between two filled neighbours: -x has the larger stack sum (two
    # strata, the thinner bottom one under the +x cell's single stratum), so
    # it wins and gives its top stratum
    o.buf("h")[10, 5:8] = 2e-4
    o.buf("strata_h")[10, 5, :2] = (2.5e-4, 1e-4)
    o.buf("strata_pig")[10, 5, 1] = pig16(P2)
    o.buf("strata_h")[10, 7, 0] = 3e-4
    o.buf("strata_pig")[10, 7, 0] = pig16(P3)
    # a tie: equal sums on -y and +y, the first in order (-y) wins
    o.buf("h")[12:15, 25] = 2e-4
    o.buf("strata_h")[12, 25, 0] = 1e-4
    o.buf("strata_pig")[12, 25, 0] = pig16(P2)
    o.buf("strata_h")[14, 25, 0] = 1e-4
    o.buf("strata_pig")[14, 25, 0] = pig16(P3)
    activate(o)
    for k in range(1, 7):
        L.of_k_strata(o._p, 4)
        row = o.buf("strata_h")[4, :, 0]
        filled = row > 0
        assert filled[2:3 + k].all() and not filled[3 + k:].any(), k
        assert np.array_equal(row[3:3 + k], o.buf("h")[4, 3:3 + k])
        assert (o.buf("strata_pig")[4, 3:3 + k, 0] == pig16(P1)).all()
        assert (o.buf("strata_cure")[4, 3:3 + k, 0] == np.float32(0.25)).all()
    assert (o.buf("strata_pig")[10, 6, 0] == pig16(P2)).all()
    assert o.buf("strata_h")[10, 6, 0] == np.float32(2e-4)
    assert (o.buf("strata_pig")[13, 25, 0] == pig16(P2)).all()
