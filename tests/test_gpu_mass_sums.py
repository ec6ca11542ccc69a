"""The device FP64 sums of the mass report (cz_mass_sums) against an exact
emulation of their summation order in numpy float64: chunks of 65536 cells,
in each chunk thread t of 256 adds cells t, t + 256, ... in order, a shuffle
tree per wave64 (offsets 32 .. 1), the four waves as (w0 + w1) + (w2 + w3);
then one block folds the chunk records the same way.  Sums must match to
the bit, counts and maxima exactly (compute_stats, src/main.rs:1708-1720:
a non-finite height is counted and its cell skipped).  The film's mass
report, which runs on these sums, is checked against the restatement in
the parity tests."""
import numpy as np
import pytest

import canvaz_film as czf

pytestmark = pytest.mark.gpu

CHUNK = 65536


def block_tree(v):
    """[..., 256] per-thread values -> [...] block sum in the kernel's order."""
    w = v.reshape(v.shape[:-1] + (4, 64)).copy()
    for off in (32, 16, 8, 4, 2, 1):
        lanes = np.arange(64)
        src = lanes + off
        ok = src < 64
        add = np.zeros_like(w)
        add[..., ok] = w[..., src[ok]]
        w = w + add
    w0 = w[..., 0]
    return (w0[..., 0] + w0[..., 1]) + (w0[..., 2] + w0[..., 3])


def thread_sums(x, nthreads=256):
    """Sequential per-thread sums over a strided layout: x [..., k*256]."""
    k = x.shape[-1] // nthreads
    r = x.reshape(x.shape[:-1] + (k, nthreads))
    acc = np.zeros(r.shape[:-2] + (nthreads,), np.float64)
    for i in range(k):
        acc = acc + r[..., i, :]
    return acc


def emulate(h, m):
    n = h.size
    nparts = (n + CHUNK - 1) // CHUNK
    fin = np.isfinite(h)
    a = np.where(fin, h, 0.0).astype(np.float64)
    b = np.where(fin, m, 0.0).astype(np.float64)
    pad = nparts * CHUNK - n
    a = np.concatenate([a, np.zeros(pad)]).reshape(nparts, CHUNK)
    b = np.concatenate([b, np.zeros(pad)]).reshape(nparts, CHUNK)
    ph, pm = block_tree(thread_sums(a)), block_tree(thread_sums(b))
    # fold: thread t takes records t, t + 256, ...
    k = (nparts + 255) // 256
    fh = np.zeros(k * 256)
    fm = np.zeros(k * 256)
    fh[:nparts], fm[:nparts] = ph, pm
    sh = block_tree(thread_sums(fh[None, :])[0])
    sm = block_tree(thread_sums(fm[None, :])[0])
    hf = np.where(fin, h, 0.0)
    sf = np.where(fin, h + m, 0.0).astype(np.float32)
    return (float(sh), float(sm), int((fin & (h > 0)).sum()), int((~fin).sum()),
            np.float32(max(0.0, hf.max(initial=0.0))), np.float32(max(0.0, sf.max(initial=0.0))))


def check(h, m):
    import torch
    got = czf.mass_sums(torch.from_numpy(h).cuda(), torch.from_numpy(m).cuda())
    want = emulate(h, m)
    assert got.sum_h == want[0], (got.sum_h, want[0])
    assert got.sum_m == want[1], (got.sum_m, want[1])
    assert (got.wet_cells, got.nan_cells) == (want[2], want[3])
    assert np.float32(got.h_max) == want[4] and np.float32(got.s_max) == want[5]
    # and the plain float64 totals agree to rounding
    fin = np.isfinite(h)
    assert abs(got.sum_h - h[fin].astype(np.float64).sum()) <= 1e-12 * max(1.0, abs(got.sum_h))
    return got


@pytest.mark.parametrize("n", [0, 1, 255, 65536, 65537, 3 * 65536 + 1000, 2048 * 2048 + 17])
def test_mass_sums_bit_exact(gpu, n):
    rng = np.random.default_rng(n + 1)
    h = (rng.random(n, np.float32) * 2e-3 - 4e-4).astype(np.float32)
    h[h < 0] = 0.0
    m = (rng.random(n, np.float32) * 3e-5).astype(np.float32)
    if n > 300:
        h[rng.integers(0, n, 7)] = np.nan
        h[rng.integers(0, n, 3)] = np.inf
        h[rng.integers(0, n, 2)] = -np.inf
    check(h, m)


def test_mass_sums_rejects_a_short_workspace(gpu):
    import ctypes as C
    import torch
    n = 200000
    h = torch.zeros(n, dtype=torch.float32, device="cuda")
    ws = torch.empty(czf.lib().cz_mass_sums_workspace(n) - 48, dtype=torch.uint8, device="cuda")
    out = torch.empty(48, dtype=torch.uint8, device="cuda")
    rc = czf.lib().cz_mass_sums(h.data_ptr(), h.data_ptr(), n, ws.data_ptr(), ws.numel(), out.data_ptr(), None)
    assert rc == -1 and b"workspace" in czf.lib().cz_last_error()


def test_film_mass_report_uses_the_device_sums(gpu):
    import scenes
    w = h = 300
    f = czf.Film(w, h, scenes.blobs_settings(w))
    st = scenes.blobs_state(w, h, seed=3, nblobs=20)
    assert f.write_state_arrays(st)
    r = f.mass_report()
    dx = float(f.dx)
    got = emulate(st["h"].ravel(), st["m"].ravel())
    assert r.wet_ml == got[0] * dx * dx * 1e6
    assert r.wet_cells == got[2] and r.hmax == got[4]
    f.close()
