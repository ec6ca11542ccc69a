"""cz_mass_sums_noalloc (the allocation-free form, partials in the library's
static device array) against cz_mass_sums with a caller workspace, through
the C ABI: equal to the bit on two calls in a row, the second on fewer
cells (so the array holds the first call's records past the second's), a
call beyond the array's maximum refused before anything runs, and the call
captured in a HIP graph and replayed, each replay equal to the eager
result."""
import ctypes as C

import numpy as np
import pytest

import canvaz_film as czf

pytestmark = pytest.mark.gpu


def inputs(n, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    h = torch.rand(n, generator=g, device="cuda", dtype=torch.float32) * 2e-3 - 4e-4
    h.clamp_(min=0.0)
    m = torch.rand(n, generator=g, device="cuda", dtype=torch.float32) * 3e-5
    if n > 100:
        h[::9973] = float("nan")
    return h, m


def with_workspace(h, m):
    import torch
    n = h.numel()
    ws = torch.empty(max(16, czf.lib().cz_mass_sums_workspace(n)), dtype=torch.uint8, device="cuda")
    out = torch.zeros(48, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    czf.check(czf.lib().cz_mass_sums(h.data_ptr(), m.data_ptr(), n, ws.data_ptr(), ws.numel(), out.data_ptr(), st))
    torch.cuda.synchronize()
    return out.cpu().numpy().tobytes()


def no_alloc(h, m, out):
    import torch
    st = torch.cuda.current_stream().cuda_stream
    czf.check(czf.lib().cz_mass_sums_noalloc(h.data_ptr(), m.data_ptr(), h.numel(), out.data_ptr(), st))


def test_no_alloc_equals_workspace_form(gpu):
    import torch
    out = torch.zeros(48, dtype=torch.uint8, device="cuda")
    for n, seed in ((2048 * 2048 + 17, 1), (70001, 2)):
        h, m = inputs(n, seed)
        no_alloc(h, m, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().tobytes()
        want = with_workspace(h, m)
        assert got == want, f"{n} cells: no-alloc {czf.MassTotals.from_buffer_copy(got).sum_h!r} vs " \
                            f"workspace {czf.MassTotals.from_buffer_copy(want).sum_h!r}"
        t = czf.MassTotals.from_buffer_copy(got)
        assert t.nan_cells == len(range(0, n, 9973)) and t.sum_h > 0.0


def test_no_alloc_rejects_beyond_its_maximum(gpu):
    import torch
    h = torch.zeros(16, dtype=torch.float32, device="cuda")
    out = torch.zeros(48, dtype=torch.uint8, device="cuda")
    rc = czf.lib().cz_mass_sums_noalloc(h.data_ptr(), h.data_ptr(), 4096 * 65536 + 1, out.data_ptr(), None)
    assert rc == -1
    assert b"maximum" in czf.lib().cz_last_error()
    assert bytes(out.cpu().numpy()) == bytes(48)


def test_no_alloc_replayed_from_a_graph(gpu):
    import torch
    h, m = inputs(300000, 3)
    want = with_workspace(h, m)
    out = torch.zeros(48, dtype=torch.uint8, device="cuda")
    no_alloc(h, m, out)  # first call on the device outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            no_alloc(h, m, out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for k in range(3):
        out.fill_(0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert out.cpu().numpy().tobytes() == want, f"replay {k}"
        # eager work between the replays
        no_alloc(m, h, out)
        torch.cuda.synchronize()
