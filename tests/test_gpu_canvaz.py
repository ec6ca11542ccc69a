"""Painting files (.canvaz, src/session.rs:59-99,248-317): a film saved and
opened again holds the same state bit for bit, the header carries the
settings and scalars open_painting restores, the file's framing decodes part
by part, and films opened from the same file step identically."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu

PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


def stepped(w, h, tilt, frames):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    s.height_m = h * s.width_m / w
    f = czf.Film(w, h, s)
    assert f.write_state_arrays(scenes.blobs_state(w, h, seed=11, nblobs=9, sigma_cells=(2.0, 9.0)))
    pig = scenes.jet_paint()
    for k in range(frames):
        deps = []
        if k == 2:
            deps = [czf.deposit_capsule((0.3 * w * f.dx, 0.4 * h * f.dx), (0.5 * w * f.dx, 0.4 * h * f.dx),
                                        3 * f.dx, 4e-9, pig, f.dx)]
        f.step(deps)
    f.sync()
    return f


def test_save_open_round_trip(gpu, tmp_path):
    w, h = 160, 128
    a = stepped(w, h, 30.0, 8)
    path = tmp_path / "a.canvaz"
    a.save_canvaz(path)
    sa = a.stats()

    # header (src/session.rs:256-265)
    raw = path.read_bytes()
    assert raw[:4] == b"PSIM"
    u = np.frombuffer(raw[4:16], "<u4")
    assert list(u) == [1, w, h]
    hdr = np.frombuffer(raw[16:36], "<f4")
    assert hdr[0] == np.float32(30.0) and hdr[1] == np.float32(a.settings.tilt_dir_deg)
    assert hdr[2] == np.float32(a.settings.cure_rate) and hdr[3] == np.float32(sa.h_max)
    pool = a.pool()
    np_ = pool.size // 2
    assert hdr[4] == np.float32(max(0.0, float(np.max(pool[:np_][np.isfinite(pool[:np_])]))))
    assert np.frombuffer(raw[36:44], "<f8")[0] == sa.total_deposited

    # pack framing (src/session.rs:59-72): 8 parts, each the state part's bytes
    off = 44
    count = int.from_bytes(raw[off:off + 4], "little")
    off += 4
    assert count == len(PARTS)
    for name in PARTS:
        want = a.buffer(name).tobytes()
        n = int.from_bytes(raw[off:off + 8], "little")
        nch = int.from_bytes(raw[off + 8:off + 12], "little")
        off += 12
        assert n == len(want) and nch == (n + (32 << 20) - 1) // (32 << 20)
        got = b""
        for _ in range(nch):
            cl = int.from_bytes(raw[off:off + 8], "little")
            got += czf.lz4_decompress(raw[off + 8:off + 8 + cl])
            off += 8 + cl
        assert got == want, name
    assert off == len(raw)

    # open into a film with other settings
    s2 = scenes.blobs_settings(w, tilt_deg=5.0)
    s2.height_m = h * s2.width_m / w
    b = czf.Film(w, h, s2)
    b.load_canvaz(path)
    b.sync()
    for name in PARTS:
        assert np.array_equal(a.buffer(name).view(np.uint32), b.buffer(name).view(np.uint32)), name
    sb = b.stats()
    assert sb.h_max == sa.h_max and sb.total_deposited == sa.total_deposited
    assert b.settings.tilt_deg == np.float32(30.0)

    # two films opened from the file step identically
    c = czf.Film(w, h, s2)
    c.load_canvaz(path)
    for _ in range(6):
        b.step()
        c.step()
    b.sync()
    c.sync()
    for name in PARTS + ("delta", "tile_flags"):
        assert np.array_equal(b.buffer(name).view(np.uint32), c.buffer(name).view(np.uint32)), name


def test_open_refuses_bad_files(gpu, tmp_path):
    a = stepped(96, 80, 10.0, 2)
    path = tmp_path / "a.canvaz"
    a.save_canvaz(path)
    other = czf.Film(80, 96, scenes.blobs_settings(80))
    with pytest.raises(czf.FilmError):
        other.load_canvaz(path)                      # another canvas size
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad.canvaz"
    bad.write_bytes(bytes(raw[:-100]))               # truncated
    with pytest.raises(czf.FilmError):
        a.load_canvaz(bad)
    raw[0] = ord("X")                                # not a painting
    bad.write_bytes(bytes(raw))
    with pytest.raises(czf.FilmError):
        a.load_canvaz(bad)
    with pytest.raises(czf.FilmError):
        a.load_canvaz(tmp_path / "missing.canvaz")
