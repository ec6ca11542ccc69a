"""Painting-file framing: the LZ4 block codec of the .canvaz format
(lz4_flex::compress_prepend_size / decompress_size_prepended, as used by
src/session.rs:59-99).  lz4_flex (0.11.6 in the reference's Cargo.lock) is not
vendored and no .canvaz file ships with the reference, so the codec is pinned
by the published LZ4 block format: hand-assembled blocks that exercise each
rule decode to their known bytes, and the compressor's blocks satisfy the
format's end-of-block rules and decode, through an independent decoder written
here from the format description, to their input."""

import numpy as np
import pytest

import canvaz_film as czf


def ref_decode(block, size):
    """LZ4 block decoder from the format description; also checks the
    end-of-block rules the format puts on an encoder's output."""
    out = bytearray()
    i = 0
    last_match_start = last_match_end = None
    while True:
        tok = block[i]
        i += 1
        nlit = tok >> 4
        if nlit == 15:
            while True:
                b = block[i]
                i += 1
                nlit += b
                if b != 255:
                    break
        out += block[i:i + nlit]
        i += nlit
        if i == len(block):
            break
        off = block[i] | (block[i + 1] << 8)
        i += 2
        m = tok & 15
        if m == 15:
            while True:
                b = block[i]
                i += 1
                m += b
                if b != 255:
                    break
        m += 4
        assert 0 < off <= len(out)
        last_match_start = len(out)
        for _ in range(m):
            out.append(out[-off])
        last_match_end = len(out)
    assert len(out) == size
    if last_match_start is not None:
        assert last_match_start <= size - 12, "the last match must start 12 bytes before the end"
        assert last_match_end <= size - 5, "the last 5 bytes must be literals"
    return bytes(out)


def framed(size, block):
    return size.to_bytes(4, "little") + bytes(block)


SPEC_VECTORS = [
    # empty input: one token with no literals
    (framed(0, [0x00]), b""),
    # three literals
    (framed(3, [0x30]) + b"abc", b"abc"),
    # one literal, a match of 9 at offset 1 (overlapping copy), 5 closing literals
    (framed(15, [0x15, ord("x"), 1, 0, 0x50]) + b"abcde", b"x" * 10 + b"abcde"),
    # 300 literals: length 15 + 255 + 30
    (framed(300, [0xF0, 255, 30]) + bytes(range(256)) + bytes(range(44)), bytes(range(256)) + bytes(range(44))),
    # 2 literals, a match of 500 at offset 2 (4 + 15 + 255 + 226), 5 closing literals
    (framed(507, [0x2F]) + b"ab" + bytes([2, 0, 255, 226, 0x50]) + b"vwxyz", b"ab" * 251 + b"vwxyz"),
]


@pytest.mark.parametrize("k", range(len(SPEC_VECTORS)))
def test_decoder_on_format_vectors(k):
    block, want = SPEC_VECTORS[k]
    assert czf.lz4_decompress(block) == want
    assert ref_decode(block[4:], len(want)) == want


def test_decoder_rejects_corrupt_blocks():
    good = framed(15, [0x15, ord("x"), 1, 0, 0x50]) + b"abcde"
    bad = [
        framed(15, [0x15, ord("x"), 2, 0, 0x50]) + b"abcde",  # offset beyond the output
        framed(15, [0x15, ord("x"), 0, 0, 0x50]) + b"abcde",  # offset 0
        good[:-2],                                            # truncated literals
        framed(16, [0x15, ord("x"), 1, 0, 0x50]) + b"abcde",  # size mismatch
        framed(15, [0xF0]),                                   # missing length byte
    ]
    for b in bad:
        with pytest.raises(czf.FilmError):
            czf.lz4_decompress(b)
    with pytest.raises(czf.FilmError):
        czf.lz4_decompress(good, cap=10)  # output does not fit


def inputs():
    rng = np.random.default_rng(7)
    yield b""
    for n in range(1, 40):
        yield bytes(rng.integers(0, 3, n, dtype=np.uint8))
    yield bytes(1 << 16)
    yield rng.bytes(100_000)
    f = np.zeros(300_000, np.float32)
    f[50_000:120_000] = rng.random(70_000, np.float32)
    yield f.tobytes()
    pat = rng.bytes(70_000)          # repeats 70 000 bytes apart: beyond the 64 KiB window
    yield pat + pat + pat[:1000]
    yield (b"paint" * 20_000) + rng.bytes(10)


INPUTS = list(inputs())


@pytest.mark.parametrize("k", range(len(INPUTS)))
def test_round_trip_and_format_rules(k):
    data = INPUTS[k]
    z = czf.lz4_compress(data)
    assert int.from_bytes(z[:4], "little") == len(data)
    assert len(z) <= czf.lib().cz_lz4_bound(len(data))
    assert ref_decode(z[4:], len(data)) == data
    assert czf.lz4_decompress(z) == data


def test_compressible_data_compresses():
    z = czf.lz4_compress(bytes(8 << 20))
    assert len(z) < (8 << 20) // 200
