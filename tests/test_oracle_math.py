"""Pins of the FP32 elementary functions and f16 conversions (czmath.h) used
identically by the restatement and the device build: error bounds against
float64 numpy, special values, and bit-exact round-to-nearest-even f16 against
numpy's float16 conversion."""
import numpy as np
import pytest

import oracle as orc

L = orc.lib()


def ulp_err(got, ref):
    got = np.asarray(got, np.float32).astype(np.float64)
    ref = np.asarray(ref, np.float64)
    spacing = np.abs(np.spacing(ref.astype(np.float32)).astype(np.float64))
    spacing = np.maximum(spacing, np.float64(np.finfo(np.float32).smallest_subnormal))
    return np.abs(got - ref) / spacing


def vec(fn, xs):
    return np.array([fn(float(x)) for x in xs], np.float32)


def test_exp_bounds():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-87, 88.7, 20000), rng.uniform(-1, 1, 5000), rng.uniform(-103, -87, 2000)])
    xs = xs.astype(np.float32)
    got = vec(L.of_expf, xs)
    ref = np.exp(xs.astype(np.float64))
    normal = ref >= np.finfo(np.float32).tiny
    assert ulp_err(got[normal], ref[normal]).max() <= 2.0
    assert np.abs(got[~normal] - ref[~normal]).max() <= 2 * np.finfo(np.float32).smallest_subnormal
    assert L.of_expf(0.0) == 1.0
    assert L.of_expf(float("-inf")) == 0.0
    assert L.of_expf(float("inf")) == float("inf")
    assert L.of_expf(-200.0) == 0.0
    assert np.isnan(L.of_expf(float("nan")))


def test_log_bounds():
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), rng.uniform(0.5, 2.0, 5000),
                         rng.uniform(1e-44, 1e-38, 1000)]).astype(np.float32)
    xs = xs[xs > 0]
    got = vec(L.of_logf, xs)
    ref = np.log(xs.astype(np.float64))
    e = ulp_err(got, ref)
    assert e.max() <= 2.0, xs[np.argmax(e)]
    assert L.of_logf(1.0) == 0.0
    assert L.of_logf(0.0) == float("-inf")
    assert np.isnan(L.of_logf(-1.0))


def test_pow_matches_definition():
    """pow(x, y) = exp(y log x) (the WGSL definition); error grows with |y log x|."""
    rng = np.random.default_rng(3)
    x = np.exp(rng.uniform(-20, 10, 5000)).astype(np.float32)
    y = rng.uniform(-3, 3, 5000).astype(np.float32)
    got = np.array([L.of_powf(float(a), float(b)) for a, b in zip(x, y)], np.float64)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    rel = np.abs(got - ref) / ref
    bound = 4e-7 * (1.0 + np.abs(y.astype(np.float64) * np.log(x.astype(np.float64))))
    assert np.all(rel <= bound)
    assert L.of_powf(0.0, 2.5) == 0.0
    assert L.of_powf(1.0, 7.0) == 1.0
    assert L.of_powf(4.0, 0.5) == pytest.approx(2.0, rel=2e-7)


def test_pow_fcap():
    """x^2.5 as x^2 sqrt(x): within 2 ULP (three roundings) of the float64 power on [0, 1] (the film's
    thinness ratio) and beyond; any other exponent is cz_powf, bit for bit."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0, 1, 20000), np.exp(rng.uniform(-30, 5, 5000))]).astype(np.float32)
    got = np.array([L.of_pow_fcap(float(a), 2.5) for a in x], np.float32)
    ref = np.power(x.astype(np.float64), 2.5)
    normal = ref >= np.finfo(np.float32).tiny
    assert ulp_err(got[normal], ref[normal]).max() <= 2.0
    assert L.of_pow_fcap(0.0, 2.5) == 0.0 and L.of_pow_fcap(1.0, 2.5) == 1.0
    for e in (2.0, 3.0, 1.7):
        for a in x[:200]:
            assert L.of_pow_fcap(float(a), e) == L.of_powf(float(a), e)


def test_feeder_power_table():
    """max(k, 1)^0.7 for the feeder counts 0..4: the table holds exactly what cz_powf
    returns, and those values are the correctly rounded powers or one ULP off."""
    for k in range(5):
        want = L.of_powf(float(max(k, 1)), 0.7)
        assert L.of_feed_pow(k) == want
        assert ulp_err([want], [max(k, 1) ** float(np.float32(0.7))])[0] <= 1.0


def test_atan_atan2():
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(-5, 5, 5000), np.exp(rng.uniform(-20, 20, 2000))]).astype(np.float32)
    got = vec(L.of_atanf, xs)
    assert ulp_err(got, np.arctan(xs.astype(np.float64))).max() <= 3.0
    ang = rng.uniform(-np.pi, np.pi, 3000)
    ys, xs2 = np.sin(ang).astype(np.float32), np.cos(ang).astype(np.float32)
    got2 = np.array([L.of_atan2f(float(a), float(b)) for a, b in zip(ys, xs2)], np.float64)
    ref2 = np.arctan2(ys.astype(np.float64), xs2.astype(np.float64))
    assert np.abs(got2 - ref2).max() <= 4e-7
    assert L.of_atan2f(0.0, -1.0) == pytest.approx(np.pi, abs=1e-6)
    assert L.of_atan2f(1.0, 0.0) == pytest.approx(np.pi / 2, abs=1e-6)
    assert L.of_atan2f(0.0, 0.0) == 0.0


def test_f16_roundtrip_all_halves():
    for h in range(0, 0x10000):
        e = (h >> 10) & 0x1F
        if e == 31 and (h & 0x3FF):
            continue  # NaN payloads
        f = L.of_h2f(h)
        assert np.float32(f) == np.uint16(h).view(np.float16).astype(np.float32)
        assert L.of_f2h(f) == h


def test_f16_rne_matches_numpy():
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2 ** 32, 200000, dtype=np.uint64).astype(np.uint32)
    xs = bits.view(np.float32)
    xs = xs[np.isfinite(xs)]
    # add the rounding ties and the subnormal / overflow edges
    halves = np.arange(0, 0x7C00, dtype=np.uint16).view(np.float16).astype(np.float32)
    ties = ((halves[:-1].astype(np.float64) + halves[1:].astype(np.float64)) / 2).astype(np.float32)
    xs = np.concatenate([xs, ties, -ties, np.float32([65504.0, 65519.0, 65520.0, 1e-8, 2.0 ** -25, 2.0 ** -24])])
    got = np.array([L.of_f2h(float(x)) for x in xs], np.uint32)
    ref = xs.astype(np.float16).view(np.uint16).astype(np.uint32)
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, (xs[bad[:5]], got[bad[:5]], ref[bad[:5]])
