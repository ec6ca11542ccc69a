"""The device's f16 conversions (used by every kernel that packs or unpacks
pigment slots) give the bits of the host restatement's software conversions,
which test_oracle_math pins to IEEE round-to-nearest-even: checked on the
device over all 2^32 f32 and 2^16 f16 bit patterns, NaN inputs excluded."""
import ctypes as C

import pytest

import canvaz_film as czf

pytestmark = pytest.mark.gpu


def test_hardware_f16_matches_software(gpu):
    a, b = C.c_uint64(99), C.c_uint64(99)
    czf.check(czf.lib().cz_f16_selftest(C.byref(a), C.byref(b)))
    assert (a.value, b.value) == (0, 0)
