"""Shared set-up of the parity tests: the same seeded state and deposits fed
to the HIP film and to the CPU restatement."""
import ctypes as C

import numpy as np

import canvaz_film as czf
import oracle as orc
import scenes
from cz_types import PART_NAMES

STATE_SHAPES = {"h": (), "m": (), "rheo": (2,), "strata_h": (4,), "strata_pig": (4, 8), "strata_cure": (4,),
                "sub_refl": (8,), "lay": ()}


def oracle_capsule(p0, p1, r, vol, pig, dx):
    d = orc.Deposit()
    orc.lib().of_deposit_capsule((C.c_float * 2)(*p0), (C.c_float * 2)(*p1), r, vol,
                                 (C.c_uint32 * 8)(*[int(v) for v in pig]), dx, C.byref(d))
    return d


def make_pair(w, h, settings, state=None, pigtab=None, cmf=None, gpu=True):
    """(HIP film or None, oracle film) on identical initial state."""
    pt = czf.default_pigment_table() if pigtab is None else pigtab
    cm = czf.cmf_table() if cmf is None else cmf
    o = orc.OracleFilm(w, h, settings, pt, cm)
    g = czf.Film(w, h, settings, pt, cm) if gpu else None
    if state is not None:
        write_oracle_state(o, state)
        if g is not None:
            assert g.write_state_arrays(state)
    return g, o


def write_oracle_state(o, state):
    """write_state on the restatement (src/film/mod.rs:1303-1320)."""
    for k in PART_NAMES:
        o.buf(k)[...] = state[k].reshape(o.buf(k).shape)
    for k in ("pres", "flux", "acc", "alpha", "delta", "tmp", "acc_s", "alpha_s"):
        o.buf(k)[...] = 0
    o.buf("tile_moved")[...] = 1


def rel_err(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor))) if a.size else 0.0
