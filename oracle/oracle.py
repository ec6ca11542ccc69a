"""ctypes front of the CPU restatement (oracle/film_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_film.so")

# The ABI structs are shared with the product library; reuse their layout.
import sys as _sys

_PKG = os.path.join(_HERE, "..", "paint-sim-3d_amd")
if _PKG not in _sys.path:
    _sys.path.insert(0, _PKG)
import cz_types  # noqa: E402

Params = cz_types.Params
Deposit = cz_types.Deposit
Settings = cz_types.Settings
Overrides = cz_types.Overrides
MassReport = cz_types.MassReport

_lib = None


def build():
    """Compile the restatement in place (make -C oracle)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        f32, u32, u64 = C.c_float, C.c_uint32, C.c_uint64
        L.of_create.restype = P
        L.of_create.argtypes = [u32, u32, C.POINTER(Settings), C.POINTER(f32), C.POINTER(f32)]
        L.of_destroy.argtypes = [P]
        L.of_buffer.restype = P
        L.of_buffer.argtypes = [P, C.c_int, C.POINTER(u64)]
        L.of_params.restype = C.POINTER(Params)
        L.of_params.argtypes = [P]
        L.of_settings.restype = C.POINTER(Settings)
        L.of_settings.argtypes = [P]
        L.of_overrides.restype = C.POINTER(Overrides)
        L.of_overrides.argtypes = [P]
        L.of_active_count.restype = u32
        L.of_omp_threads.restype = C.c_int
        L.of_omp_threads.argtypes = []
        L.of_active_count.argtypes = [P]
        L.of_implicit_steps.restype = u32
        L.of_implicit_steps.argtypes = [P]
        L.of_cg_s.restype = C.POINTER(f32)
        L.of_cg_s.argtypes = [P]
        L.of_encode_frame.argtypes = [P, u32, u32, u32, u32]
        L.of_update_params.argtypes = [P, f32, u32]
        L.of_choose_substeps.restype = u32
        L.of_choose_substeps.argtypes = [P, f32]
        L.of_gravity_vec.argtypes = [P, C.POINTER(f32)]
        L.of_upload_deposits.restype = u32
        L.of_upload_deposits.argtypes = [P, C.POINTER(Deposit), u32]
        L.of_step.restype = u32
        L.of_step.argtypes = [P, C.POINTER(Deposit), u32, f32]
        L.of_poll_stats.argtypes = [P]
        L.of_host_state.argtypes = [P, C.POINTER(f32), C.POINTER(C.c_double)]
        L.of_set_host_state.argtypes = [P, f32, f32, f32]
        L.of_mass_report.argtypes = [P, C.POINTER(MassReport)]
        L.of_shade.argtypes = [P, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16)]
        L.of_journal_begin.argtypes = [P, u32, u32]
        L.of_journal_head.restype = u32
        L.of_journal_head.argtypes = [P]
        L.of_journal_touch_rect.argtypes = [P, f32, f32, f32, f32]
        L.of_journal_restore.argtypes = [P, u32]
        L.of_upload_contacts.argtypes = [P, C.POINTER(cz_types.Contact), u32]
        L.of_set_knife.argtypes = [P, C.POINTER(cz_types.Knife)]
        L.of_take_scraps.restype = u32
        L.of_take_scraps.argtypes = [P, C.c_void_p, u32]
        L.of_take_reservoir.argtypes = [P, C.POINTER(u32)]
        L.of_brush_pass.argtypes = [P]
        L.of_knife_pass.argtypes = [P]
        for k in ("of_k_mark_compact", "of_k_laplacian", "of_k_plug", "of_k_fluxes", "of_k_cure", "of_k_nose",
                  "of_k_limiter", "of_k_apply", "of_k_clear_acc", "of_k_cg_setup", "of_k_shed", "of_k_slump",
                  "of_k_commit_tmp"):
            getattr(L, k).argtypes = [P]
        L.of_canvas_reflectance.argtypes = [C.POINTER(f32), C.POINTER(f32), C.POINTER(f32)]
        L.of_fn_shed_excess.restype = f32
        L.of_fn_shed_excess.argtypes = [P, u32]
        L.of_k_strata.argtypes = [P, C.c_int]
        L.of_cg_iterate.argtypes = [P, u32]
        L.of_cg_apply.argtypes = [P, C.POINTER(f32), C.POINTER(f32)]
        L.of_fn_feeders.restype = u32
        L.of_fn_feeders.argtypes = [P, u32]
        L.of_fn_front_normal.argtypes = [P, u32, C.POINTER(f32)]
        L.of_fn_wet_gain.restype = f32
        L.of_fn_wet_gain.argtypes = [f32, f32]
        L.of_fn_front_shape.restype = f32
        L.of_fn_front_shape.argtypes = [P, u32, f32, f32]
        for k in ("of_fn_wet_threshold", "of_fn_cell_excess"):
            getattr(L, k).restype = f32
            getattr(L, k).argtypes = [P, u32]
        L.of_fn_hold_level.restype = f32
        L.of_fn_hold_level.argtypes = [P, f32]
        L.of_fn_cox_speed.restype = f32
        L.of_fn_cox_speed.argtypes = [P, f32, f32]
        L.of_fn_rim_translation.restype = f32
        L.of_fn_rim_translation.argtypes = [P, u32, u32, f32, f32, f32, C.POINTER(f32)]
        for name in ("of_expf", "of_logf", "of_atanf"):
            getattr(L, name).restype = f32
            getattr(L, name).argtypes = [f32]
        L.of_powf.restype = f32
        L.of_powf.argtypes = [f32, f32]
        L.of_pow_fcap.restype = f32
        L.of_pow_fcap.argtypes = [f32, f32]
        L.of_feed_pow.restype = f32
        L.of_feed_pow.argtypes = [u32]
        L.of_atan2f.restype = f32
        L.of_atan2f.argtypes = [f32, f32]
        L.of_f2h.restype = u32
        L.of_f2h.argtypes = [f32]
        L.of_h2f.restype = f32
        L.of_h2f.argtypes = [u32]
        L.of_km_layer.argtypes = [f32, f32, f32, C.POINTER(f32), C.POINTER(f32)]
        L.of_composite.restype = f32
        L.of_composite.argtypes = [f32, f32, f32]
        L.of_layer_ks.argtypes = [P, C.POINTER(f32), u32, C.POINTER(f32), C.POINTER(f32)]
        L.of_hb_flux.restype = f32
        L.of_hb_flux.argtypes = [f32, f32, f32, f32, f32, C.POINTER(f32), C.POINTER(f32)]
        L.of_band_flux.restype = f32
        L.of_band_flux.argtypes = [f32, f32, f32, f32, f32]
        L.of_settings_default.argtypes = [C.POINTER(Settings)]
        L.of_overrides_default.argtypes = [C.POINTER(Overrides)]
        L.of_pigment_table.argtypes = [C.POINTER(f32)] * 6
        L.of_cmf_table.argtypes = [C.POINTER(f32)]
        L.of_weave.argtypes = [u32, u32, f32, f32, C.POINTER(f32)]
        L.of_substrate_refl.argtypes = [u64, u32, C.POINTER(u32)]
        L.of_height_at.restype = f32
        L.of_height_at.argtypes = [C.POINTER(Deposit), f32, f32]
        L.of_reference_height.restype = f32
        L.of_reference_height.argtypes = [C.POINTER(Deposit), f32, f32]
        L.of_sampled_volume.restype = f32
        L.of_sampled_volume.argtypes = [C.POINTER(Deposit), f32]
        L.of_inside_fraction.restype = f32
        L.of_inside_fraction.argtypes = [C.POINTER(Deposit), f32, f32]
        fp2 = C.POINTER(f32)
        L.of_deposit_ribbon.argtypes = [fp2, fp2, fp2, fp2, f32, f32, u32, f32, C.POINTER(u32), f32, C.POINTER(Deposit)]
        L.of_deposit_capsule.argtypes = [fp2, fp2, f32, f32, C.POINTER(u32), f32, C.POINTER(Deposit)]
        L.of_deposit_stroke.argtypes = [fp2, fp2, f32, f32, C.POINTER(u32), f32, C.POINTER(Deposit)]
        L.of_km_mix_f64.argtypes = [C.POINTER(C.c_uint16), fp2, fp2, u64, u32, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# (id, dtype, per-cell shape) of every buffer the tests look at
BUFFERS = {
    "h": (0, np.float32, ()), "m": (1, np.float32, ()), "rheo": (2, np.float32, (2,)),
    "strata_h": (3, np.float32, (4,)), "strata_pig": (4, np.uint32, (4, 8)),
    "strata_cure": (5, np.float32, (4,)), "sub_refl": (6, np.uint32, (8,)), "lay": (7, np.uint32, ()),
    "delta": (16, np.uint32, ()), "fnorm": (17, np.float32, (2,)), "pres": (18, np.float32, (2,)),
    "h_frame": (19, np.float32, ()), "cg_m": (20, np.float32, (4,)), "cg_p": (21, np.float32, ()),
    "tile_moved": (22, np.uint32, None), "tile_flags": (23, np.uint32, None), "tmp": (24, np.float32, ()),
    "flux": (25, np.float32, (4,)), "acc": (26, np.float32, (4,)), "alpha": (27, np.float32, ()),
    "stats": (28, np.float32, None), "pool": (29, np.float32, None), "tile_list": (30, np.uint32, None),
    "acc_s": (31, np.float32, (4, 4)), "alpha_s": (32, np.float32, (4,)),
    "jmeta": (38, np.uint32, None), "jmark": (39, np.uint32, None), "jdata": (40, np.uint32, None),
    "n_cell": (41, np.float32, ()),
    "cg_x": (100, np.float32, ()), "cg_d": (101, np.float32, ()), "cg_r": (102, np.float32, ()),
}


class OracleFilm:
    """The restated Film: same buffers, same frame, on the CPU."""

    def __init__(self, w, h, settings, pigtab, cmf):
        L = lib()
        self.w, self.h = w, h
        pt = np.ascontiguousarray(pigtab, np.float32)
        cm = np.ascontiguousarray(cmf, np.float32)
        self._p = L.of_create(w, h, C.byref(settings), _fp(pt), _fp(cm))
        if not self._p:
            raise MemoryError("of_create failed")

    def close(self):
        if self._p:
            lib().of_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def buf(self, name):
        """numpy view (no copy) of a buffer."""
        which, dt, shape = BUFFERS[name]
        nb = C.c_uint64()
        ptr = lib().of_buffer(self._p, which, C.byref(nb))
        n = nb.value // np.dtype(dt).itemsize
        arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,))
        if shape is None:
            return arr
        return arr.reshape((self.h, self.w) + shape)

    @property
    def params(self):
        return lib().of_params(self._p).contents

    @property
    def settings(self):
        return lib().of_settings(self._p).contents

    @property
    def overrides(self):
        return lib().of_overrides(self._p).contents

    def active_count(self):
        return lib().of_active_count(self._p)

    def implicit_steps(self):
        return lib().of_implicit_steps(self._p)

    def cg_s(self):
        p = lib().of_cg_s(self._p)
        return np.array([p[i] for i in range(8)], np.float32)

    def host_state(self):
        a = (C.c_float * 3)()
        b = (C.c_double * 2)()
        lib().of_host_state(self._p, a, b)
        return {"h_max": a[0], "mu_min": a[1], "h_max_pending": a[2],
                "total_deposited": b[0], "weave_volume": b[1]}

    def set_host_state(self, h_max, mu_min, pending):
        lib().of_set_host_state(self._p, h_max, mu_min, pending)

    def update_params(self, frame_dt, substeps):
        lib().of_update_params(self._p, frame_dt, substeps)

    def choose_substeps(self, frame_dt):
        return lib().of_choose_substeps(self._p, frame_dt)

    def gravity_vec(self):
        g = (C.c_float * 3)()
        lib().of_gravity_vec(self._p, g)
        return [g[0], g[1], g[2]]

    def upload_deposits(self, deps):
        arr, n = cz_types.deposit_array(deps)
        return lib().of_upload_deposits(self._p, arr, n)

    def encode_frame(self, substeps, implicit_steps=None, cg_iters=None, run_implicit=None):
        s = self.settings
        lib().of_encode_frame(self._p, substeps,
                              self.implicit_steps() if implicit_steps is None else implicit_steps,
                              s.implicit_iters if cg_iters is None else cg_iters,
                              s.implicit_hydro if run_implicit is None else run_implicit)

    def shade(self):
        """shade_cell over the grid (shade.wgsl:29-83): (color, detail) as f16
        bits, [h, w, 4] each."""
        color = np.zeros(self.w * self.h * 4, np.uint16)
        detail = np.zeros(self.w * self.h * 4, np.uint16)
        u16p = C.POINTER(C.c_uint16)
        lib().of_shade(self._p, color.ctypes.data_as(u16p), detail.ctypes.data_as(u16p))
        return color.reshape(self.h, self.w, 4), detail.reshape(self.h, self.w, 4)

    def journal_begin(self, gen, cap):
        """Journal.gen = gen (src/session.rs:141); the ring of cap tiles is
        made at the first call."""
        lib().of_journal_begin(self._p, gen, cap)

    def journal_head(self):
        return lib().of_journal_head(self._p)

    def touch_rect(self, minx, miny, maxx, maxy):
        lib().of_journal_touch_rect(self._p, minx, miny, maxx, maxy)

    def restore_journal(self, target):
        lib().of_journal_restore(self._p, target)

    def upload_contacts(self, contacts):
        arr, n = cz_types.contact_array(contacts)
        lib().of_upload_contacts(self._p, arr, n)

    def set_knife(self, knife):
        lib().of_set_knife(self._p, C.byref(knife) if knife is not None else None)

    def brush_pass(self):
        """brush_scrape alone, on the uploaded contacts."""
        lib().of_brush_pass(self._p)

    def knife_pass(self):
        """knife alone, with the set knife."""
        lib().of_knife_pass(self._p)

    def take_reservoir(self):
        a = np.zeros((512, 17), np.uint32)
        lib().of_take_reservoir(self._p, a.ctypes.data_as(C.POINTER(C.c_uint32)))
        return a

    def take_scraps(self, cap=1 << 18):
        a = np.zeros(cap, cz_types.SCRAP_DTYPE)
        n = lib().of_take_scraps(self._p, a.ctypes.data_as(C.c_void_p), cap)
        return a[:min(n, cap)]

    def step(self, deps=(), frame_dt=1.0 / 60.0):
        arr, n = cz_types.deposit_array(deps)
        return lib().of_step(self._p, arr, n, frame_dt)

    def poll_stats(self):
        lib().of_poll_stats(self._p)

    def mass_report(self):
        r = MassReport()
        lib().of_mass_report(self._p, C.byref(r))
        return r
