"""Host-side mirror of the reference's `Film` (src/film/mod.rs:445-1462) over
the C ABI of include/canvaz_film.h (libcanvaz_film.so, gfx950).

Method names follow the reference: Film(w, h, settings) = Film::new,
update_params, choose_substeps, gravity_vec, upload_deposits, encode_frame,
step (= App::encode_sim, src/main.rs:1875-1909), read_state, write_state,
dry_now, clear, plus the mass report of compute_stats (src/main.rs:1679).

There is no CPU fallback: if the HIP library cannot be loaded, every entry
raises.
"""
import ctypes as C
import os

import numpy as np

from cz_types import (Contact, Deposit, HostState, Knife, MassReport, MassTotals, Overrides, Params, Request, Result, Scene,
                      ServerConfig, ServerStats, Settings, Stats, PART_NAMES, SCRAP_DTYPE,
                      contact_array, default_overrides, default_settings, deposit_array)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcanvaz_film.so")

_lib = None


class FilmError(RuntimeError):
    pass


def lib():
    """Load libcanvaz_film.so (raises when it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FilmError(f"{LIB_PATH} is missing: build it with paint-sim-3d_amd/build.py")
    # PyTorch ships its own HIP runtime under the same SONAME as the one the
    # library links (libamdhip64.so.7).  Loaded after the library, torch maps
    # a second runtime, and whichever of the two opens the device second sees
    # no device.  Loading torch first makes the library bind to torch's copy,
    # so one runtime serves the process and device pointers pass between them.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P, f32, u32, u64, i32 = C.c_void_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int
    fp, up = C.POINTER(f32), C.POINTER(u32)
    sig = {
        "cz_last_error": (C.c_char_p, []),
        "cz_version": (C.c_char_p, []),
        "cz_settings_default": (None, [C.POINTER(Settings)]),
        "cz_overrides_default": (None, [C.POINTER(Overrides)]),
        "cz_overrides_from_env": (None, [C.POINTER(Overrides)]),
        "cz_pigment_table": (None, [fp, fp, fp, fp, fp, fp]),
        "cz_default_pigments": (None, [fp, fp, fp, fp, fp]),
        "cz_cmf_table": (None, [fp]),
        "cz_weave": (None, [u32, u32, f32, f32, fp]),
        "cz_substrate_refl": (None, [u64, u32, up]),
        "cz_canvas_reflectance": (None, [fp, fp, fp]),
        "cz_substrate_words": (None, [C.POINTER(Settings), fp, up]),
        "cz_deposit_ribbon": (None, [fp, fp, fp, fp, f32, f32, u32, f32, up, f32, C.POINTER(Deposit)]),
        "cz_deposit_capsule": (None, [fp, fp, f32, f32, up, f32, C.POINTER(Deposit)]),
        "cz_deposit_stroke": (None, [fp, fp, f32, f32, up, f32, C.POINTER(Deposit)]),
        "cz_deposit_laid_by": (None, [C.POINTER(Deposit), u32]),
        "cz_deposit_sampled_volume": (f32, [C.POINTER(Deposit), f32]),
        "cz_deposit_height_at": (f32, [C.POINTER(Deposit), f32, f32]),
        "cz_deposit_inside_fraction": (f32, [C.POINTER(Deposit), f32, f32]),
        "cz_paint_pack": (None, [fp, i32, up]),
        "cz_gravity_vec_host": (None, [C.POINTER(Settings), fp]),
        "cz_update_params_host": (None, [C.POINTER(Settings), C.POINTER(Overrides), u32, u32, f32, u32,
                                         C.POINTER(HostState), C.POINTER(Params), up]),
        "cz_choose_substeps_host": (u32, [C.POINTER(Settings), u32, f32, C.POINTER(HostState)]),
        "cz_film_create": (i32, [u32, u32, C.POINTER(Settings), fp, fp, C.POINTER(P)]),
        "cz_film_destroy": (None, [P]),
        "cz_film_dims": (i32, [P, up, up, fp]),
        "cz_film_set_settings": (i32, [P, C.POINTER(Settings)]),
        "cz_film_get_settings": (i32, [P, C.POINTER(Settings)]),
        "cz_film_set_overrides": (i32, [P, C.POINTER(Overrides)]),
        "cz_film_gravity_vec": (i32, [P, fp]),
        "cz_film_update_params": (i32, [P, f32, u32]),
        "cz_film_choose_substeps": (u32, [P, f32]),
        "cz_film_get_params": (i32, [P, C.POINTER(Params)]),
        "cz_film_set_params": (i32, [P, C.POINTER(Params)]),
        "cz_film_upload_deposits": (i32, [P, C.POINTER(Deposit), u32, up]),
        "cz_film_encode_frame": (i32, [P, u32]),
        "cz_film_step": (i32, [P, C.POINTER(Deposit), u32, f32, up]),
        "cz_film_sync": (i32, [P]),
        "cz_film_stats": (i32, [P, C.POINTER(Stats)]),
        "cz_film_pool": (i32, [P, fp, u64]),
        "cz_film_mass_report": (i32, [P, C.POINTER(MassReport)]),
        "cz_film_state_sizes": (i32, [P, C.POINTER(u64)]),
        "cz_film_read_state": (i32, [P, C.POINTER(C.c_void_p)]),
        "cz_film_write_state": (i32, [P, C.POINTER(C.c_void_p), C.POINTER(u64)]),
        "cz_film_buffer_size": (i32, [P, i32, C.POINTER(u64)]),
        "cz_film_read_buffer": (i32, [P, i32, C.c_void_p, u64]),
        "cz_film_write_buffer": (i32, [P, i32, C.c_void_p, u64]),
        "cz_film_device_ptr": (i32, [P, i32, C.POINTER(C.c_void_p), C.POINTER(u64)]),
        "cz_film_dry_now": (i32, [P]),
        "cz_film_clear": (i32, [P]),
        "cz_film_stream": (C.c_void_p, [P]),
        "cz_film_use_graphs": (i32, [P, i32]),
        "cz_film_graph_counts": (i32, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "cz_f16_selftest": (i32, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "cz_film_substep_launch": (i32, [P, i32]),
        "cz_film_save_canvaz": (i32, [P, C.c_char_p]),
        "cz_film_bin_on_device": (i32, [P, i32]),
        "cz_film_shade_output": (i32, [P, i32]),
        "cz_film_journal_gen": (i32, [P, C.c_uint32]),
        "cz_film_upload_contacts": (i32, [P, C.POINTER(Contact), C.c_uint32]),
        "cz_film_set_knife": (i32, [P, C.POINTER(Knife)]),
        "cz_film_take_reservoir": (i32, [P, C.POINTER(C.c_uint32)]),
        "cz_film_take_scraps": (i32, [P, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
        "cz_film_journal_budget": (i32, [P, C.c_uint64]),
        "cz_film_journal_head": (i32, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "cz_film_touch_rect": (i32, [P, C.c_float, C.c_float, C.c_float, C.c_float]),
        "cz_film_restore_journal": (i32, [P, C.c_uint32]),
        "cz_film_load_canvaz": (i32, [P, C.c_char_p]),
        "cz_lz4_compress": (i32, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
        "cz_lz4_decompress": (i32, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
        "cz_lz4_bound": (C.c_uint64, [C.c_uint64]),
        "cz_film_flux_shape": (i32, [P, i32]),
        "cz_film_profile": (i32, [P, i32]),
        "cz_film_profile_read": (i32, [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(u64)]),
        "cz_film_profile_only": (i32, [P, C.c_char_p]),
        "cz_film_reset_counters": (i32, [P]),
        "cz_scene_default": (None, [C.POINTER(Scene)]),
        "cz_film_seed_blobs": (i32, [P, C.POINTER(Scene)]),
        "cz_film_reshape": (i32, [P, u32, u32, C.POINTER(Settings)]),
        "cz_scene_cells_host": (i32, [u32, u32, f32, C.POINTER(Scene), u32, u32, fp, fp, fp, up]),
        "cz_server_config_default": (None, [C.POINTER(ServerConfig)]),
        "cz_server_create": (i32, [C.POINTER(ServerConfig), C.POINTER(P)]),
        "cz_server_destroy": (None, [P]),
        "cz_server_run": (i32, [P, C.POINTER(Request), u32, C.POINTER(Result), C.POINTER(ServerStats)]),
        "cz_server_film": (P, [P, C.c_uint64]),
        "cz_batch_create": (i32, [u32, C.POINTER(P)]),
        "cz_batch_destroy": (None, [P]),
        "cz_batch_add": (i32, [P, P]),
        "cz_batch_remove": (i32, [P, P]),
        "cz_batch_size": (u32, [P]),
        "cz_batch_step": (i32, [P, f32]),
        "cz_batch_sync": (i32, [P]),
        "cz_batch_use_graphs": (i32, [P, i32]),
        "cz_batch_graph_counts": (i32, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "cz_batch_flux_shape": (i32, [P, i32]),
        "cz_batch_stream": (C.c_void_p, [P]),
        "cz_stream_probe": (i32, [C.c_void_p, C.c_void_p, u64, i32, C.c_void_p]),
        "cz_mass_sums_workspace": (u64, [u64]),
        "cz_mass_sums": (i32, [C.c_void_p, C.c_void_p, u64, C.c_void_p, u64, C.c_void_p, C.c_void_p]),
        "cz_mass_sums_noalloc": (i32, [C.c_void_p, C.c_void_p, u64, C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise FilmError(f"canvaz_film error {rc}: {lib().cz_last_error().decode()}")
    return rc


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


# ---- tables and deposit geometry (host functions of the reference) --------

def default_pigment_table():
    """pigment_table(default_pigments()) (src/color/mod.rs:190-258) -> f32[480]."""
    L = lib()
    refl = np.zeros(12 * 16, np.float32)
    sc, ti, vi, yi = (np.zeros(12, np.float32) for _ in range(4))
    L.cz_default_pigments(_fp(refl), _fp(sc), _fp(ti), _fp(vi), _fp(yi))
    out = np.zeros(480, np.float32)
    L.cz_pigment_table(_fp(refl), _fp(sc), _fp(ti), _fp(vi), _fp(yi), _fp(out))
    return out


def pigment_table(refl, scatter, tint, viscosity, yield_stress):
    out = np.zeros(480, np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (refl, scatter, tint, viscosity, yield_stress)]
    lib().cz_pigment_table(*[_fp(x) for x in a], _fp(out))
    return out


def cmf_table():
    out = np.zeros(48, np.float32)
    lib().cz_cmf_table(_fp(out))
    return out


def paint_pack(frac16, normalize=True):
    f = np.ascontiguousarray(frac16, np.float32)
    out = np.zeros(8, np.uint32)
    lib().cz_paint_pack(_fp(f), 1 if normalize else 0, _up(out))
    return out


def update_params_host(settings, overrides, w, h, frame_dt, substeps, h_max, mu_min, pending):
    """update_params (src/film/mod.rs:869-917) without a device: (Params, implicit_steps)."""
    hs = HostState(h_max, mu_min, pending, 0)
    p = Params()
    ni = C.c_uint32()
    lib().cz_update_params_host(C.byref(settings), C.byref(overrides), w, h, frame_dt, substeps, C.byref(hs),
                                C.byref(p), C.byref(ni))
    return p, ni.value


def choose_substeps_host(settings, w, frame_dt, h_max, mu_min, pending):
    hs = HostState(h_max, mu_min, pending, 0)
    return lib().cz_choose_substeps_host(C.byref(settings), w, frame_dt, C.byref(hs))


def _v2(p):
    return (C.c_float * 2)(float(p[0]), float(p[1]))


def _pig8(pig):
    return (C.c_uint32 * 8)(*[int(v) for v in pig])


def lz4_compress(data):
    """One LZ4 block of `data` with its u32 size prepended (lz4_flex::compress_prepend_size)."""
    src = bytes(data)
    cap = lib().cz_lz4_bound(len(src))
    out = C.create_string_buffer(cap)
    n = C.c_uint64()
    check(lib().cz_lz4_compress(src, len(src), out, cap, C.byref(n)))
    return out.raw[:n.value]


def lz4_decompress(data, cap=None):
    """Inverse of lz4_compress (lz4_flex::decompress_size_prepended)."""
    src = bytes(data)
    if cap is None:
        cap = int.from_bytes(src[:4], "little") if len(src) >= 4 else 0
    out = C.create_string_buffer(max(cap, 1))
    n = C.c_uint64()
    check(lib().cz_lz4_decompress(src, len(src), out, cap, C.byref(n)))
    return out.raw[:n.value]


def deposit_capsule(p0, p1, r, volume, pig, dx):
    d = Deposit()
    lib().cz_deposit_capsule(_v2(p0), _v2(p1), r, volume, _pig8(pig), dx, C.byref(d))
    return d


def deposit_stroke(p0, p1, r, volume, pig, dx):
    d = Deposit()
    lib().cz_deposit_stroke(_v2(p0), _v2(p1), r, volume, _pig8(pig), dx, C.byref(d))
    return d


def deposit_ribbon(p0, p1, n0, n1, r0, r1, caps, volume, pig, dx):
    d = Deposit()
    lib().cz_deposit_ribbon(_v2(p0), _v2(p1), _v2(n0), _v2(n1), r0, r1, caps, volume, _pig8(pig), dx, C.byref(d))
    return d


def laid_by(d, method):
    lib().cz_deposit_laid_by(C.byref(d), method)
    return d


# ---- the film --------------------------------------------------------------

BUFFER_IDS = {
    "h": 0, "m": 1, "rheo": 2, "strata_h": 3, "strata_pig": 4, "strata_cure": 5, "sub_refl": 6, "lay": 7,
    "delta": 16, "fnorm": 17, "pres": 18, "h_frame": 19, "cg_m": 20, "cg_p": 21, "tile_moved": 22,
    "tile_flags": 23, "tmp": 24, "flux": 25, "acc": 26, "alpha": 27, "stats": 28, "pool": 29, "tile_list": 30,
    "acc_s": 31, "alpha_s": 32, "dep_start": 33, "dep_count": 34, "dep_items": 35,
    "color": 36, "detail": 37, "jmeta": 38, "jmark": 39, "jdata": 40, "n_cell": 41,
}
_U32 = {"strata_pig", "sub_refl", "lay", "delta", "tile_moved", "tile_flags", "tile_list", "dep_start", "dep_count",
        "dep_items", "jmeta", "jmark", "jdata"}
_CELL_SHAPE = {"h": (), "m": (), "rheo": (2,), "strata_h": (4,), "strata_pig": (4, 8), "strata_cure": (4,),
               "sub_refl": (8,), "lay": (), "delta": (), "color": (4,), "detail": (4,), "fnorm": (2,), "pres": (2,), "h_frame": (), "n_cell": (), "cg_m": (4,),
               "cg_p": (), "tmp": (), "flux": (4,), "acc": (4,), "alpha": (), "acc_s": (4, 4), "alpha_s": (4,)}


class Film:
    """One canvas on the current HIP device (Film::new, src/film/mod.rs:579)."""

    def __init__(self, w, h, settings=None, pigtab=None, cmf=None):
        L = lib()
        s = settings if settings is not None else default_settings()
        pt = None if pigtab is None else np.ascontiguousarray(pigtab, np.float32)
        cm = None if cmf is None else np.ascontiguousarray(cmf, np.float32)
        ptr = C.c_void_p()
        check(L.cz_film_create(w, h, C.byref(s), _fp(pt) if pt is not None else None,
                               _fp(cm) if cm is not None else None, C.byref(ptr)))
        self._p = ptr
        self.w, self.h = w, h
        d = C.c_float()
        check(L.cz_film_dims(self._p, None, None, C.byref(d)))
        self.dx = d.value  # the FP32 cell size the step uses (width_m / w in f32)

    def close(self):
        if getattr(self, "_p", None):
            lib().cz_film_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def settings(self):
        s = Settings()
        check(lib().cz_film_get_settings(self._p, C.byref(s)))
        return s

    @settings.setter
    def settings(self, s):
        check(lib().cz_film_set_settings(self._p, C.byref(s)))

    def set_overrides(self, o):
        check(lib().cz_film_set_overrides(self._p, C.byref(o)))

    @property
    def params(self):
        p = Params()
        check(lib().cz_film_get_params(self._p, C.byref(p)))
        return p

    def set_params(self, p):
        check(lib().cz_film_set_params(self._p, C.byref(p)))

    def gravity_vec(self):
        g = (C.c_float * 3)()
        check(lib().cz_film_gravity_vec(self._p, g))
        return [g[0], g[1], g[2]]

    def update_params(self, frame_dt, substeps):
        check(lib().cz_film_update_params(self._p, frame_dt, substeps))

    def choose_substeps(self, frame_dt):
        return lib().cz_film_choose_substeps(self._p, frame_dt)

    def upload_deposits(self, deps):
        arr, n = deposit_array(deps)
        sent = C.c_uint32()
        check(lib().cz_film_upload_deposits(self._p, arr, n, C.byref(sent)))
        return sent.value

    def encode_frame(self, substeps):
        check(lib().cz_film_encode_frame(self._p, substeps))

    def step(self, deps=(), frame_dt=1.0 / 60.0):
        arr, n = deposit_array(deps)
        sent = C.c_uint32()
        check(lib().cz_film_step(self._p, arr, n, frame_dt, C.byref(sent)))
        return sent.value

    def sync(self):
        check(lib().cz_film_sync(self._p))

    def stats(self):
        st = Stats()
        check(lib().cz_film_stats(self._p, C.byref(st)))
        return st

    def mass_report(self):
        r = MassReport()
        check(lib().cz_film_mass_report(self._p, C.byref(r)))
        return r

    def buffer(self, name):
        """Host copy of a state part or solver buffer, shaped [H, W, ...] for per-cell ones."""
        which = BUFFER_IDS[name]
        nb = C.c_uint64()
        check(lib().cz_film_buffer_size(self._p, which, C.byref(nb)))
        dt = np.uint16 if name in ("color", "detail") else (np.uint32 if name in _U32 else np.float32)
        a = np.empty(nb.value // np.dtype(dt).itemsize, dt)
        check(lib().cz_film_read_buffer(self._p, which, a.ctypes.data_as(C.c_void_p), nb.value))
        if name in _CELL_SHAPE:
            return a.reshape((self.h, self.w) + _CELL_SHAPE[name])
        return a

    def write_buffer(self, name, arr):
        which = BUFFER_IDS[name]
        a = np.ascontiguousarray(arr)
        check(lib().cz_film_write_buffer(self._p, which, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def state_sizes(self):
        sz = (C.c_uint64 * 8)()
        check(lib().cz_film_state_sizes(self._p, sz))
        return list(sz)

    def read_state(self):
        """read_state (src/film/mod.rs:1281): 8 byte strings in state_buffers order."""
        sizes = self.state_sizes()
        bufs = [np.empty(s, np.uint8) for s in sizes]
        ptrs = (C.c_void_p * 8)(*[b.ctypes.data for b in bufs])
        check(lib().cz_film_read_state(self._p, ptrs))
        return [b.tobytes() for b in bufs]

    def write_state(self, parts):
        """write_state (src/film/mod.rs:1303); returns False on a size mismatch."""
        if len(parts) != 8:
            return False
        arrs = [np.frombuffer(p, np.uint8) if isinstance(p, (bytes, bytearray)) else
                np.ascontiguousarray(p).view(np.uint8).reshape(-1) for p in parts]
        ptrs = (C.c_void_p * 8)(*[a.ctypes.data for a in arrs])
        sizes = (C.c_uint64 * 8)(*[a.nbytes for a in arrs])
        rc = lib().cz_film_write_state(self._p, ptrs, sizes)
        if rc == -2:
            return False
        check(rc)
        return True

    def write_state_arrays(self, state):
        """write_state from a dict of per-part numpy arrays (names of PART_NAMES)."""
        return self.write_state([np.ascontiguousarray(state[k]) for k in PART_NAMES])

    def reshape(self, w, h, settings):
        """Re-dimension within the created capacity (a fresh film of w x h)."""
        check(lib().cz_film_reshape(self._p, w, h, C.byref(settings)))
        self.w, self.h = w, h
        d = C.c_float()
        check(lib().cz_film_dims(self._p, None, None, C.byref(d)))
        self.dx = d.value

    def flux_shape(self, shape):
        """Substep launch shape: -1 = tuned online, 0 = LDS window per tile,
        1 = block per (tile, face), 2 = flux + apply fused, 3 and 4 = 0 and 2
        with 1024 threads per tile, 5 = the whole substep in one launch, 6 =
        the flux of 1 between the 1024-thread lap/plug and apply of 3."""
        check(lib().cz_film_flux_shape(self._p, shape))

    def pool(self):
        """pool8 of the last frame: (w/4)*(h/4) surface maxima, then as many relief maxima."""
        n = 2 * ((self.w + 3) // 4) * ((self.h + 3) // 4)
        a = np.empty(n, np.float32)
        check(lib().cz_film_pool(self._p, _fp(a), n))
        return a

    def shade_output(self, on=True):
        """Run the shade pass at the end of every frame (buffers "color", "detail": f16 bits)."""
        check(lib().cz_film_shade_output(self._p, 1 if on else 0))

    # ---- tool contacts and the knife (src/film/mod.rs:1030-1173) ----
    def upload_contacts(self, contacts):
        """upload_contacts: Contact segments for the next frame's brush pass."""
        arr, n = contact_array(contacts)
        check(lib().cz_film_upload_contacts(self._p, arr, n))

    def set_knife(self, knife):
        """set_knife: a Knife (KnifeParams) for the following frames, or None."""
        check(lib().cz_film_set_knife(self._p, C.byref(knife) if knife is not None else None))

    def take_reservoir(self):
        """Paint the brush pass took since the last call, [512 bristles][17] u32
        (volume, 16 pigment volumes; units of 1e-15 m^3); the sums restart."""
        a = np.zeros((512, 17), np.uint32)
        check(lib().cz_film_take_reservoir(self._p, a.ctypes.data_as(C.POINTER(C.c_uint32))))
        return a

    def take_scraps(self, cap=1 << 18):
        """The knife's scraps since the last call (structured array, ascending
        cell order, taken before laid paint of a cell); the list restarts."""
        a = np.zeros(cap, SCRAP_DTYPE)
        n = C.c_uint32()
        check(lib().cz_film_take_scraps(self._p, a.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return a[:min(n.value, cap)]

    # ---- undo journal (src/film/mod.rs:357-443, 1246-1275) ----
    def journal_gen(self, gen):
        """Journal.gen (src/session.rs:139-143): 0 = off."""
        check(lib().cz_film_journal_gen(self._p, gen))

    def journal_budget(self, nbytes):
        """PAINT_UNDO_MB in bytes, before the first nonzero generation."""
        check(lib().cz_film_journal_budget(self._p, nbytes))

    def journal_head(self):
        """(head, cap): the next record's sequence number (Journal::head) and the ring's size in tiles."""
        h, c = C.c_uint32(), C.c_uint32()
        check(lib().cz_film_journal_head(self._p, C.byref(h), C.byref(c)))
        return h.value, c.value

    def touch_rect(self, minx, miny, maxx, maxy):
        """Film::touch_rect: journal the tiles of this rectangle (metres) in the next frame."""
        check(lib().cz_film_touch_rect(self._p, minx, miny, maxx, maxy))

    def restore_journal(self, target):
        """Film::restore_journal: tiles with records of generation >= target go back to the oldest."""
        check(lib().cz_film_restore_journal(self._p, target))

    def bin_on_device(self, mode):
        """Deposit binning: -1 by deposit count, 0 on the host, 1 on the device."""
        check(lib().cz_film_bin_on_device(self._p, mode))

    def save_canvaz(self, path):
        """Write the film as a painting file (save_painting, src/session.rs:248-280)."""
        check(lib().cz_film_save_canvaz(self._p, os.fsencode(path)))

    def load_canvaz(self, path):
        """Read a painting file of this film's size (open_painting, src/session.rs:282-317)."""
        check(lib().cz_film_load_canvaz(self._p, os.fsencode(path)))

    def substep_launch(self, mode):
        """-1 = by active-set size, 0 = three launches per substep, 1 = one launch per frame."""
        check(lib().cz_film_substep_launch(self._p, mode))

    def use_graphs(self, on=True):
        """Replay frames from captured hipGraphs (the default) or launch directly."""
        check(lib().cz_film_use_graphs(self._p, 1 if on else 0))

    def graph_counts(self):
        """(graphs captured, frames replayed from them)."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        check(lib().cz_film_graph_counts(self._p, C.byref(a), C.byref(b)))
        return a.value, b.value

    def seed_blobs(self, scene):
        """Replace the state with a seeded blob canvas generated on the device."""
        check(lib().cz_film_seed_blobs(self._p, C.byref(scene)))

    def dry_now(self):
        check(lib().cz_film_dry_now(self._p))

    def clear(self):
        check(lib().cz_film_clear(self._p))

    def device_ptr(self, name):
        p = C.c_void_p()
        nb = C.c_uint64()
        check(lib().cz_film_device_ptr(self._p, BUFFER_IDS[name], C.byref(p), C.byref(nb)))
        return p.value, nb.value

    def stream(self):
        return lib().cz_film_stream(self._p)

    def profile(self, on=True):
        check(lib().cz_film_profile(self._p, 1 if on else 0))

    def profile_only(self, label=None):
        check(lib().cz_film_profile_only(self._p, label.encode() if label else None))

    def reset_counters(self):
        check(lib().cz_film_reset_counters(self._p))

    def profile_read(self, label):
        avg = C.c_double()
        cnt = C.c_uint64()
        check(lib().cz_film_profile_read(self._p, label.encode(), C.byref(avg), C.byref(cnt)))
        return avg.value, cnt.value


class UndoHistory:
    """The undo levels of a painting session (src/session.rs:9-51, 139-211).
    A stroke level keeps the journal head at its start and is undone with
    restore_journal; a full level keeps a copy of the state parts and is
    undone with write_state.  Every level and every undo starts a new
    journal generation."""

    MAX_LEVELS = 64
    MAX_FULL = 3

    def __init__(self, film):
        self.film = film
        self.gen = 0
        self.levels = []  # (gen, "journal", start head) or (gen, "full", state parts)

    def _next_gen(self):
        self.gen += 1
        self.film.journal_gen(self.gen)
        return self.gen

    def _trim(self):
        while len(self.levels) > self.MAX_LEVELS:
            self.levels.pop(0)
        fulls = [i for i, lv in enumerate(self.levels) if lv[1] == "full"]
        if len(fulls) > self.MAX_FULL:
            del self.levels[:fulls[len(fulls) - self.MAX_FULL - 1] + 1]

    def begin_stroke(self):
        start, _ = self.film.journal_head()
        gen = self._next_gen()
        self.levels.append((gen, "journal", start))
        self._trim()

    def begin_full(self):
        parts = self.film.read_state()
        gen = self._next_gen()
        self.levels.append((gen, "full", parts))
        self._trim()

    def can_undo(self):
        return bool(self.levels)

    def rollback(self):
        """Undo the newest level (App::rollback); FilmError with the
        reference's message when it cannot."""
        if not self.levels:
            raise FilmError("Nothing to undo")
        gen, kind, val = self.levels.pop()
        if kind == "journal":
            head, cap = self.film.journal_head()
            if (head - val) & 0xFFFFFFFF > cap:
                self.levels = [lv for lv in self.levels if lv[1] == "full"]
                raise FilmError("That stroke changed too much of the canvas to undo")
            self.film.restore_journal(gen)
        elif not self.film.write_state(val):
            raise FilmError("Undo failed")
        self._next_gen()


class Batch:
    """Films stepped together, each as `Film.step` would step it alone
    without deposits, with one launch per kernel phase over all members."""

    def __init__(self, max_members=256):
        ptr = C.c_void_p()
        check(lib().cz_batch_create(max_members, C.byref(ptr)))
        self._p = ptr
        self.members = []

    def add(self, film):
        check(lib().cz_batch_add(self._p, film._p))
        self.members.append(film)

    def remove(self, film):
        check(lib().cz_batch_remove(self._p, film._p))
        self.members.remove(film)

    def __len__(self):
        return lib().cz_batch_size(self._p)

    def step(self, frame_dt=1.0 / 60.0):
        check(lib().cz_batch_step(self._p, frame_dt))

    def sync(self):
        check(lib().cz_batch_sync(self._p))

    def use_graphs(self, on=True):
        check(lib().cz_batch_use_graphs(self._p, 1 if on else 0))

    def graph_counts(self):
        """(graphs captured, frames replayed from them)."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        check(lib().cz_batch_graph_counts(self._p, C.byref(a), C.byref(b)))
        return a.value, b.value

    def flux_shape(self, shape):
        check(lib().cz_batch_flux_shape(self._p, shape))

    def stream(self):
        return lib().cz_batch_stream(self._p)

    def close(self):
        if getattr(self, "_p", None):
            lib().cz_batch_destroy(self._p)
            self._p = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mass_sums(h, m, workspace=None):
    """cz_mass_sums over two device tensors (float32, same number of cells) on
    torch's current stream: the FP64 totals of compute_stats's loop
    (src/main.rs:1708-1720).  workspace: a uint8 device tensor of at least
    cz_mass_sums_workspace(n) bytes (allocated here when None)."""
    import torch
    n = h.numel()
    assert m.numel() == n and h.dtype == torch.float32 and m.dtype == torch.float32 and h.is_cuda and m.is_cuda
    need = lib().cz_mass_sums_workspace(n)
    if workspace is None:
        workspace = torch.empty(max(need, 16), dtype=torch.uint8, device=h.device)
    out = torch.empty(48, dtype=torch.uint8, device=h.device)
    st = torch.cuda.current_stream().cuda_stream
    check(lib().cz_mass_sums(h.data_ptr(), m.data_ptr(), n, workspace.data_ptr(), workspace.numel(), out.data_ptr(), st))
    return MassTotals.from_buffer_copy(out.cpu().numpy().tobytes())


def scene(seed=1, **kw):
    """cz_scene with the plain workload's defaults, fields overridden by kw."""
    sc = Scene()
    lib().cz_scene_default(C.byref(sc))
    sc.seed = seed
    for k, v in kw.items():
        setattr(sc, k, v)
    return sc


def scene_cells_host(w, h, dx, sc, j0=0, j1=None):
    """Host twin of the device generator for rows [j0, j1): (h, m, rheo, pig0)."""
    j1 = h if j1 is None else j1
    rows = j1 - j0
    hv = np.zeros((rows, w), np.float32)
    mv = np.zeros((rows, w), np.float32)
    rv = np.zeros((rows, w, 2), np.float32)
    pv = np.zeros((rows, w, 8), np.uint32)
    check(lib().cz_scene_cells_host(w, h, dx, C.byref(sc), j0, j1, _fp(hv), _fp(mv), _fp(rv), _up(pv)))
    return hv, mv, rv, pv


class _Borrowed(Film):
    """A film owned by the library (a server's kept canvas): readable, not destroyed here."""

    def __init__(self, ptr, w, h):
        self._p = C.c_void_p(ptr)
        self.w, self.h = w, h
        d = C.c_float()
        check(lib().cz_film_dims(self._p, None, None, C.byref(d)))
        self.dx = d.value

    def close(self):
        self._p = None


class Server:
    """The serving loop: requests of seeded canvases, admitted under a
    device-memory budget (first come first served with max_wait_s = 0, else
    least work first with aged requests first, evicting long-running canvases
    to host memory when an aged one does not fit), stepped together, retired
    when done."""

    def __init__(self, hbm_budget=200 << 30, max_members=256, frame_dt=1.0 / 60.0, keep_films=False, max_wait_s=0.0,
                 host_budget=None):
        cfg = ServerConfig()
        lib().cz_server_config_default(C.byref(cfg))
        cfg.hbm_budget, cfg.max_members, cfg.frame_dt, cfg.keep_films = hbm_budget, max_members, frame_dt, int(keep_films)
        cfg.max_wait_s = max_wait_s
        if host_budget is not None:
            cfg.host_budget = host_budget
        ptr = C.c_void_p()
        check(lib().cz_server_create(C.byref(cfg), C.byref(ptr)))
        self._p = ptr

    def run(self, requests):
        """requests: sequence of Request in arrival order -> (results array, ServerStats)."""
        n = len(requests)
        reqs = (Request * max(n, 1))(*requests)
        res = (Result * max(n, 1))()
        st = ServerStats()
        check(lib().cz_server_run(self._p, reqs, n, res, C.byref(st)))
        return list(res)[:n], st

    def film(self, rid, w, h):
        ptr = lib().cz_server_film(self._p, rid)
        return _Borrowed(ptr, w, h) if ptr else None

    def close(self):
        if getattr(self, "_p", None):
            lib().cz_server_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request(rid, w, h, steps, arrival_s, settings, sc):
    r = Request()
    r.id, r.w, r.h, r.steps, r.arrival_s = rid, w, h, steps, arrival_s
    r.settings = settings
    r.scene = sc
    return r
