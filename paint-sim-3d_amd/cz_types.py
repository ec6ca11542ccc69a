"""ctypes mirrors of the POD structs of include/canvaz_film.h.

Params (176 B) and Deposit (96 B) are byte-identical to the reference's
`Params` / `Deposit` (src/film/mod.rs:5-39); Settings mirrors `FilmSettings`
(src/film/mod.rs:281-313).
"""
import ctypes as C

f32, u32 = C.c_float, C.c_uint32


class Params(C.Structure):
    _fields_ = [("size", u32 * 2), ("tiles", u32 * 2), ("dx", f32), ("dt", f32), ("inv_dx", f32),
                ("frame_dt", f32), ("g", f32 * 4), ("phys", f32 * 4), ("ein", f32 * 4), ("cure", f32 * 4),
                ("misc", f32 * 4), ("flags", u32 * 4), ("front", f32 * 4), ("front2", f32 * 4),
                ("front3", f32 * 4)]


class Deposit(C.Structure):
    _fields_ = [("p0", f32 * 2), ("p1", f32 * 2), ("n0", f32 * 2), ("n1", f32 * 2), ("r0", f32), ("r1", f32),
                ("h0", f32), ("vol", f32), ("kind", u32), ("pad", u32 * 3), ("pig", u32 * 8)]


class Contact(C.Structure):
    """`Contact` (src/film/mod.rs:239-248); pad is the kind: 0 bristle, 1 roller, 2 sponge, 3 body."""
    _fields_ = [("p0", f32 * 2), ("p1", f32 * 2), ("r", f32), ("zb", f32), ("bristle", u32), ("pad", u32)]


class Knife(C.Structure):
    """`KnifeParams` (src/film/mod.rs:250-270)."""
    _fields_ = [("a", f32 * 2), ("b", f32 * 2), ("dir", f32 * 2), ("sweep", f32), ("height", f32),
                ("half_width", f32), ("active", u32), ("frame", u32), ("pad1", u32), ("bands", u32),
                ("pad2", u32), ("pad3", u32), ("pad4", u32), ("pig", (u32 * 8) * 8), ("fill", f32 * 8),
                ("smear", f32 * 8)]


class Scrap(C.Structure):
    """`Scrap` (src/film/mod.rs:272-279): vol > 0 taken off the cell, vol < 0 laid on it."""
    _fields_ = [("cell", u32), ("vol", f32), ("frame", u32), ("pad1", u32), ("pig", u32 * 8)]


class Settings(C.Structure):
    _fields_ = [("width_m", f32), ("height_m", f32), ("rho", f32), ("sigma", f32), ("n_index", f32),
                ("m_cap", f32), ("contact_angle_deg", f32), ("hysteresis_deg", f32), ("h_star", f32),
                ("raw_canvas", u32), ("weave_um", f32), ("cure_rate", f32), ("tau_cure", f32),
                ("lnmu_cure", f32), ("v_cap", f32), ("tilt_deg", f32), ("tilt_dir_deg", f32), ("gravity", f32),
                ("medium_mu", f32), ("debug_view", u32), ("h_min", f32), ("m_cap_hydro", f32),
                ("max_substeps", u32), ("min_substeps", u32), ("implicit_iters", u32), ("implicit_hydro", u32),
                ("slump", u32), ("capillary", u32), ("capillary_implicit", u32), ("n_per_cell", u32),
                ("has_canvas_rgb", u32), ("canvas_rgb", f32 * 3)]


class Overrides(C.Structure):
    _fields_ = [("no_implicit", u32), ("no_cap", u32), ("capcond", f32), ("fcap", f32), ("no_hold", u32),
                ("kshape", f32), ("no_transl", u32), ("no_shape", u32), ("kline", f32), ("has_hyst", u32),
                ("hyst_deg", f32)]


class Stats(C.Structure):
    _fields_ = [("h_max", f32), ("mu_min", f32), ("active_tiles", u32), ("substeps_last", u32),
                ("implicit_steps", u32), ("frame", u32), ("cg_report", f32 * 4), ("total_deposited", C.c_double),
                ("weave_volume", C.c_double), ("act_cells", C.c_uint64), ("act_cell_substeps", C.c_uint64),
                ("frames_counted", C.c_uint64), ("act_cell_implicit", C.c_uint64),
                ("one_launch_frames", C.c_uint64), ("one_launch_cells", C.c_uint64),
                ("one_launch_cell_substeps", C.c_uint64), ("flux_shape_cell_substeps", C.c_uint64 * 7)]


class Scene(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nblobs", u32), ("sigma_lo", f32), ("sigma_hi", f32), ("peak_lo", f32),
                ("peak_hi", f32), ("tau_lo", f32), ("tau_hi", f32), ("mu_lo", f32), ("mu_hi", f32),
                ("alpha", f32), ("weave_um", f32), ("n_lo", f32), ("n_hi", f32)]


class Request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("w", u32), ("h", u32), ("steps", u32), ("pad", u32), ("arrival_s", C.c_double),
                ("settings", Settings), ("scene", Scene)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("cell_steps", C.c_uint64), ("status", C.c_int32), ("pad", u32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_members", u32), ("frame_dt", f32), ("keep_films", C.c_int32),
                ("max_wait_s", f32), ("host_budget", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("frames", C.c_uint64), ("admitted", C.c_uint64),
                ("cell_steps", C.c_uint64), ("peak_members", C.c_uint64), ("peak_bytes", C.c_uint64),
                ("films_created", C.c_uint64), ("films_reused", C.c_uint64), ("evictions", C.c_uint64),
                ("resumes", C.c_uint64), ("peak_host_bytes", C.c_uint64)]


class HostState(C.Structure):
    _fields_ = [("h_max", f32), ("mu_min", f32), ("h_max_pending", f32), ("pad", u32)]


class MassTotals(C.Structure):
    _fields_ = [("sum_h", C.c_double), ("sum_m", C.c_double), ("wet_cells", C.c_uint64), ("nan_cells", C.c_uint64),
                ("h_max", f32), ("s_max", f32), ("pad", u32 * 2)]


class MassReport(C.Structure):
    _fields_ = [("wet_ml", C.c_double), ("cured_ml", C.c_double), ("deposited_ml", C.c_double),
                ("err_pct", C.c_double), ("hmax", f32), ("pad", u32), ("wet_cells", C.c_uint64),
                ("nan_cells", C.c_uint64)]


assert C.sizeof(Params) == 176
assert C.sizeof(Deposit) == 96
assert C.sizeof(Contact) == 32
assert C.sizeof(Knife) == 384
assert C.sizeof(Scrap) == 48
assert C.sizeof(MassTotals) == 48

PART_NAMES = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


def default_settings():
    """FilmSettings::default (src/film/mod.rs:315-350)."""
    s = Settings()
    s.width_m, s.height_m, s.rho, s.sigma, s.n_index = 1.2, 0.8, 1000.0, 0.035, 1.0
    s.m_cap, s.contact_angle_deg, s.hysteresis_deg, s.h_star = 1.0e-8, 45.0, 15.0, 2.0e-5
    s.raw_canvas, s.weave_um, s.cure_rate, s.tau_cure, s.lnmu_cure = 1, 25.0, 1.0 / 600.0, 300.0, 5.0
    s.v_cap, s.tilt_deg, s.tilt_dir_deg, s.gravity, s.medium_mu = 1.0, 0.0, -90.0, 9.81, 0.08
    s.debug_view, s.h_min, s.m_cap_hydro = 0, 1.5e-5, 1.0e-7
    s.max_substeps, s.min_substeps, s.implicit_iters, s.implicit_hydro = 64, 4, 40, 1
    s.slump, s.capillary, s.capillary_implicit = 0, 1, 0
    s.n_per_cell = 0
    s.has_canvas_rgb = 0   # canvas_rgb: None
    return s


def default_overrides():
    o = Overrides()
    o.capcond, o.fcap, o.kshape, o.kline = 80.0, 2.5, 1.0, 0.0
    return o


def contact_array(cs):
    """(ctypes array, count) from a sequence of Contact."""
    cs = list(cs)
    arr = (Contact * max(len(cs), 1))()
    for i, c in enumerate(cs):
        arr[i] = c
    return arr, len(cs)


SCRAP_DTYPE = [("cell", "<u4"), ("vol", "<f4"), ("frame", "<u4"), ("pad1", "<u4"), ("pig", "<u4", (8,))]


def deposit_array(deps):
    """(ctypes array, count) from a sequence of Deposit or a numpy structured/bytes buffer."""
    deps = list(deps)
    n = len(deps)
    arr = (Deposit * max(n, 1))()
    for i, d in enumerate(deps):
        arr[i] = d
    return arr, n
