"""ctypes binding of ``libgpuocean_amd.so`` (the C ABI in include/gpuocean_amd.h).

The library is built ahead of time (``python -m gpuocean_amd.build``).  There is
no fallback: if the library is missing or cannot be loaded, every simulator
constructor raises.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpuocean_amd.so")

c_int32 = ctypes.c_int32
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
c_uint32 = ctypes.c_uint32
c_fp = ctypes.POINTER(ctypes.c_float)


class CDKLM16Params(ctypes.Structure):
    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("rk_order", c_int32), ("one_dimensional", c_int32),
        ("has_land", c_int32), ("f_is_array", c_int32), ("angle_is_zero", c_int32), ("chunk_rows", c_int32),
        ("dx", c_float), ("dy", c_float), ("g", c_float), ("f", c_float), ("r", c_float),
        ("theta", c_float), ("desing_eps", c_float), ("flux_slope_eps", c_float),
        ("depth_cutoff", c_float), ("flux_balancer", c_float), ("land_value", c_float),
        ("north_x", c_float), ("north_y", c_float), ("wind_stress_factor", c_float), ("rho_o", c_float),
        ("divc_ok", c_int32), ("waves_per_simd", c_int32),
    ]


class Forcing(ctypes.Structure):
    _fields_ = [
        ("wind_x_cur", c_void_p), ("wind_x_next", c_void_p), ("wind_y_cur", c_void_p), ("wind_y_next", c_void_p),
        ("wind_x_nx", c_int32), ("wind_x_ny", c_int32), ("wind_y_nx", c_int32), ("wind_y_ny", c_int32),
        ("wind_t", c_float),
        ("pressure_cur", c_void_p), ("pressure_next", c_void_p),
        ("pressure_nx", c_int32), ("pressure_ny", c_int32), ("pressure_t", c_float),
    ]


class Member(ctypes.Structure):
    """gpo_member: one member of a batched ensemble."""
    _fields_ = [
        ("params", CDKLM16Params), ("dt", c_float),
        ("h0_eta", c_void_p), ("h0_hu", c_void_p), ("h0_hv", c_void_p),
        ("h1_eta", c_void_p), ("h1_hu", c_void_p), ("h1_hv", c_void_p),
        ("pitch", c_int32), ("Hi", c_void_p), ("Hi_pitch", c_int32), ("f_cells", c_void_p),
        ("boundary_conditions", c_uint32),
        ("noise", c_int32), ("xi", c_void_p), ("xi_rows", c_int32), ("xi_cols", c_int32), ("xi_pitch", c_int32),
        ("coarse", c_void_p), ("coarse_pitch", c_int32), ("seed", ctypes.c_uint64), ("stream_id", c_uint32), ("draw", c_uint32),
        ("soar_weights", c_float * 25), ("perturbation_scale", c_float),
        ("g", c_float), ("f", c_float), ("beta", c_float),
        ("bc_each_step", c_int32), ("sponge", c_int32 * 4),
        ("frs_ns_cur", c_void_p), ("frs_ns_next", c_void_p), ("frs_ns_nx", c_int32),
        ("frs_ew_cur", c_void_p), ("frs_ew_next", c_void_p), ("frs_ew_ny", c_int32), ("frs_t", c_float),
        ("drifters", c_int32), ("n_drifters", c_int32), ("positions", c_void_p), ("pos_pitch", c_int32),
        ("sensitivity", c_float), ("Bm", c_void_p), ("Bm_pitch", c_int32),
        ("coarse_nx", c_int32), ("coarse_ny", c_int32), ("coarse_dx", c_float), ("coarse_dy", c_float),
        ("coarse_offset_i", c_int32), ("coarse_offset_j", c_int32),
        ("frs_n_times", c_int32), ("frs_times", c_void_p), ("frs_ns_levels", c_void_p), ("frs_ew_levels", c_void_p),
        ("t", ctypes.c_double),
    ]


class Frs(ctypes.Structure):
    """gpo_frs: flow-relaxation data of a differentiable run."""
    _fields_ = [
        ("sponge", c_int32 * 4), ("ns_cur", c_void_p), ("ns_next", c_void_p), ("ns_nx", c_int32),
        ("ew_cur", c_void_p), ("ew_next", c_void_p), ("ew_ny", c_int32), ("t", c_float),
    ]


class AdjMember(ctypes.Structure):
    """gpo_adj_member: one member of a batched differentiable ensemble."""
    _fields_ = [
        ("params", CDKLM16Params), ("dt", c_float), ("boundary_conditions", c_uint32),
        ("Hi", c_void_p), ("Hi_pitch", c_int32), ("f_cells", c_void_p), ("pitch", c_int32),
        ("plane", ctypes.c_int64), ("state", c_void_p), ("stage", c_void_p), ("ckpt", c_void_p),
        ("qs", c_void_p), ("q1s", c_void_p), ("lam", c_void_p), ("workspace", c_void_p),
        ("has_frs", c_int32), ("frs", Frs),
        ("has_noise", c_int32), ("seed", ctypes.c_uint64), ("stream_id", c_uint32), ("draw0", c_uint32),
        ("soar_weights", c_float * 25), ("g", c_float), ("f", c_float), ("beta", c_float),
        ("xi", c_void_p), ("xi_rows", c_int32), ("xi_cols", c_int32), ("coarse", c_void_p), ("coarse_seg", c_void_p),
        ("pre", c_void_p),
        ("n_drifters", c_int32), ("Bm", c_void_p), ("Bm_pitch", c_int32), ("pos", c_void_p),
        ("pos_ckpt", c_void_p), ("pos_seg", c_void_p), ("lam_pos", c_void_p),
    ]


class ObsMember(ctypes.Structure):
    """gpo_obs_member: one member seen by the observation operator."""
    _fields_ = [
        ("eta", c_void_p), ("hu", c_void_p), ("hv", c_void_p), ("pitch", c_int32),
        ("Bm", c_void_p), ("Bm_pitch", c_int32), ("nx", c_int32), ("ny", c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("points", c_void_p), ("cells", c_void_p),
    ]


OBS_FLOW, OBS_STATE, OBS_BUOY, OBS_INNOVATION, OBS_GHOSTED = 0, 1, 2, 3, 16


class GpoError(RuntimeError):
    pass


_lib = None

# (name, restype, argtypes)
_SIGNATURES = [
    ("gpo_cdklm16_prepare", c_int32, [ctypes.POINTER(CDKLM16Params), c_void_p]),
    ("gpo_cdklm16_substep", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_float, c_int32,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_int32, c_void_p, ctypes.POINTER(Forcing), c_uint32, c_int32, c_void_p]),
    ("gpo_cdklm16_rk2_steps", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_int32, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_int32, c_void_p, c_int32, c_void_p, c_uint32, c_int32, c_void_p]),
    ("gpo_boundary_conditions", c_int32,
     [c_int32, c_int32, c_uint32, ctypes.POINTER(c_int32), c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_float, c_void_p]),
    ("gpo_bathymetry_init", c_int32,
     [c_int32, c_int32, c_uint32, c_float, c_void_p, c_int32, c_void_p, c_int32, c_void_p]),
    ("gpo_coriolis_cells", c_int32,
     [c_int32, c_int32, c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p]),
    ("gpo_max_dt", c_int32,
     [c_int32, c_int32, c_float, c_float, c_float, c_float, c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_int32, c_void_p, c_void_p]),
    ("gpo_mass_sum", c_int32, [c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p]),
    ("gpo_philox4x32_10", c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    ("gpo_random_normal", c_int32, [c_void_p, c_int32, c_int32, c_int32, ctypes.c_uint64, c_uint32, c_uint32,
                                    c_void_p]),
    ("gpo_soar", c_int32, [c_int32, c_int32, c_void_p, c_float, c_int32, c_int32, c_void_p, c_int32, c_void_p,
                           c_int32, c_int32, c_void_p]),
    ("gpo_geostrophic_balance", c_int32,
     [c_int32, c_int32, c_float, c_float, c_float, c_float, c_float, c_float, c_float, c_void_p, c_int32,
      c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_float, c_void_p]),
    ("gpo_bicubic_interpolation", c_int32,
     [c_int32, c_int32, c_float, c_float, c_int32, c_int32, c_float, c_float, c_int32, c_int32, c_float, c_float,
      c_float, c_float, c_float, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_int32, c_float, c_void_p]),
    ("gpo_drift", c_int32,
     [c_int32, c_int32, c_float, c_float, c_float, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_int32, c_float, c_void_p, c_void_p, c_void_p,
      c_void_p, c_int32, c_int32, c_float, c_float, c_void_p]),
    ("gpo_observe_underlying_flow", c_int32,
     [c_int32, c_int32, c_float, c_float, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_void_p,
      c_int32, c_float, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_void_p]),
    ("gpo_observe", c_int32,
     [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_float, c_void_p, c_void_p, c_void_p]),
    ("gpo_ensemble_create", c_int32,
     [ctypes.POINTER(Member), c_int32, c_int32, c_int32, c_int32, ctypes.POINTER(c_void_p), c_void_p]),
    ("gpo_ensemble_update", c_int32, [c_void_p, c_fp, ctypes.POINTER(c_uint32), c_void_p]),
    ("gpo_ensemble_set_times", c_int32, [c_void_p, ctypes.POINTER(ctypes.c_double), c_void_p]),
    ("gpo_ensemble_info", c_int32, [c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32)]),
    ("gpo_ensemble_step", c_int32, [c_void_p, c_int32, c_void_p]),
    ("gpo_ensemble_part", c_int32, [c_void_p, c_int32, c_void_p]),
    ("gpo_ensemble_select", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gpo_ensemble_prepare_sets", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gpo_ensemble_run_set", c_int32, [c_void_p, c_int32, c_int32, c_uint32, c_void_p]),
    ("gpo_ensemble_capture", c_int32, [c_void_p, c_int32, c_void_p]),
    ("gpo_ensemble_replay", c_int32, [c_void_p, c_int32, c_void_p]),
    ("gpo_ensemble_draws", c_int32, [c_void_p, ctypes.POINTER(c_uint32), c_void_p]),
    ("gpo_ensemble_destroy", c_int32, [c_void_p]),
    ("gpo_adjoint_workspace_bytes", ctypes.c_size_t, [c_int32, c_int32, c_int32]),
    ("gpo_tlm_workspace_bytes", ctypes.c_size_t, [c_int32, c_int32]),
    ("gpo_cdklm16_substep_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), ctypes.c_double, c_int32] + [c_void_p] * 9 +
     [c_int32, c_void_p, c_int32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    ("gpo_cdklm16_substep_tlm_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), ctypes.c_double, c_int32] + [c_void_p] * 18 +
     [c_int32, c_void_p, c_int32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    ("gpo_cdklm16_substep_adj_f32", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_float, c_int32] + [c_void_p] * 13 +
     [c_int32, c_int32, c_void_p, c_int32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    ("gpo_cdklm16_substep_adj_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), ctypes.c_double, c_int32] + [c_void_p] * 13 +
     [c_int32, c_int32, c_void_p, c_int32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    ("gpo_cdklm16_trajectory_f32", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_int32, c_float, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_int32,
      c_void_p, c_int32, c_void_p, c_uint32, ctypes.POINTER(Frs), c_void_p]),
    ("gpo_cdklm16_adjoint_steps_f32", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_int32, c_float, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
      ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_int32, c_void_p, c_int32, c_void_p, c_uint32,
      ctypes.POINTER(Frs), c_void_p, c_void_p, c_void_p]),
    ("gpo_cdklm16_adjoint_steps_f32_noalloc", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_int32, c_float, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
      ctypes.POINTER(c_void_p), c_int32, c_void_p, c_int32, c_void_p, c_uint32, ctypes.POINTER(Frs), c_void_p,
      c_void_p]),
    ("gpo_frs_adj_f32", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_uint32, ctypes.POINTER(c_int32), c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p]),
    ("gpo_frs_adj_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_uint32, ctypes.POINTER(c_int32), c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p]),
    ("gpo_frs_apply_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_uint32, ctypes.POINTER(Frs), c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p]),
    ("gpo_adj_batch_create", c_int32,
     [ctypes.POINTER(AdjMember), c_int32, c_int32, c_int32, ctypes.POINTER(c_void_p), c_void_p]),
    ("gpo_adj_batch_forward", c_int32, [c_void_p, c_int32, c_void_p]),
    ("gpo_adj_batch_checkpoint", c_int32, [c_void_p, c_int32, c_int32, c_void_p]),
    ("gpo_adj_batch_trajectory", c_int32, [c_void_p, c_int32, c_void_p]),
    ("gpo_adj_batch_sweep", c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
    ("gpo_adj_batch_destroy", c_int32, [c_void_p]),
    ("gpo_adj_batch_set_reverse", c_int32, [c_void_p, c_int32, c_int32, c_void_p]),
    ("gpo_adj_batch_slot", c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p]),
    ("gpo_balance_adj_f32", c_int32,
     [ctypes.POINTER(CDKLM16Params), c_float, c_float, c_float, c_void_p, c_int32, c_void_p, c_int32,
      c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p]),
    ("gpo_balance_adj_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), ctypes.c_double, ctypes.c_double, ctypes.c_double, c_void_p, c_int32,
      c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p]),
    ("gpo_balance_f64", c_int32,
     [ctypes.POINTER(CDKLM16Params), ctypes.c_double, ctypes.c_double, ctypes.c_double, c_void_p, c_int32,
      c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32,
      c_void_p]),
    ("gpo_drift_f64", c_int32,
     [c_int32, c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_int32, c_uint32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_double,
      c_void_p]),
    ("gpo_drift_adj_f32", c_int32,
     [c_int32, c_int32, c_float, c_float, c_float, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32,
      c_uint32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p]),
    ("gpo_drift_adj_f64", c_int32,
     [c_int32, c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, c_void_p, c_void_p, c_void_p, c_int32,
      c_void_p, c_int32, c_uint32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_double,
      c_void_p]),
    ("gpo_point_misfit_f32", c_int32,
     [c_int32, c_int32, c_float, c_float, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_int32,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gpo_point_misfit_f64", c_int32,
     [c_int32, c_int32, ctypes.c_double, ctypes.c_double, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
      c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gpo_last_error", ctypes.c_char_p, []),
    ("gpo_version", ctypes.c_char_p, []),
]

# Optional symbols bound when the library exports them.
_OPTIONAL = []


def declared_symbols():
    """Names of every function declared in include/gpuocean_amd.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "gpuocean_amd.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(gpo_[a-z0-9_]+)\s*\(", text)))


def lib():
    """The loaded library; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpoError(f"{LIB_PATH} is missing: build it with `python -m gpuocean_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES + _OPTIONAL:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def register(name, restype, argtypes):
    _OPTIONAL.append((name, restype, argtypes))
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def check(rc, what=""):
    if rc != 0:
        msg = lib().gpo_last_error().decode(errors="replace")
        raise GpoError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device (or host) address of a torch tensor, or None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())
