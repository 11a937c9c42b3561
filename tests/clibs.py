"""ctypes handles for the checker libraries (test infrastructure).

* ``oracle()``  -- oracle/liboracle.so, the CPU restatement (prefix ``orc_``).
* ``ref()``     -- oracle/_ref/libfysics_ref.so, the reference compiled from its
                   own sources (present only where it was built); ``None`` otherwise.

Both are checkers: product code never loads them.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfysics_ref.so")

c_float_p = C.POINTER(C.c_float)
c_double_p = C.POINTER(C.c_double)
c_u8_p = C.POINTER(C.c_ubyte)
c_int_p = C.POINTER(C.c_int)


class Physics(C.Structure):
    """fy_physics (fysics.h:23-35)."""
    _fields_ = [(n, C.c_double) for n in (
        "delta_beta", "energy_kev", "distance_mm", "pixel_um", "unsharp_sigma",
        "unsharp_coeff", "psf_sigma_vox", "psf_sigma_z_vox")]


class NoiseModel(C.Structure):
    """fy_noise_model (fysics.h:191-197)."""
    _fields_ = [("g", C.c_double), ("b", C.c_double), ("noise_ref", C.c_double),
                ("ref_intensity", C.c_double), ("n_bins_used", C.c_int)]


class HistState(C.Structure):
    """fy_hist_state (fysics.h:54)."""
    _fields_ = [("hist", C.c_long * 256), ("total", C.c_long)]


class Dering(C.Structure):
    """fy_dering (fysics.h:406-415)."""
    _fields_ = [("Z", C.c_long), ("Y", C.c_long), ("X", C.c_long),
                ("cy", C.c_double), ("cx", C.c_double),
                ("slab_z", C.c_int), ("nslab", C.c_int), ("ns", C.c_int), ("nr", C.c_int),
                ("sum", c_float_p), ("cnt", C.POINTER(C.c_uint)), ("ring", c_float_p),
                ("have_rings", C.c_int)]


class PipelineCfg(C.Structure):
    """fy_pipeline_cfg (fysics.h:436-491)."""
    _fields_ = [
        ("delta_beta", C.c_double), ("energy_kev", C.c_double), ("distance_mm", C.c_double),
        ("pixel_um", C.c_double), ("unsharp_sigma", C.c_double), ("unsharp_coeff", C.c_double),
        ("machine_current_start", C.c_double), ("machine_current_stop", C.c_double),
        ("window_lo", C.c_double), ("window_hi", C.c_double),
        ("do_deconv", C.c_int), ("use_matched_deconv", C.c_int),
        ("psf_sigma_vox", C.c_double), ("deconv_tikhonov", C.c_double),
        ("guided_eps", C.c_double),
        ("do_air_zero", C.c_int), ("air_cut_u8", C.c_int), ("air_cut_band", C.c_int),
        ("air_thresh", C.c_double), ("air_cut_aggr", C.c_double),
        ("air_min_component", C.c_int), ("denoise_k", C.c_double), ("scratch_passes", C.c_int),
        ("do_normalize", C.c_int), ("norm_lo", C.c_int), ("norm_hi", C.c_int),
        ("do_zdrift", C.c_int), ("zdrift_factor", c_float_p),
        ("do_dering", C.c_int), ("calib_budget_gb", C.c_double),
        ("dering_cy", C.c_double), ("dering_cx", C.c_double),
        ("psf_p5", C.c_double), ("psf_med", C.c_double),
        ("do_musica", C.c_int), ("musica_p", C.c_double), ("musica_levels", C.c_int),
        ("musica_core", C.c_double),
        ("guided_subsample", C.c_int), ("no_dither", C.c_int), ("halo", C.c_int),
        ("have_norm", C.c_int), ("have_zdrift", C.c_int), ("have_dec_range", C.c_int),
        ("have_dering", C.c_int), ("dering", C.POINTER(Dering)),
        ("psf_z_ratio", C.c_double),
        ("have_eps_r", C.c_int), ("eps_fn_a", C.c_double), ("eps_fn_b", C.c_double),
        ("eps_fn_med", C.c_double),
        ("dec_lo", C.c_double), ("dec_hi", C.c_double), ("vol_z", C.c_long),
        ("noise_floor", C.c_double), ("occ", c_u8_p),
        ("occ_ncz", C.c_long), ("occ_ncy", C.c_long), ("occ_ncx", C.c_long),
        ("occ_cz", C.c_long), ("occ_cy", C.c_long), ("occ_cx", C.c_long)]


class CubeParams(C.Structure):
    """fy_cube_params (include/fysics_gpu.h), as orc_chain_cube reads it."""
    _fields_ = [
        ("phys", Physics), ("deconv", C.c_int), ("deconv_param", C.c_double),
        ("deconv_cap", C.c_double), ("fz2_scale", C.c_double), ("window_lo", C.c_double),
        ("window_hi", C.c_double), ("have_dec_range", C.c_int), ("dec_lo", C.c_double),
        ("dec_hi", C.c_double), ("guided_eps", C.c_double), ("guided_radius", C.c_int),
        ("guided_s", C.c_int), ("no_dither", C.c_int), ("recenter", C.c_int), ("norm_lo", C.c_int),
        ("norm_hi", C.c_int), ("dering", C.POINTER(Dering)), ("zdrift_factor", c_float_p),
        ("vol_z", C.c_long), ("fp64", C.c_int),
        ("air_zero", C.c_int), ("air_cut_u8", C.c_int), ("air_cut_band", C.c_int),
        ("air_thresh", C.c_double), ("air_cut_aggr", C.c_double), ("air_min_component", C.c_int),
        ("scratch_passes", C.c_int)]


def fp(a):
    return a.ctypes.data_as(c_float_p)


def dp(a):
    return a.ctypes.data_as(c_double_p)


def up(a):
    return a.ctypes.data_as(c_u8_p)


def lp(a):
    return a.ctypes.data_as(C.POINTER(C.c_long))


# signature table shared by the fy_ (reference) and orc_ (oracle) names
_SIGS = {
    "next_fft_size": (C.c_int, [C.c_int]),
    "fft1d": (None, [c_float_p, c_float_p, C.c_int, C.c_int]),
    "fft3d": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fft3d_normalize": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int]),
    "paganin_transfer": (C.c_double, [C.c_double, C.POINTER(Physics)]),
    "unsharp_transfer": (C.c_double, [C.c_double, C.POINTER(Physics)]),
    "recon_transfer": (C.c_double, [C.c_double, C.POINTER(Physics)]),
    "auto_reg": (C.c_double, [C.POINTER(Physics), C.c_double]),
    "auto_deltabeta_scale": (C.c_double, [C.POINTER(Physics)]),
    "kernel_halo": (C.c_int, [C.POINTER(Physics)]),
    "deconvolve": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int,
                             C.POINTER(Physics), C.c_double]),
    "deconvolve_matched": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(Physics), C.c_double]),
    "u8_to_float": (None, [c_u8_p, c_float_p, C.c_size_t]),
    "float_to_u8": (None, [c_float_p, c_u8_p, C.c_size_t]),
    "u8_to_phys": (None, [c_u8_p, c_float_p, C.c_size_t, C.c_double, C.c_double]),
    "phys_to_u8": (None, [c_float_p, c_u8_p, C.c_size_t, C.c_double, C.c_double]),
    "phys_to_u8_level": (C.c_float, [C.c_double, C.c_double, C.c_double]),
    "dering_init": (C.c_int, [C.POINTER(Dering), C.c_long, C.c_long, C.c_long,
                              C.c_double, C.c_double, C.c_int, C.c_int]),
    "dering_free": (None, [C.POINTER(Dering)]),
    "dering_tile_init": (C.c_int, [C.POINTER(Dering), C.POINTER(Dering)]),
    "dering_accumulate_u8": (None, [C.POINTER(Dering), c_u8_p, C.c_long, C.c_long,
                                    C.c_long, C.c_long, C.c_long, C.c_int]),
    "dering_merge_tile": (None, [C.POINTER(Dering), C.c_int, C.POINTER(Dering)]),
    "dering_finalize": (C.c_long, [C.POINTER(Dering), C.c_int, C.c_uint, C.c_double, C.c_double]),
    "dering_apply": (None, [C.POINTER(Dering), c_float_p, C.c_long, C.c_long, C.c_long,
                            C.c_long, C.c_long, C.c_long, C.c_double]),
    "estimate_noise": (C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_double, C.POINTER(NoiseModel)]),
    "guided_eps_for_noise": (C.c_double, [C.c_double]),
    "noise_aniso": (C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p]),
    "zdrift_apply": (None, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, c_float_p]),
    "hist_accumulate_u8": (None, [C.POINTER(HistState), c_u8_p, C.c_size_t]),
    "hist_percentile_u8": (C.c_int, [C.POINTER(HistState), C.c_double]),
    "hist_percentile_u8_inner": (C.c_int, [C.POINTER(HistState), C.c_double]),
    "norm_apply_u8": (None, [c_u8_p, c_float_p, C.c_size_t, C.c_ubyte, C.c_ubyte]),
    "zdrift_accumulate": (None, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, c_double_p,
                                 C.POINTER(C.c_long), C.c_float]),
    "zdrift_finalize": (None, [c_double_p, C.POINTER(C.c_long), C.c_int, c_float_p]),
    "flat_noise": (C.c_double, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "calib_commit_eps": (None, [C.POINTER(PipelineCfg), c_double_p, c_double_p, C.c_int]),
    "aniso_ratio": (C.c_double, [C.c_double, C.c_double]),
    "calib_commit_aniso": (None, [C.POINTER(PipelineCfg), c_double_p, C.c_int]),
    "valley_depth": (C.c_double, [C.POINTER(C.c_long), c_int_p, c_int_p, c_int_p]),
    "phase_correlate": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p,
                                  C.c_int]),
    "downscale2x": (None, [c_u8_p, C.c_int, C.c_int, C.c_int, c_u8_p, c_int_p, c_int_p, c_int_p, C.c_int,
                           C.c_float]),
    "process_buffer": (C.c_int, [C.POINTER(PipelineCfg), c_u8_p] + [C.c_long] * 14 +
                       [c_u8_p, c_int_p]),
}

_ORC_ONLY = {
    "orc_fft1d_f32": (None, [c_float_p, c_float_p, C.c_int, C.c_int]),
    "orc_fft1d_f64": (None, [c_double_p, c_double_p, C.c_int, C.c_int]),
    "orc_fft3d_f32": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_fft3d_f64": (None, [c_double_p, c_double_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_fft3d_normalize_f32": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int]),
    "orc_deconv_radial_f32": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(Physics), C.c_double, C.c_int, C.c_double,
                                        C.c_double]),
    "orc_deconv_radial_f64": (C.c_int, [c_float_p, c_double_p, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(Physics), C.c_double, C.c_int, C.c_double,
                                        C.c_double]),
    "orc_deconv_lut": (None, [C.c_int, C.c_int, C.c_int, C.POINTER(Physics), C.c_double, C.c_int,
                              C.c_double, C.c_double, c_float_p, c_double_p]),
    "orc_matched_fz2_scale": (C.c_double, [C.POINTER(Physics)]),
    "orc_box_mean_f32": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_box_mean_f64": (None, [c_double_p, c_double_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_guided_exact_f32": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double]),
    "orc_guided_fast_f32": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_int]),
    "orc_guided_exact_f64": (C.c_int, [c_double_p, c_double_p, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_double]),
    "orc_guided_fast_f64": (C.c_int, [c_double_p, c_double_p, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_double, C.c_int]),
    "orc_dither01": (C.c_float, [C.c_long, C.c_long, C.c_long]),
    "orc_quant_u8": (C.c_ubyte, [C.c_float, C.c_long, C.c_long, C.c_long, C.c_int]),
    "orc_eps_for_tile": (C.c_double, [C.POINTER(PipelineCfg), C.c_double, C.c_double,
                                      C.c_long, C.c_long, C.c_long, C.c_long]),
    "orc_chain_cube": (C.c_int, [C.POINTER(CubeParams), c_u8_p] + [C.c_long] * 12 +
                       [c_u8_p, c_int_p, c_float_p]),
    "orc_chain_cube_f64": (C.c_int, [C.POINTER(CubeParams), c_u8_p] + [C.c_long] * 12 + [c_u8_p, c_double_p]),
    "orc_zdrift_u8_sub": (None, [c_u8_p, C.c_long, C.c_long, C.c_long, C.c_long, c_double_p,
                                 C.POINTER(C.c_long)]),
    "orc_air_cc_despeckle": (None, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_air_zero": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.POINTER(CubeParams)]),
    "orc_process_buffer_f": (C.c_int, [C.POINTER(PipelineCfg), c_u8_p] + [C.c_long] * 14 +
                             [c_u8_p, c_int_p, c_float_p]),
}

class Zarr(C.Structure):
    """fy_zarr (fysics.h:376-381)."""
    _fields_ = [("root", C.c_char * 1024), ("shape", C.c_long * 3), ("chunk", C.c_long * 3),
                ("fill", C.c_ubyte)]


_REF_ONLY = {
    "fy_zarr_open": (C.c_int, [C.POINTER(Zarr), C.c_char_p]),
    "fy_zarr_read": (C.c_int, [C.POINTER(Zarr), C.c_long, C.c_long, C.c_long, C.c_long, C.c_long, C.c_long,
                               C.c_void_p]),
    "fy_guided_denoise_ws": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double, c_float_p]),
    "fy_guided_denoise_fast_ws": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_double, C.c_int, c_float_p]),
    "fy_guided_ws_floats": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "fy_box_smooth": (None, [c_float_p, c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ref_deconv_radial": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(Physics), C.c_double, C.c_int, C.c_double,
                                    C.c_double]),
    "ref_quant_u8": (C.c_ubyte, [C.c_float, C.c_long, C.c_long, C.c_long, C.c_int]),
    "ref_eps_for_tile": (C.c_double, [C.POINTER(PipelineCfg), C.c_double, C.c_double,
                                      C.c_long, C.c_long, C.c_long, C.c_long]),
    "fy_dering_tile_reset": (None, [C.POINTER(Dering)]),
    "fy_hist_init": (None, [C.POINTER(HistState)]),
    "ref_air_cc_despeckle": (None, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ref_process_tile": (C.c_int, [c_float_p, c_float_p, c_u8_p, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(PipelineCfg)]),
    "fy_correct_zdrift": (C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_float]),
}


def _declare(lib, prefix, extra):
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, prefix + name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in extra.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


_cache = {}


def oracle():
    if "orc" not in _cache:
        _cache["orc"] = _declare(C.CDLL(ORACLE_SO), "orc_", _ORC_ONLY)
    return _cache["orc"]


def ref():
    if "ref" not in _cache:
        _cache["ref"] = _declare(C.CDLL(REF_SO), "fy_", _REF_ONLY) if os.path.exists(REF_SO) else None
    return _cache["ref"]


def physics(delta_beta=1000.0, energy_kev=78.0, distance_mm=220.0, pixel_um=2.4,
            unsharp_sigma=1.2, unsharp_coeff=4.0, psf_sigma_vox=0.0, psf_sigma_z_vox=0.0):
    return Physics(delta_beta, energy_kev, distance_mm, pixel_um, unsharp_sigma,
                   unsharp_coeff, psf_sigma_vox, psf_sigma_z_vox)


def lcg_uniform(n, seed):
    """The reference tests' LCG (tests/test_fysics.c:99,226): ((rng>>8)&0xFFFF)/65535."""
    out = np.empty(n, np.float32)
    rng = seed & 0xFFFFFFFF
    # vectorised LCG jump-ahead is overkill for test sizes; loop in chunks
    a, c = 1664525, 1013904223
    vals = np.empty(n, np.uint32)
    for i in range(n):
        rng = (rng * a + c) & 0xFFFFFFFF
        vals[i] = rng
    out[:] = ((vals >> 8) & 0xFFFF).astype(np.float32) / np.float32(65535.0)
    return out
