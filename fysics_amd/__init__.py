"""fysics_amd -- MI355X-native volume-restoration hot path of fysics.

Python mirror of the C API in include/fysics.h and include/fysics_gpu.h, bound
with ctypes to the in-tree libfysics_amd.so (C host code + HIP kernels for
gfx950).  Function names, argument meaning and the 0/1 return convention are
the reference's (fysics.h); numpy arrays stand in for the host pointers and
torch tensors (or raw integers) for device pointers.

There is no Python or CPU fallback: if the shared library is missing or no GPU
is present, calls raise.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["lib", "Physics", "NoiseModel", "Dering", "PipelineCfg", "CubeParams", "Cube", "physics",
           "pipeline_cfg",
           "deconvolve", "deconvolve_matched", "guided", "guided_fast", "box_smooth", "fft1d",
           "fft3d", "process_buffer", "LIB_PATH"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfysics_amd.so")

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
    """fy_cube_params (include/fysics_gpu.h)."""
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


class Cube(C.Structure):
    """fy_cube (include/fysics_gpu.h): device u8 region, inner tile, outputs."""
    _fields_ = [("d_in", C.c_void_p), ("rz0", C.c_long), ("ry0", C.c_long), ("rx0", C.c_long),
                ("hz", C.c_long), ("hy", C.c_long), ("hx", C.c_long), ("z0", C.c_long),
                ("y0", C.c_long), ("x0", C.c_long), ("tz", C.c_long), ("ty", C.c_long),
                ("tx", C.c_long), ("d_out", C.c_void_p), ("d_float", C.c_void_p),
                ("params", C.POINTER(CubeParams))]


class Zarr(C.Structure):
    """fy_zarr (fysics.h:376-381): local zarr v2 raw-u8 volume handle."""
    _fields_ = [("root", C.c_char * 1024), ("shape", C.c_long * 3), ("chunk", C.c_long * 3),
                ("fill", C.c_ubyte)]


class ExportStats(C.Structure):
    """fy_export_stats (include/fysics_gpu.h)."""
    _fields_ = [("tiles", C.c_long), ("written", C.c_long), ("air", C.c_long), ("io_fail", C.c_long),
                ("wall_s", C.c_double), ("read_s", C.c_double), ("write_s", C.c_double)]


_P = C.POINTER(Physics)
_VP = C.c_void_p
_L = C.c_long
_ZP = C.POINTER(Zarr)

# name -> (restype, argtypes); every function declared in include/fysics*.h
SIGNATURES = {
    "fy_is_pow2": (C.c_int, [C.c_int]),
    "fy_next_pow2": (C.c_int, [C.c_int]),
    "fy_next_fft_size": (C.c_int, [C.c_int]),
    "fy_fft1d": (None, [c_float_p, c_float_p, C.c_int, C.c_int]),
    "fy_fft3d": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fy_fft3d_normalize": (None, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int]),
    "fy_u8_to_float": (None, [c_u8_p, c_float_p, C.c_size_t]),
    "fy_float_to_u8": (None, [c_float_p, c_u8_p, C.c_size_t]),
    "fy_u8_to_phys": (None, [c_u8_p, c_float_p, C.c_size_t, C.c_double, C.c_double]),
    "fy_phys_to_u8": (None, [c_float_p, c_u8_p, C.c_size_t, C.c_double, C.c_double]),
    "fy_phys_to_u8_level": (C.c_float, [C.c_double, C.c_double, C.c_double]),
    "fy_paganin_transfer": (C.c_double, [C.c_double, _P]),
    "fy_unsharp_transfer": (C.c_double, [C.c_double, _P]),
    "fy_recon_transfer": (C.c_double, [C.c_double, _P]),
    "fy_auto_reg": (C.c_double, [_P, C.c_double]),
    "fy_auto_deltabeta_scale": (C.c_double, [_P]),
    "fy_deconvolve": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, _P, C.c_double]),
    "fy_deconvolve_matched": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, _P,
                                        C.c_double]),
    "fy_kernel_halo": (C.c_int, [_P]),
    "fy_guided_denoise": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double]),
    "fy_guided_denoise_ws": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double, c_float_p]),
    "fy_guided_ws_floats": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "fy_guided_denoise_fast_ws": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_double, C.c_int, c_float_p]),
    "fy_box_smooth": (None, [c_float_p, c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fy_guided_eps_for_noise": (C.c_double, [C.c_double]),
    "fy_estimate_noise": (C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.POINTER(NoiseModel)]),
    "fy_noise_aniso": (C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p]),
    "fy_zdrift_apply": (None, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, c_float_p]),
    "fy_zdrift_accumulate": (None, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int, c_double_p,
                                    C.POINTER(C.c_long), C.c_float]),
    "fy_zdrift_finalize": (None, [c_double_p, C.POINTER(C.c_long), C.c_int, c_float_p]),
    "fy_correct_zdrift": (C.c_int, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_float]),
    "fy_hist_init": (None, [C.POINTER(HistState)]),
    "fy_hist_accumulate_u8": (None, [C.POINTER(HistState), c_u8_p, C.c_size_t]),
    "fy_hist_merge": (None, [C.POINTER(HistState), C.POINTER(HistState)]),
    "fy_hist_percentile_u8": (C.c_int, [C.POINTER(HistState), C.c_double]),
    "fy_hist_percentile_u8_inner": (C.c_int, [C.POINTER(HistState), C.c_double]),
    "fy_norm_finalize": (None, [C.POINTER(HistState), C.c_double, C.c_double, c_u8_p, c_u8_p]),
    "fy_norm_apply_u8": (None, [c_u8_p, c_float_p, C.c_size_t, C.c_ubyte, C.c_ubyte]),
    "fy_flat_noise": (C.c_double, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fy_phase_correlate": (C.c_int, [c_float_p, c_float_p, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p,
                                     C.c_int]),
    "fy_gpu_phase_correlate": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p, C.c_int,
                                         _VP]),
    "fy_downscale2x": (None, [c_u8_p, C.c_int, C.c_int, C.c_int, c_u8_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int), C.c_int, C.c_float]),
    "fy_gpu_downscale2x": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, _VP, C.c_int, C.c_float, _VP]),
    "fy_valley_depth": (C.c_double, [C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "fy_dering_init": (C.c_int, [C.POINTER(Dering), _L, _L, _L, C.c_double, C.c_double, C.c_int,
                                 C.c_int]),
    "fy_dering_free": (None, [C.POINTER(Dering)]),
    "fy_dering_tile_init": (C.c_int, [C.POINTER(Dering), C.POINTER(Dering)]),
    "fy_dering_tile_reset": (None, [C.POINTER(Dering)]),
    "fy_dering_accumulate_u8": (None, [C.POINTER(Dering), c_u8_p, _L, _L, _L, _L, _L, C.c_int]),
    "fy_dering_merge_tile": (None, [C.POINTER(Dering), C.c_int, C.POINTER(Dering)]),
    "fy_dering_finalize": (C.c_long, [C.POINTER(Dering), C.c_int, C.c_uint, C.c_double,
                                      C.c_double]),
    "fy_dering_apply": (None, [C.POINTER(Dering), c_float_p, _L, _L, _L, _L, _L, _L, C.c_double]),
    "fy_process_buffer": (C.c_int, [C.POINTER(PipelineCfg), c_u8_p] + [_L] * 14 + [c_u8_p, c_int_p]),
    # fysics_gpu.h
    "fy_gpu_deconvolve": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, _P, C.c_double, _VP]),
    "fy_gpu_deconvolve_matched": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, _P, C.c_double,
                                            _VP]),
    "fy_gpu_deconv_radial": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, _P, C.c_double,
                                       C.c_int, C.c_double, C.c_double, _VP]),
    "fy_gpu_deconv_radial_f64": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, _P, C.c_double,
                                           C.c_int, C.c_double, C.c_double, _VP]),
    "fy_gpu_guided": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                _VP]),
    "fy_gpu_process_buffer": (C.c_int, [C.POINTER(PipelineCfg), _VP] + [_L] * 14 +
                              [_VP, c_int_p, _VP]),
    "fy_gpu_process_cubes": (C.c_int, [C.POINTER(Cube), C.c_int, _VP]),
    "fy_gpu_graph_cubes": (_VP, [C.POINTER(Cube), C.c_int, _VP]),
    "fy_graph_launch": (C.c_int, [_VP]),
    "fy_graph_destroy": (None, [_VP]),
    "fy_cube_workspace_bytes": (C.c_size_t, [C.POINTER(Cube)]),
    "fy_serve_create": (_VP, [C.c_size_t, C.c_int, C.c_int, _VP]),
    "fy_serve_submit": (C.c_int, [_VP, C.POINTER(Cube), _L]),
    "fy_serve_submit_host": (C.c_int, [_VP, C.POINTER(Cube), _VP, _VP, C.c_long]),
    "fy_serve_poll": (C.c_int, [_VP, C.POINTER(C.c_long), C.c_int]),
    "fy_serve_outstanding": (C.c_int, [_VP]),
    "fy_serve_destroy": (None, [_VP]),
    "fy_cube_params_from_cfg": (C.c_int, [C.POINTER(PipelineCfg), _L, _L, _L, _L, _L, _L,
                                          C.POINTER(CubeParams)]),
    "fy_calib_create": (_VP, [_L, _VP]),
    "fy_calib_accumulate_u8": (C.c_int, [_VP, _VP, _L, _L, _L, _L, C.c_int]),
    "fy_calib_read": (C.c_int, [_VP, C.POINTER(HistState), c_double_p, C.POINTER(C.c_long)]),
    "fy_calib_reset": (C.c_int, [_VP]),
    "fy_calib_destroy": (None, [_VP]),
    "fy_gpu_flat_noise": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, c_double_p, _VP]),
    "fy_gpu_zarr_open": (C.c_int, [_ZP, C.c_char_p]),
    "fy_gpu_zarr_create": (C.c_int, [_ZP, C.c_char_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "fy_gpu_zarr_read": (C.c_int, [_ZP, _L, _L, _L, _L, _L, _L, _VP]),
    "fy_gpu_zarr_write_chunk": (C.c_int, [_ZP, _L, _L, _L, _VP, _L, _L, _L]),
    "fy_gpu_zarr_cache_new": (C.c_void_p, [C.c_size_t]),
    "fy_gpu_zarr_cache_free": (None, [C.c_void_p]),
    "fy_gpu_zarr_read_cached": (C.c_int, [_ZP, C.c_void_p, _L, _L, _L, _L, _L, _L, _VP]),
    "fy_gpu_process_chunk": (C.c_int, [_ZP, C.POINTER(PipelineCfg), _L, _L, _L, C.c_int, _VP, c_int_p, c_int_p,
                                   c_int_p, c_int_p]),
    "fy_gpu_export": (C.c_int, [_ZP, _ZP, C.POINTER(PipelineCfg), C.c_int, _L, _L, C.c_int, C.c_int,
                                C.POINTER(ExportStats), _VP]),
    "fy_calib_commit_eps": (None, [C.POINTER(PipelineCfg), c_double_p, c_double_p, C.c_int]),
    "fy_aniso_ratio": (C.c_double, [C.c_double, C.c_double]),
    "fy_calib_commit_aniso": (None, [C.POINTER(PipelineCfg), c_double_p, C.c_int]),
    "fy_gpu_tuning": (C.c_int, [C.c_int]),
    "fy_gpu_tuning_report": (C.c_int, [C.c_char_p, C.c_int]),
    "fy_gpu_tuning_save": (C.c_int, [C.c_char_p]),
    "fy_gpu_tuning_load": (C.c_int, [C.c_char_p]),
    "fy_gpu_tuning_clear": (None, []),
    "fy_gpu_sync": (C.c_int, [_VP]),
    "fy_gpu_stream_release": (C.c_int, [_VP]),
    "fy_gpu_table_cache": (C.c_long, [C.POINTER(C.c_long), C.c_long]),
    "fy_gpu_fft_planes": (C.c_int, [C.c_int]),
    "fy_gpu_batch_mix": (C.c_int, [C.c_int]),
    "fy_gpu_fft_radix": (C.c_int, [C.c_int]),
    "fy_gpu_deconvolve_noalloc": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, _P, C.c_double, _VP]),
    "fy_gpu_profile": (None, [C.c_int]),
    "fy_gpu_profile_read": (C.c_int, [C.c_char_p, C.c_int, c_double_p, C.POINTER(C.c_long),
                                      C.c_int]),
    "fy_gpu_profile_reset": (None, []),
    "fy_gpu_last_error": (C.c_char_p, []),
}

_lib = None


def lib():
    """The loaded libfysics_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m fysics_amd.build` "
                               "(there is no fallback implementation)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def physics(delta_beta=1000.0, energy_kev=78.0, distance_mm=220.0, pixel_um=2.4, unsharp_sigma=1.2,
            unsharp_coeff=4.0, psf_sigma_vox=0.0, psf_sigma_z_vox=0.0):
    return Physics(delta_beta, energy_kev, distance_mm, pixel_um, unsharp_sigma, unsharp_coeff,
                   psf_sigma_vox, psf_sigma_z_vox)


def pipeline_cfg(**kw):
    c = PipelineCfg()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _fp(a):
    return a.ctypes.data_as(c_float_p)


def _check(rc, what):
    if rc != 0:
        err = lib().fy_gpu_last_error()
        raise RuntimeError(f"{what} failed (rc={rc}): {err.decode() if err else ''}")


def deconvolve(vol, p, reg=-1.0):
    """fy_deconvolve on a host volume (paganin.c:301)."""
    v = _f32(vol)
    out = np.empty_like(v)
    _check(lib().fy_deconvolve(_fp(v), _fp(out), *v.shape, C.byref(p), reg), "fy_deconvolve")
    return out


def deconvolve_matched(vol, p, tikhonov=-1.0):
    v = _f32(vol)
    out = np.empty_like(v)
    _check(lib().fy_deconvolve_matched(_fp(v), _fp(out), *v.shape, C.byref(p), tikhonov),
           "fy_deconvolve_matched")
    return out


def guided(vol, radius=2, eps=1e-3):
    v = _f32(vol)
    out = np.empty_like(v)
    _check(lib().fy_guided_denoise(_fp(v), _fp(out), *v.shape, radius, eps), "fy_guided_denoise")
    return out


def guided_fast(vol, radius=2, eps=1e-3, s=2):
    v = _f32(vol)
    out = np.empty_like(v)
    _check(lib().fy_guided_denoise_fast_ws(_fp(v), _fp(out), *v.shape, radius, eps, s, None),
           "fy_guided_denoise_fast_ws")
    return out


def box_smooth(vol, r):
    v = _f32(vol)
    out = np.empty_like(v)
    lib().fy_box_smooth(_fp(v), _fp(out), None, *v.shape, r)
    return out


def fft1d(re, im, sign):
    re = _f32(re).copy()
    im = _f32(im).copy()
    lib().fy_fft1d(_fp(re), _fp(im), re.size, sign)
    return re, im


def fft3d(re, im, sign):
    re = _f32(re).copy()
    im = _f32(im).copy()
    lib().fy_fft3d(_fp(re), _fp(im), *re.shape, sign)
    return re, im


def process_buffer(cfg, u8buf, origin, inner_origin, inner_shape, vol_yx):
    """fy_process_buffer: u8buf is the halo'd region at global `origin`."""
    u = np.ascontiguousarray(u8buf, dtype=np.uint8)
    out = np.empty(inner_shape, np.uint8)
    air = C.c_int(0)
    rc = lib().fy_process_buffer(C.byref(cfg), u.ctypes.data_as(c_u8_p), *origin, *u.shape,
                                 *inner_origin, *inner_shape, *vol_yx,
                                 out.ctypes.data_as(c_u8_p), C.byref(air))
    _check(rc, "fy_process_buffer")
    return out, bool(air.value)


def profile_read():
    """Per-kernel event timings: {name: (total_ms, launches)}."""
    names = C.create_string_buffer(8192)
    ms = (C.c_double * 128)()
    cnt = (C.c_long * 128)()
    n = lib().fy_gpu_profile_read(names, 8192, ms, cnt, 128)
    keys = names.raw.split(b"\0")
    return {keys[i].decode(): (ms[i], cnt[i]) for i in range(n)}


def tuning_report():
    """The fast guided filter's tuned tile shapes, one line per shape."""
    n = lib().fy_gpu_tuning_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().fy_gpu_tuning_report(buf, n + 1)
    return buf.value.decode()
