"""Precision study at 512^3: the fp32 GPU chain against the fp64 oracle.

The input is a 512^3 u8 phantom (sheets, anisotropic Paganin+unsharp forward
model with sigma_z/sigma_xy 1.17, radial noise).  The GPU runs the chain in
fp32: dewindow -> Wiener deconvolution (delta/beta x0.35, reg 0.015, cap 8,
sigma ratio 1.17) -> rewindow + clamp -> fast guided filter (r=2, s=2) ->
dithered u8.  The oracle runs the same steps in fp64 (orc_deconv_radial_f64 keeps
the 4097-entry LUT semantics; orc_guided_fast_f64), from the same fp32
dewindowed input.

Bars: relative L2 <= 1e-5 for the deconvolved tile and for the guided tile;
u8 within +-1 LSB, differing voxels < 1e-3.  The all-FP64 GPU variant of the same
chain (fp32 only in its emitted tile) is held to relative L2 1e-7 and < 1e-5
differing u8 voxels.  The measured figures are printed.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import clibs
import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu

N = 512
WLO, WHI = -0.1, 1.2
EPS = 2e-3


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def quant_u8_np(v):
    """Dithered quantiser of pipeline.c:53-65 at global (z,y,x) = local, vectorised:
    32-bit hash, then floor(v*255 + d) with the two roundings of fp32."""
    z = np.arange(N, dtype=np.uint32)[:, None, None]
    y = np.arange(N, dtype=np.uint32)[None, :, None]
    x = np.arange(N, dtype=np.uint32)[None, None, :]
    with np.errstate(over="ignore"):
        h = (z * np.uint32(73856093)) ^ (y * np.uint32(19349663)) ^ (x * np.uint32(83492791))
        h ^= h >> np.uint32(16)
        h *= np.uint32(0x7feb352d)
        h ^= h >> np.uint32(15)
        h *= np.uint32(0x846ca68b)
        h ^= h >> np.uint32(16)
    d = (h >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    del h
    t = v.astype(np.float32) * np.float32(255.0)
    t += d
    q = np.where(t <= 0, 0, np.minimum(t.astype(np.int64), 255))
    return q.astype(np.uint8)


def test_precision_512_fp32_gpu_vs_fp64_oracle(orc):
    rng = np.random.default_rng(4)
    vol = phantom.sheets((N, N, N), rng, n_sheets=12)
    u8 = phantom.noisy_u8(phantom.forward(vol, z_ratio=1.17, **phantom.CONFIG1_PHYS), rng, phantom.WINDOW,
                          0.02, 4.0, radial_rise=3.0)
    del vol
    phys = dict(phantom.CONFIG1_PHYS)
    phys["delta_beta"] *= 0.35

    # --- GPU, fp32
    P = F.CubeParams()
    P.phys = F.physics(**phys)
    P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, 1.17 ** 2
    P.window_lo, P.window_hi = WLO, WHI
    d_in = torch.from_numpy(u8).cuda()
    d_out = torch.empty_like(d_in)
    d_f = torch.empty(u8.shape, dtype=torch.float32, device="cuda")
    z = (0, 0, 0)
    lib = F.lib()

    def run():
        c = F.Cube(d_in.data_ptr(), *z, N, N, N, *z, N, N, N, d_out.data_ptr(), d_f.data_ptr(), C.pointer(P))
        assert lib.fy_gpu_process_cubes(C.byref(c), 1, None) == 0, lib.fy_gpu_last_error()
        torch.cuda.synchronize()

    P.guided_eps = 0.0                  # deconvolved tile (after rewindow + clamp)
    run()
    g_dec = d_f.cpu().numpy()
    P.guided_eps, P.guided_radius, P.guided_s = EPS, 2, 2
    run()
    g_gui = d_f.cpu().numpy()
    g_u8 = d_out.cpu().numpy()
    P.fp64 = 1                          # the all-FP64 variant (its fp32-emitted guided tile)
    run()
    g64_gui = d_f.cpu().numpy()
    g64_u8 = d_out.cpu().numpy()
    del d_in, d_out, d_f

    # --- oracle, fp64 from the same fp32 dewindowed input
    f = (u8.astype(np.float32) * np.float32(1.0 / 255.0)) * np.float32(WHI - WLO)
    f += np.float32(WLO)
    o = np.empty(f.shape, np.float64)
    op = clibs.Physics(phys["delta_beta"], phys["energy_kev"], phys["distance_mm"], phys["pixel_um"],
                       phys["unsharp_sigma"], phys["unsharp_coeff"], 0.0, 0.0)
    assert orc.orc_deconv_radial_f64(clibs.fp(f), o.ctypes.data_as(clibs.c_double_p), N, N, N, C.byref(op),
                                     0.015, 0, 8.0, 1.17 ** 2) == 0
    del f
    o -= WLO
    o *= 1.0 / (WHI - WLO)
    np.clip(o, 0.0, 1.0, out=o)
    e_dec = rel_l2(g_dec, o)
    m_dec = float(np.abs(g_dec - o).max())
    og = np.empty_like(o)
    assert orc.orc_guided_fast_f64(o.ctypes.data_as(clibs.c_double_p), og.ctypes.data_as(clibs.c_double_p),
                                   N, N, N, 2, EPS, 2) == 0
    del o
    e_gui = rel_l2(g_gui, og)
    m_gui = float(np.abs(g_gui - og).max())
    e64 = rel_l2(g64_gui, og)
    o_u8 = quant_u8_np(og)
    del og
    du = np.abs(g_u8.astype(np.int16) - o_u8.astype(np.int16))
    frac = float((du > 0).mean())
    du64 = np.abs(g64_u8.astype(np.int16) - o_u8.astype(np.int16))
    frac64 = float((du64 > 0).mean())
    print(f"\n512^3 fp32 GPU vs fp64 oracle: deconv relL2 {e_dec:.3e} max|d| {m_dec:.3e}; "
          f"guided relL2 {e_gui:.3e} max|d| {m_gui:.3e}; u8 max|d| {int(du.max())} differing {frac:.2e}"
          f"\n512^3 fp64 GPU vs fp64 oracle: guided (fp32-emitted) relL2 {e64:.3e}; u8 max|d| {int(du64.max())} "
          f"differing {frac64:.2e}")
    assert e_dec <= 1e-5, e_dec
    assert e_gui <= 1e-5, e_gui
    assert du.max() <= 1
    assert frac < 1e-3
    assert e64 <= 1e-7, e64               # fp32 rounding of the emitted tile only
    assert du64.max() <= 1 and frac64 < 1e-5
