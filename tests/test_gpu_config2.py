"""The headline workload (256^3 u8 cube: anisotropic forward blur, noise rising
3x to the rim, 20 concentric rings) through the bench's device chain, against
the oracle on the same inputs.

Bars: ring detection sums, counts, votes and radii bit-exact (the finalized
ring tables equal); the float tile after the guided filter within relL2 1e-5
of the oracle's fp32 chain and of its fp64 twin (deconvolution and guided
filter in double); the dithered u8 within +-1 on < 0.1 % of voxels of both.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import clibs
import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu


def eps_for_cube(n):
    """Radially scaled guided eps of the generator's noise model at the cube's
    centre (pipeline.c:447-456), as the benchmark uses."""
    fn0 = np.sqrt(0.02 * 110.0 + 4.0) / 255.0
    rmax = np.sqrt(2) * (n - 1) / 2
    fn = lambda r: fn0 * (1 + 2 * r / rmax)
    fn_med = fn(0.5 * rmax)
    return (4.2 * fn_med) ** 2 * min(max((fn(0.0) / fn_med) ** 2, 0.4), 8.0)


def test_config2_chain_vs_oracle(orc):
    n = 256
    print("config 2 chain 256^3", flush=True)
    u8 = phantom.config2_volume((n, n, n), seed=2)
    lib = F.lib()
    d, t = F.Dering(), F.Dering()
    od, ot = clibs.Dering(), clibs.Dering()
    assert lib.fy_dering_init(C.byref(d), n, n, n, -1.0, -1.0, 512, 8) == 0
    assert lib.fy_dering_tile_init(C.byref(t), C.byref(d)) == 0
    assert orc.orc_dering_init(C.byref(od), n, n, n, -1.0, -1.0, 512, 8) == 0
    assert orc.orc_dering_tile_init(C.byref(ot), C.byref(od)) == 0
    lib.fy_dering_accumulate_u8(C.byref(t), u8.ctypes.data_as(F.c_u8_p), 0, 0, n, n, n, 2)
    orc.orc_dering_accumulate_u8(C.byref(ot), clibs.up(u8), 0, 0, n, n, n, 2)
    nb = d.ns * d.nr
    assert np.array_equal(np.ctypeslib.as_array(t.cnt, (nb,)), np.ctypeslib.as_array(ot.cnt, (nb,)))
    assert np.array_equal(np.ctypeslib.as_array(t.sum, (nb,)), np.ctypeslib.as_array(ot.sum, (nb,)))
    lib.fy_dering_merge_tile(C.byref(d), 0, C.byref(t))
    orc.orc_dering_merge_tile(C.byref(od), 0, C.byref(ot))
    n1 = lib.fy_dering_finalize(C.byref(d), 15, 100, 0.5, 6.0)
    n2 = orc.orc_dering_finalize(C.byref(od), 15, 100, 0.5, 6.0)
    assert n1 == n2 and n1 >= 10 and d.have_rings == od.have_rings == 1
    assert np.array_equal(np.ctypeslib.as_array(d.ring, (d.nslab * d.nr,)),
                          np.ctypeslib.as_array(od.ring, (od.nslab * od.nr,)))

    P = F.CubeParams()
    P.phys = F.physics(delta_beta=300.0 * 0.35, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                       unsharp_sigma=0.8, unsharp_coeff=0.5)
    P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, 1.17 ** 2
    P.window_lo, P.window_hi = -0.1, 1.2
    P.guided_eps, P.guided_radius, P.guided_s = eps_for_cube(n), 2, 2
    P.dering = C.pointer(d)
    d_in = torch.from_numpy(u8).cuda()
    d_q = torch.empty_like(d_in)
    d_f = torch.empty(u8.shape, dtype=torch.float32, device="cuda")
    cube = F.Cube(d_in.data_ptr(), 0, 0, 0, n, n, n, 0, 0, 0, n, n, n, d_q.data_ptr(), d_f.data_ptr(), C.pointer(P))
    torch.cuda.synchronize()
    assert lib.fy_gpu_process_cubes(C.byref(cube), 1, None) == 0, lib.fy_gpu_last_error()
    assert lib.fy_gpu_sync(None) == 0
    got_q, got_f = d_q.cpu().numpy(), d_f.cpu().numpy()

    op = clibs.CubeParams()
    C.memmove(C.byref(op), C.byref(P), C.sizeof(op))
    op.dering = C.pointer(od)
    want_q = np.empty(u8.shape, np.uint8)
    want_f = np.empty(u8.shape, np.float32)
    air = C.c_int(0)
    assert orc.orc_chain_cube(C.byref(op), clibs.up(u8), 0, 0, 0, n, n, n, 0, 0, 0, n, n, n, clibs.up(want_q),
                              C.byref(air), clibs.fp(want_f)) == 0
    r = np.linalg.norm((got_f - want_f).astype(np.float64)) / np.linalg.norm(want_f.astype(np.float64))
    diff = np.abs(got_q.astype(int) - want_q.astype(int))
    # the fp64 oracle: deconvolution and guided filter in double (same LUT)
    q64 = np.empty(u8.shape, np.uint8)
    f64 = np.empty(u8.shape, np.float64)
    assert orc.orc_chain_cube_f64(C.byref(op), clibs.up(u8), 0, 0, 0, n, n, n, 0, 0, 0, n, n, n, clibs.up(q64),
                                  clibs.dp(f64)) == 0
    r64 = np.linalg.norm(got_f - f64) / np.linalg.norm(f64)
    d64 = np.abs(got_q.astype(int) - q64.astype(int))
    print(f"config 2: rings {n1}, float relL2 {r:.2e} (fp32 oracle) {r64:.2e} (fp64 oracle), "
          f"u8 off by one {(diff > 0).mean():.2e} / {(d64 > 0).mean():.2e}", flush=True)
    assert r <= 1e-5 and r64 <= 1e-5
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    assert d64.max() <= 1 and (d64 > 0).mean() < 1e-3
    for x in (t, d):
        lib.fy_dering_free(C.byref(x))
    for x in (ot, od):
        orc.orc_dering_free(C.byref(x))
