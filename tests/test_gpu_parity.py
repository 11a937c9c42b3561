"""GPU parity: the HIP path (through the C ABI of libfysics_amd.so) against the
oracle (CPU restatement, oracle/liboracle.so) on the same seeded inputs.

Bars (stated per test): relative L2 <= 1e-5 for fp32 volumes against the fp64
oracle (which keeps the reference's LUT semantics), bit-exact for integer and
byte work (quantisation, conversions, ring sums/counts/votes), u8 chain output
within +-1 LSB.
"""
import ctypes as C

import numpy as np
import pytest

import clibs
import fysics_amd as F

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


LENGTHS = [1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024,
           1536, 2048, 3072, 4096]


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_vs_fp64_dft(n, sign):
    rng = np.random.default_rng(n)
    re = rng.standard_normal(n).astype(np.float32)
    im = rng.standard_normal(n).astype(np.float32)
    gr, gi = F.fft1d(re, im, sign)
    x = re.astype(np.float64) + 1j * im
    want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    got = gr.astype(np.float64) + 1j * gi
    assert rel_l2(got, want) < 2e-6 * max(1.0, np.log2(n))


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 24), (3, 6, 4), (1, 1, 16), (16, 32, 12)])
def test_fft3d_roundtrip_and_values(shape):
    rng = np.random.default_rng(7)
    re = rng.random(shape).astype(np.float32)
    im = np.zeros(shape, np.float32)
    fr, fi = F.fft3d(re, im, -1)
    want = np.fft.fftn(re.astype(np.float64))
    assert rel_l2(fr + 1j * fi, want) < 1e-5
    br, bi = F.fft3d(fr, fi, +1)
    lib = F.lib()
    lib.fy_fft3d_normalize(br.ctypes.data_as(F.c_float_p), bi.ctypes.data_as(F.c_float_p), *shape)
    assert np.abs(br - re).max() < 1e-4           # tests/test_fysics.c:52,84


DECONV_SHAPES = [(22, 27, 30), (16, 16, 16), (5, 7, 9), (1, 4, 6), (2, 3, 5), (33, 17, 40),
                 (64, 64, 64), (96, 96, 96), (80, 70, 90)]


@pytest.mark.parametrize("shape", DECONV_SHAPES)
def test_deconvolve_vs_oracle(orc, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.random(shape).astype(np.float32)
    p = F.physics()
    got = F.deconvolve(x, p, 0.015)
    want64 = np.empty(shape, np.float64)
    pc = clibs.physics()
    assert orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(want64), *shape, C.byref(pc), 0.015, 0, 8.0, 1.0) == 0
    want32 = np.empty(shape, np.float32)
    assert orc.orc_deconvolve(clibs.fp(x), clibs.fp(want32), *shape, C.byref(pc), 0.015) == 0
    assert rel_l2(got, want64) <= REL_L2
    assert rel_l2(got, want32) <= REL_L2


# every FFT length from 96 up on each axis (plans of three or more stages, and
# column passes with fewer than 16 lines per workgroup)
LONG_AXES = [(96, 6, 6), (384, 4, 4), (512, 4, 4), (768, 4, 4), (1024, 3, 3), (2048, 2, 2), (4096, 2, 2)]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("base", LONG_AXES)
def test_deconvolve_long_axes_vs_oracle(orc, base, axis):
    shape = tuple(np.roll(base, axis))
    rng = np.random.default_rng(sum(shape) + axis)
    x = rng.random(shape).astype(np.float32)
    p = F.physics()
    got = F.deconvolve(x, p, 0.015)
    want64 = np.empty(shape, np.float64)
    pc = clibs.physics()
    assert orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(want64), *shape, C.byref(pc), 0.015, 0, 8.0, 1.0) == 0
    assert rel_l2(got, want64) <= REL_L2


@pytest.mark.parametrize("shape", [(22, 27, 30), (5, 7, 9), (64, 64, 64), (96, 40, 24), (512, 4, 4), (4, 4, 768)])
def test_deconvolve_f64_vs_oracle(orc, shape):
    """The all-FP64 variant against the fp64 oracle (same fp64 LUT weights):
    relative L2 <= 1e-12."""
    import torch
    rng = np.random.default_rng(sum(shape) + 64)
    x = rng.random(shape).astype(np.float32)
    p = F.physics()
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.empty(shape, dtype=torch.float64, device="cuda")
    assert F.lib().fy_gpu_deconv_radial_f64(d_in.data_ptr(), d_out.data_ptr(), *shape, C.byref(p), 0.015, 0,
                                            8.0, 1.17 ** 2, None) == 0
    torch.cuda.synchronize()
    want64 = np.empty(shape, np.float64)
    pc = clibs.physics()
    assert orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(want64), *shape, C.byref(pc), 0.015, 0, 8.0,
                                     1.17 ** 2) == 0
    assert rel_l2(d_out.cpu().numpy(), want64) <= 1e-12


@pytest.mark.parametrize("shape", [(20, 24, 28), (48, 48, 48), (7, 9, 11)])
def test_deconvolve_matched_aniso_vs_oracle(orc, shape):
    rng = np.random.default_rng(3)
    x = rng.random(shape).astype(np.float32)
    p = F.physics(psf_sigma_vox=1.1, psf_sigma_z_vox=1.3)
    got = F.deconvolve_matched(x, p, 0.05)
    pc = clibs.physics(psf_sigma_vox=1.1, psf_sigma_z_vox=1.3)
    want64 = np.empty(shape, np.float64)
    s = orc.orc_matched_fz2_scale(C.byref(pc))
    assert orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(want64), *shape, C.byref(pc), 0.05, 1, 0.0, s) == 0
    assert rel_l2(got, want64) <= REL_L2


def test_deconvolve_in_place_aliasing():
    rng = np.random.default_rng(5)
    x = rng.random((12, 20, 18)).astype(np.float32)
    p = F.physics()
    out = F.deconvolve(x, p, 0.015)
    y = x.copy()
    lib = F.lib()
    assert lib.fy_deconvolve(F._fp(y), F._fp(y), *y.shape, C.byref(p), 0.015) == 0
    assert np.array_equal(y, out)


GUIDED_SHAPES = [(20, 18, 22), (32, 32, 32), (5, 6, 7), (3, 40, 33), (64, 48, 40)]


@pytest.mark.parametrize("shape", GUIDED_SHAPES)
@pytest.mark.parametrize("r", [1, 2, 3])
def test_guided_exact_vs_oracle(orc, shape, r):
    rng = np.random.default_rng(11)
    x = (0.3 + 0.4 * (np.arange(shape[2]) > shape[2] // 2) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    got = F.guided(x, r, 4e-4)
    want = np.empty(shape, np.float64)
    assert orc.orc_guided_exact_f64(clibs.dp(x.astype(np.float64)), clibs.dp(want), *shape, r, 4e-4) == 0
    assert rel_l2(got, want) <= REL_L2


@pytest.mark.parametrize("shape", GUIDED_SHAPES)
@pytest.mark.parametrize("s", [2, 3])
def test_guided_fast_vs_oracle(orc, shape, s):
    rng = np.random.default_rng(12)
    x = (0.3 + 0.4 * (np.arange(shape[2]) > shape[2] // 2) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    got = F.guided_fast(x, 2, 4e-4, s)
    want = np.empty(shape, np.float64)
    assert orc.orc_guided_fast_f64(clibs.dp(x.astype(np.float64)), clibs.dp(want), *shape, 2, 4e-4, s) == 0
    assert rel_l2(got, want) <= REL_L2


@pytest.mark.parametrize("r", [0, 1, 2, 3, 5])
def test_box_smooth_vs_oracle(orc, r):
    rng = np.random.default_rng(13)
    x = rng.random((17, 19, 23)).astype(np.float32)
    got = F.box_smooth(x, r)
    want = np.empty(x.shape, np.float64)
    orc.orc_box_mean_f64(clibs.dp(x.astype(np.float64)), clibs.dp(want), *x.shape, r)
    assert rel_l2(got, want) <= REL_L2


def test_conversions_bit_exact(orc):
    lib = F.lib()
    u = np.arange(256, dtype=np.uint8).repeat(7)
    n = u.size
    a = np.empty(n, np.float32); b = np.empty(n, np.float32)
    lib.fy_u8_to_float(u.ctypes.data_as(F.c_u8_p), F._fp(a), n)
    orc.orc_u8_to_float(clibs.up(u), clibs.fp(b), n)
    assert np.array_equal(a, b)
    lib.fy_u8_to_phys(u.ctypes.data_as(F.c_u8_p), F._fp(a), n, -0.03, 0.21)
    orc.orc_u8_to_phys(clibs.up(u), clibs.fp(b), n, -0.03, 0.21)
    assert np.array_equal(a, b)
    back = np.empty(n, np.uint8)
    lib.fy_phys_to_u8(F._fp(a), back.ctypes.data_as(F.c_u8_p), n, -0.03, 0.21)
    assert np.array_equal(back, u)                # tests/test_fysics.c:439-447
    rng = np.random.default_rng(0)
    f = (rng.random(100000) * 1.4 - 0.2).astype(np.float32)
    q1 = np.empty(f.size, np.uint8); q2 = np.empty(f.size, np.uint8)
    lib.fy_float_to_u8(F._fp(f), q1.ctypes.data_as(F.c_u8_p), f.size)
    orc.orc_float_to_u8(clibs.fp(f), clibs.up(q2), f.size)
    assert np.array_equal(q1, q2)


def test_kernel_halo_matches_oracle(orc):
    for kw in [dict(), dict(delta_beta=300, energy_kev=110, distance_mm=1200, pixel_um=4.3,
                            unsharp_sigma=0.8, unsharp_coeff=0.5)]:
        assert F.lib().fy_kernel_halo(C.byref(F.physics(**kw))) == orc.orc_kernel_halo(C.byref(clibs.physics(**kw)))


def ring_phantom():
    """tests/test_fysics.c:121-142: spiral winding + rings at R=60 (+10) and R=110 (-8)."""
    nz, ny, nx = 32, 256, 256
    cy = cx = 127.5
    z = np.arange(nz)[:, None, None]
    yy, xx = np.meshgrid(np.arange(ny) - cy, np.arange(nx) - cx, indexing="ij")
    r = np.sqrt(yy ** 2 + xx ** 2)
    th = np.arctan2(yy, xx)
    v = 128 + 30 * np.cos(2 * np.pi * (r - 192 * th / (2 * np.pi)) / 24)
    v = v + 10 * np.exp(-(r - 60) ** 2 / (2 * 0.8 ** 2)) - 8 * np.exp(-(r - 110) ** 2 / (2 * 0.8 ** 2))
    rng = np.random.default_rng(777)
    v = v[None] + rng.integers(0, 256, size=(nz, ny, nx)) / 25.5 - 5.0 + 0 * z
    return np.clip(v, 1, 255).astype(np.uint8)


@pytest.mark.parametrize("ss", [1, 2])
def test_dering_accumulate_finalize_bit_exact(orc, ss):
    vol = ring_phantom()
    nz, ny, nx = vol.shape
    lib = F.lib()
    d = F.Dering(); t = F.Dering()
    od = clibs.Dering(); ot = clibs.Dering()
    assert lib.fy_dering_init(C.byref(d), nz, ny, nx, -1, -1, 64, 8) == 0
    assert lib.fy_dering_tile_init(C.byref(t), C.byref(d)) == 0
    assert orc.orc_dering_init(C.byref(od), nz, ny, nx, -1, -1, 64, 8) == 0
    assert orc.orc_dering_tile_init(C.byref(ot), C.byref(od)) == 0
    # two partial accumulations (two z halves) then one merge, like a tile loop
    for z0 in (0, 16):
        sub = np.ascontiguousarray(vol[z0:z0 + 16])
        lib.fy_dering_accumulate_u8(C.byref(t), sub.ctypes.data_as(F.c_u8_p), 0, 0, 16, ny, nx, ss)
        orc.orc_dering_accumulate_u8(C.byref(ot), clibs.up(sub), 0, 0, 16, ny, nx, ss)
    nb = d.ns * d.nr
    assert np.array_equal(np.ctypeslib.as_array(t.cnt, (nb,)), np.ctypeslib.as_array(ot.cnt, (nb,)))
    assert np.array_equal(np.ctypeslib.as_array(t.sum, (nb,)), np.ctypeslib.as_array(ot.sum, (nb,)))
    lib.fy_dering_merge_tile(C.byref(d), 0, C.byref(t))
    orc.orc_dering_merge_tile(C.byref(od), 0, C.byref(ot))
    n1 = lib.fy_dering_finalize(C.byref(d), 15, 50, 2.0, 15.0)
    n2 = orc.orc_dering_finalize(C.byref(od), 15, 50, 2.0, 15.0)
    assert n1 == n2 and d.have_rings == od.have_rings == 1
    ring = np.ctypeslib.as_array(d.ring, (d.nslab * d.nr,))
    assert np.array_equal(ring, np.ctypeslib.as_array(od.ring, (od.nslab * od.nr,)))
    assert ring[59:62].max() > 4.0 and ring[109:112].min() < -3.0
    f = (vol * np.float32(1 / 255)).astype(np.float32)
    g = f.copy()
    lib.fy_dering_apply(C.byref(d), F._fp(f), 0, 0, 0, nz, ny, nx, 1.0 / 255.0)
    orc.orc_dering_apply(C.byref(od), clibs.fp(g), 0, 0, 0, nz, ny, nx, 1.0 / 255.0)
    assert np.array_equal(f, g)
    for x in (t, d):
        lib.fy_dering_free(C.byref(x))
    for x in (ot, od):
        orc.orc_dering_free(C.byref(x))


def _chain_cfg(lib_mod, **kw):
    base = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                unsharp_sigma=0.8, unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2,
                do_deconv=1, guided_eps=1.2e-3, guided_subsample=0)
    base.update(kw)
    c = lib_mod.PipelineCfg()
    for k, v in base.items():
        setattr(c, k, v)
    return c


CHAIN_CASES = [
    dict(),
    dict(no_dither=1),
    dict(use_matched_deconv=1, psf_sigma_vox=1.0, psf_z_ratio=1.17, deconv_tikhonov=0.05),
    dict(guided_subsample=1),
    dict(do_deconv=0),
    dict(guided_eps=0.0),
    dict(have_norm=1, norm_lo=40, norm_hi=200),
    dict(have_dec_range=1, dec_lo=0.02, dec_hi=0.9),
    dict(have_eps_r=1, eps_fn_a=0.01, eps_fn_b=0.0002, eps_fn_med=0.012),
]


@pytest.mark.parametrize("case", range(len(CHAIN_CASES)))
def test_process_buffer_vs_oracle(orc, case):
    from fysics_amd import phantom
    rng = np.random.default_rng(case)
    vol = phantom.sheets((40, 44, 48), rng, n_sheets=8)
    u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
    kw = CHAIN_CASES[case]
    cfg = _chain_cfg(F, **kw)
    ocfg = _chain_cfg(clibs, **kw)
    origin = (100, 200, 300)
    inner_o = (108, 206, 310)
    inner = (24, 30, 28)
    got, air = F.process_buffer(cfg, u8, origin, inner_o, inner, (512, 640))
    want = np.empty(inner, np.uint8)
    oair = C.c_int(0)
    assert orc.orc_process_buffer(C.byref(ocfg), clibs.up(u8), *origin, *u8.shape, *inner_o, *inner,
                                  512, 640, clibs.up(want), C.byref(oair)) == 0
    assert air == bool(oair.value) == False
    diff = np.abs(got.astype(int) - want.astype(int))
    assert diff.max() <= 1
    assert (diff > 0).mean() < 1e-3


def test_process_buffer_all_air():
    cfg = _chain_cfg(F)
    u8 = np.zeros((20, 20, 20), np.uint8)
    got, air = F.process_buffer(cfg, u8, (0, 0, 0), (2, 2, 2), (16, 16, 16), (20, 20))
    assert air and not got.any()


def test_process_buffer_rejects_unsupported_stages():
    cfg = _chain_cfg(F, do_musica=1)
    u8 = np.full((12, 12, 12), 100, np.uint8)
    with pytest.raises(RuntimeError):
        F.process_buffer(cfg, u8, (0, 0, 0), (0, 0, 0), (12, 12, 12), (12, 12))


@pytest.mark.parametrize("shape,sig", [((40, 40, 40), 0.05), ((48, 40, 56), 0.02), ((24, 24, 24), 0.1)])
def test_estimate_noise_vs_oracle(orc, shape, sig):
    """Local moments are formed in the reference's summation order and the
    percentiles are exact order statistics, so the model matches the oracle to
    the last bits (fp64 fit on identical inputs)."""
    rng = np.random.default_rng(int(sig * 1000))
    zz, yy, xx = np.meshgrid(*[np.arange(k) for k in shape], indexing="ij")
    s = 0.5 + 0.3 * np.sin(xx * 0.10) * np.cos(yy * 0.09)
    v = (s + sig * (0.5 + s) * rng.standard_normal(shape)).astype(np.float32)
    a = F.NoiseModel()
    b = clibs.NoiseModel()
    assert F.lib().fy_estimate_noise(F._fp(v), *shape, 5, 10.0, 0.5, C.byref(a)) == 0
    assert orc.orc_estimate_noise(clibs.fp(v), *shape, 5, 10.0, 0.5, C.byref(b)) == 0
    assert a.n_bins_used == b.n_bins_used
    assert np.allclose([a.g, a.b, a.noise_ref], [b.g, b.b, b.noise_ref], rtol=1e-12, atol=1e-18)


def test_estimate_noise_golden():
    gold = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                              "reference_vectors.npz"))
    v = gold["noise_in"]
    nm = F.NoiseModel()
    assert F.lib().fy_estimate_noise(F._fp(v), *v.shape, 5, 10.0, 0.5, C.byref(nm)) == 0
    g = gold["noise_model"]
    assert nm.n_bins_used == int(g[4])
    assert np.allclose([nm.g, nm.b, nm.noise_ref], g[:3], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("blur", [True, False])
def test_noise_aniso_vs_oracle(orc, blur):
    rng = np.random.default_rng(42)
    v = (0.5 + 0.05 * (rng.random((64, 64, 64)) - 0.5)).astype(np.float32)
    if blur:                                              # tests/test_fysics.c:257-278
        b = v.copy()
        b[1:-1] = 0.25 * v[:-2] + 0.5 * v[1:-1] + 0.25 * v[2:]
        v = b
    rz, rxy = C.c_double(), C.c_double()
    oz, oxy = C.c_double(), C.c_double()
    assert F.lib().fy_noise_aniso(F._fp(v), *v.shape, C.byref(rz), C.byref(rxy)) == 0
    assert orc.orc_noise_aniso(clibs.fp(v), *v.shape, C.byref(oz), C.byref(oxy)) == 0
    assert np.allclose([rz.value, rxy.value], [oz.value, oxy.value], rtol=1e-10)
    assert (rz.value > rxy.value + 0.1) == blur
