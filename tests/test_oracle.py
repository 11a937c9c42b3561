"""Pins the CPU restatement in oracle/ (CPU only).

Each oracle function is checked against (a) the golden vectors written by the
compiled reference (tests/golden/reference_vectors.npz), (b) the live reference
build oracle/_ref when present, and (c) independent mathematics: numpy's FFT,
brute-force box means, the nabu closed form, properties of the dithered
quantiser.  The checks are chosen so that a dropped term, a wrong sign or index
or a transposed axis in the restatement fails at least one of them.
"""
import ctypes as C
import os

import numpy as np
import pytest

import clibs

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
PHYS_A = dict(delta_beta=1000.0, energy_kev=78.0, distance_mm=220.0, pixel_um=2.4,
              unsharp_sigma=1.2, unsharp_coeff=4.0)
PHYS_B = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
              unsharp_sigma=0.8, unsharp_coeff=0.5)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ------------------------------------------------------------------ FFT (fft.c)
@pytest.mark.parametrize("n", [16, 48, 96, 192, 256, 3, 6, 12])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_golden_and_numpy(orc, n, sign):
    re, im = GOLD[f"fft_{n}_{sign}_in"]
    a, b = re.copy(), im.copy()
    orc.orc_fft1d_f32(clibs.fp(a), clibs.fp(b), n, sign)
    gr, gi = GOLD[f"fft_{n}_{sign}_out"]
    assert rel_l2(a + 1j * b, gr + 1j * gi) < 2e-6
    x = re.astype(np.float64) + 1j * im
    want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    a64, b64 = re.astype(np.float64), im.astype(np.float64)
    orc.orc_fft1d_f64(clibs.dp(a64), clibs.dp(b64), n, sign)
    assert rel_l2(a64 + 1j * b64, want) < 1e-14
    assert rel_l2(a + 1j * b, want) < 2e-6


def test_fft_naive_dft_small(orc):
    """tests/test_fysics.c:26-40 and :56-72 (n = 16 and 48 against a direct sum)."""
    for n in (16, 48):
        t = np.arange(n)
        re = (np.sin(t * 0.7) + 0.3 * np.cos(t * 1.3)).astype(np.float32)
        im = (0.2 * np.sin(t * 0.31)).astype(np.float32)
        k = t[:, None]
        dft = (np.exp(-2j * np.pi * k * t[None] / n) * (re + 1j * im)[None]).sum(1)
        a, b = re.copy(), im.copy()
        orc.orc_fft1d_f32(clibs.fp(a), clibs.fp(b), n, -1)
        assert np.max(np.abs(a - dft.real) + np.abs(b - dft.imag)) < 1e-3


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 24), (6, 4, 3)])
def test_fft3d_roundtrip_and_numpy(orc, shape):
    rng = np.random.default_rng(2)
    re = rng.random(shape).astype(np.float32)
    im = np.zeros(shape, np.float32)
    a, b = re.copy(), im.copy()
    orc.orc_fft3d_f32(clibs.fp(a), clibs.fp(b), *shape, -1)
    assert rel_l2(a + 1j * b, np.fft.fftn(re.astype(np.float64))) < 2e-6
    orc.orc_fft3d_f32(clibs.fp(a), clibs.fp(b), *shape, 1)
    orc.orc_fft3d_normalize_f32(clibs.fp(a), clibs.fp(b), *shape)
    assert np.abs(a - re).max() < 1e-4


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 24), (6, 4, 3), (3, 16, 1)])
def test_fft3d_f64_numpy(orc, shape):
    """orc_fft3d_f64: sign -1 is numpy's fftn, sign +1 the unnormalised inverse
    (N * ifftn), on complex data, so a conjugated axis or a skipped axis fails."""
    rng = np.random.default_rng(12)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n = int(np.prod(shape))
    for sign, want in ((-1, np.fft.fftn(x)), (1, np.fft.ifftn(x) * n)):
        a, b = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
        orc.orc_fft3d_f64(clibs.dp(a), clibs.dp(b), *shape, sign)
        assert rel_l2(a + 1j * b, want) < 1e-14


def test_next_fft_size(orc):
    cases = {176: 192, 128: 128, 160: 192, 97: 128, 1: 1, 2: 2, 3: 3, 5: 6, 64: 64, 96: 96,
             256: 256, 512: 512, 158: 192, 200: 256, 300: 384}
    for n, want in cases.items():
        assert orc.orc_next_fft_size(n) == want


# ------------------------------------------------------------------ transfers (paganin.c)
def test_transfers_golden_and_nabu(orc):
    p = clibs.physics(**PHYS_A)
    f = GOLD["transfer_f"]
    for name, key in (("paganin_transfer", "transfer_T"), ("unsharp_transfer", "transfer_U"),
                      ("recon_transfer", "transfer_H")):
        got = np.array([getattr(orc, "orc_" + name)(x, C.byref(p)) for x in f])
        assert np.allclose(got, GOLD[key], rtol=1e-14, atol=0)
    # nabu formula (tests/test_fysics.c:327-330)
    lam, dist, fu = 1.23984199e-3 / 78.0, 220000.0, 0.1 / 2.4
    assert abs(orc.orc_paganin_transfer(0.1, C.byref(p)) - 1.0 / (1.0 + 1000.0 * lam * dist * np.pi * fu * fu)) < 1e-12
    scales = [orc.orc_auto_deltabeta_scale(C.byref(clibs.physics(**k))) for k in (PHYS_A, PHYS_B)]
    assert np.allclose(scales, GOLD["dbscale"], rtol=1e-14)
    halos = [orc.orc_kernel_halo(C.byref(clibs.physics(**k))) for k in (PHYS_A, PHYS_B)]
    assert halos == list(GOLD["halo"])


# ------------------------------------------------------------------ deconvolution (paganin.c)
def _lut_weight_numpy(fr2, us, lut):
    t = fr2 * us
    ti = np.minimum(t.astype(np.int64), 4095)
    return lut[ti] + (lut[ti + 1] - lut[ti]) * (t - ti)


def deconv_numpy_f64(x, weight_fn, cap, s, use_lut=True):
    """Independent restatement through numpy's FFT: reflect-pad to the supported
    size, full complex 3-D FFT, radial weight through the 4097-entry LUT over
    u = s*fz^2 + fy^2 + fx^2 evaluated in float64, inverse, crop."""
    nz, ny, nx = x.shape
    nxt = lambda n: min([m for m in [2 ** k for k in range(14)] + [3 * 2 ** k for k in range(13)] if m >= n])
    pz, py, px = nxt(nz), nxt(ny), nxt(nx)
    refl = lambda n, p: np.array([i if i < n else 2 * n - 1 - i for i in range(p)]) if n > 1 else np.zeros(p, int)
    xp = x.astype(np.float64)[np.ix_(refl(nz, pz), refl(ny, py), refl(nx, px))]
    fz = np.fft.fftfreq(pz)[:, None, None] ** 2 * s
    fy = np.fft.fftfreq(py)[None, :, None] ** 2
    fx = (np.arange(px // 2 + 1) / px)[None, None, :] ** 2
    umax = fz.max() + fy.max() + fx.max()
    us = 4096 / umax
    knots = np.array([weight_fn(np.sqrt(i / us)) for i in range(4097)])
    if cap > 0:
        knots = np.minimum(knots, cap)
    u = fz + fy + fx
    w = _lut_weight_numpy(u, us, knots) if use_lut else np.minimum(weight_fn(np.sqrt(u)), cap if cap > 0 else np.inf)
    spec = np.fft.rfftn(xp) * w
    return np.fft.irfftn(spec, s=(pz, py, px), axes=(0, 1, 2))[:nz, :ny, :nx]


def wiener_np(p, reg):
    lam = 1.23984199e-3 / p["energy_kev"]
    def w(fr):
        fu = fr / p["pixel_um"]
        t = 1.0 / (1.0 + p["delta_beta"] * lam * p["distance_mm"] * 1000.0 * np.pi * fu * fu)
        un = 1.0 + p["unsharp_coeff"] * (1.0 - np.exp(-2 * np.pi ** 2 * p["unsharp_sigma"] ** 2 * fr * fr))
        h = np.maximum(t * un, 1e-4)
        return h / (h * h + reg)
    return w


def test_deconv_golden(orc):
    x, y = GOLD["deconv_in"], GOLD["deconv_out"]
    p = clibs.physics(**PHYS_A)
    a = np.empty_like(x)
    assert orc.orc_deconvolve(clibs.fp(x), clibs.fp(a), *x.shape, C.byref(p), 0.015) == 0
    assert rel_l2(a, y) < 1e-6
    # tests/test_fysics.c:237: half-spectrum path == full-complex path within LUT tolerance
    full = deconv_numpy_f64(x, wiener_np(PHYS_A, 0.015), 8.0, 1.0, use_lut=False)
    assert np.abs(a - full).max() < 1e-3 * max(np.abs(full).max(), 1)


@pytest.mark.parametrize("shape", [(22, 27, 30), (5, 7, 9), (1, 4, 6), (3, 1, 5), (12, 6, 1)])
def test_deconv_f64_vs_numpy_lut(orc, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.random(shape).astype(np.float32)
    p = clibs.physics(**PHYS_A)
    got = np.empty(shape, np.float64)
    assert orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(got), *shape, C.byref(p), 0.015, 0, 8.0, 1.0) == 0
    want = deconv_numpy_f64(x, wiener_np(PHYS_A, 0.015), 8.0, 1.0)
    assert rel_l2(got, want) < 1e-12
    f32 = np.empty(shape, np.float32)
    assert orc.orc_deconvolve(clibs.fp(x), clibs.fp(f32), *shape, C.byref(p), 0.015) == 0
    assert rel_l2(f32, want) < 2e-6


def test_deconv_aniso_and_matched_golden(orc):
    x, y = GOLD["aniso_in"], GOLD["aniso_out"]
    p = clibs.physics(**PHYS_B)
    a = np.empty_like(x)
    assert orc.orc_deconv_radial_f32(clibs.fp(x), clibs.fp(a), *x.shape, C.byref(p), 0.015, 0, 8.0, 1.17 ** 2) == 0
    assert rel_l2(a, y) < 1e-6
    d = np.empty(x.shape, np.float64)
    orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(d), *x.shape, C.byref(p), 0.015, 0, 8.0, 1.17 ** 2)
    assert rel_l2(d, deconv_numpy_f64(x, wiener_np(PHYS_B, 0.015), 8.0, 1.17 ** 2)) < 1e-12
    x, y = GOLD["matched_in"], GOLD["matched_out"]
    pm = clibs.physics(**PHYS_B, psf_sigma_vox=1.1, psf_sigma_z_vox=1.3)
    a = np.empty_like(x)
    assert orc.orc_deconvolve_matched(clibs.fp(x), clibs.fp(a), *x.shape, C.byref(pm), 0.05) == 0
    assert rel_l2(a, y) < 1e-6
    assert abs(orc.orc_matched_fz2_scale(C.byref(pm)) - (1.3 / 1.1) ** 2) < 1e-15


def test_deconv_lut_knots(orc):
    """LUT entries are the weight at sqrt(i/us), capped, rounded to fp32."""
    p = clibs.physics(**PHYS_A)
    lut = np.empty(4097, np.float32)
    us = C.c_double()
    orc.orc_deconv_lut(32, 24, 48, C.byref(p), 0.015, 0, 8.0, 1.0, clibs.fp(lut), C.byref(us))
    umax = 0.25 + 0.25 + 0.25
    assert abs(us.value - 4096 / umax) < 1e-9
    w = wiener_np(PHYS_A, 0.015)
    want = np.minimum([w(np.sqrt(i / us.value)) for i in range(4097)], 8.0).astype(np.float32)
    assert np.array_equal(lut, want)


# ------------------------------------------------------------------ guided / box (guided.c)
def box_mean_numpy(x, r):
    """Brute force: mean over the in-bounds part of each (2r+1)^3 window."""
    nz, ny, nx = x.shape
    c = np.pad(x.astype(np.float64), r)
    o = np.pad(np.ones(x.shape), r)
    s = np.zeros(x.shape)
    cnt = np.zeros(x.shape)
    for dz in range(2 * r + 1):
        for dy in range(2 * r + 1):
            for dx in range(2 * r + 1):
                s += c[dz:dz + nz, dy:dy + ny, dx:dx + nx]
                cnt += o[dz:dz + nz, dy:dy + ny, dx:dx + nx]
    return s / cnt


def guided_numpy(x, r, eps):
    x = x.astype(np.float64)
    m = box_mean_numpy(x, r)
    v = np.maximum(box_mean_numpy(x * x, r) - m * m, 0)
    a = v / (v + eps)
    b = m * (1 - a)
    return box_mean_numpy(a, r) * x + box_mean_numpy(b, r)


@pytest.mark.parametrize("r", [1, 2, 3])
def test_box_mean(orc, r):
    g = GOLD["guided_in"]
    a = np.empty_like(g)
    orc.orc_box_mean_f32(clibs.fp(g), clibs.fp(a), *g.shape, r)
    assert rel_l2(a, GOLD[f"box_r{r}"]) < 1e-6
    d = np.empty(g.shape)
    orc.orc_box_mean_f64(clibs.dp(g.astype(np.float64)), clibs.dp(d), *g.shape, r)
    assert rel_l2(d, box_mean_numpy(g, r)) < 1e-13


def test_guided_golden_and_bruteforce(orc):
    g = GOLD["guided_in"]
    a = np.empty_like(g)
    orc.orc_guided_exact_f32(clibs.fp(g), clibs.fp(a), *g.shape, 2, 4e-4)
    assert rel_l2(a, GOLD["guided_exact"]) < 1e-6
    orc.orc_guided_fast_f32(clibs.fp(g), clibs.fp(a), *g.shape, 2, 4e-4, 2)
    assert rel_l2(a, GOLD["guided_fast"]) < 1e-6
    d = np.empty(g.shape)
    orc.orc_guided_exact_f64(clibs.dp(g.astype(np.float64)), clibs.dp(d), *g.shape, 2, 4e-4)
    assert rel_l2(d, guided_numpy(g, 2, 4e-4)) < 1e-12


def test_guided_fast_f64_structure(orc):
    """fast(s) == exact guided coefficients on the decimated grid, trilinear
    upsampled (guided.c:152-228), re-derived with numpy."""
    rng = np.random.default_rng(4)
    x = rng.random((9, 10, 11))
    s, r, eps = 2, 2, 1e-2
    xl = x[::s, ::s, ::s]
    m = box_mean_numpy(xl, 1)
    v = np.maximum(box_mean_numpy(xl * xl, 1) - m * m, 0)
    a = v / (v + eps)
    b = m * (1 - a)
    ma, mb = box_mean_numpy(a, 1), box_mean_numpy(b, 1)

    def coords(n, ln):
        u = np.arange(n, dtype=np.float32) / np.float32(s)
        i0 = np.minimum(u.astype(int), ln - 2)
        return i0, np.minimum(u - i0, 1.0).astype(np.float64)
    (z0, fz), (y0, fy), (x0, fx) = coords(9, 5), coords(10, 5), coords(11, 6)

    def tri(f):
        out = np.zeros(x.shape)
        for dz, wz in ((0, 1 - fz), (1, fz)):
            for dy, wy in ((0, 1 - fy), (1, fy)):
                for dx, wx in ((0, 1 - fx), (1, fx)):
                    out += (wz[:, None, None] * wy[None, :, None] * wx[None, None, :] *
                            f[np.ix_(z0 + dz, y0 + dy, x0 + dx)])
        return out
    want = tri(ma) * x + tri(mb)
    got = np.empty(x.shape)
    orc.orc_guided_fast_f64(clibs.dp(x), clibs.dp(got), *x.shape, r, eps, s)
    assert rel_l2(got, want) < 1e-13


# ------------------------------------------------------------------ quantiser, conversions
def test_quant_golden(orc):
    v, c = GOLD["quant_v"], GOLD["quant_c"]
    q = np.array([orc.orc_quant_u8(float(a), int(z), int(y), int(x), 0) for a, (z, y, x) in zip(v, c)])
    assert np.array_equal(q, GOLD["quant_q"])
    q = np.array([orc.orc_quant_u8(float(a), int(z), int(y), int(x), 1) for a, (z, y, x) in zip(v, c)])
    assert np.array_equal(q, GOLD["quant_q_nd"])


def test_dither_is_unbiased(orc):
    """floor(v*255 + d) with d ~ U[0,1) has mean v*255 (pipeline.c:49-52)."""
    v = 0.4237
    q = np.array([orc.orc_quant_u8(v, z, y, 7, 0) for z in range(60) for y in range(60)])
    d = np.array([orc.orc_dither01(z, y, 7) for z in range(60) for y in range(60)])
    assert abs(q.mean() - v * 255) < 0.02
    assert d.min() >= 0 and d.max() < 1 and abs(d.mean() - 0.5) < 0.02


def test_conversions_roundtrip(orc):
    """tests/test_fysics.c:439-453."""
    u = np.arange(256, dtype=np.uint8)
    ph = np.empty(256, np.float32)
    orc.orc_u8_to_phys(clibs.up(u), clibs.fp(ph), 256, -0.03, 0.21)
    back = np.empty(256, np.uint8)
    orc.orc_phys_to_u8(clibs.fp(ph), clibs.up(back), 256, -0.03, 0.21)
    assert np.array_equal(back, u)
    assert abs(ph[0] + 0.03) < 1e-6 and abs(ph[255] - 0.21) < 1e-6
    assert abs(orc.orc_phys_to_u8_level(0.0, -0.03, 0.145) - orc.orc_phys_to_u8_level(0.0, -0.03, 0.21)) > 1.0


def test_float_to_u8_golden_and_properties(orc):
    """orc_float_to_u8 against the reference's fy_float_to_u8 (stream.c:31-36) on
    the rounding edges of every level, out-of-range and random values; and what
    the rounding fixes: u8 -> float -> u8 is the identity, and the map is
    monotone, 0 below 0.5/255 and 255 from 254.5/255 up."""
    fu, fq = GOLD["f2u8_in"], GOLD["f2u8_out"]
    got = np.empty(fu.size, np.uint8)
    orc.orc_float_to_u8(clibs.fp(fu), clibs.up(got), fu.size)
    assert np.array_equal(got, fq)
    u = np.arange(256, dtype=np.uint8)
    f = np.empty(256, np.float32)
    orc.orc_u8_to_float(clibs.up(u), clibs.fp(f), 256)
    back = np.empty(256, np.uint8)
    orc.orc_float_to_u8(clibs.fp(f), clibs.up(back), 256)
    assert np.array_equal(back, u)
    xs = np.sort(np.concatenate([fu, np.linspace(-0.1, 1.1, 5001, dtype=np.float32)])).astype(np.float32)
    q = np.empty(xs.size, np.uint8)
    orc.orc_float_to_u8(clibs.fp(xs), clibs.up(q), xs.size)
    assert np.all(np.diff(q.astype(int)) >= 0)
    assert np.all(q[xs < 0.49 / 255] == 0) and np.all(q[xs > 254.51 / 255] == 255)


def test_zdrift_apply_golden_and_properties(orc):
    """orc_zdrift_apply against the reference's fy_zdrift_apply (zdrift.c:64-71):
    slab z of the chunk is scaled by factor[z0 + z]; unit factors leave it
    unchanged and a factor of 2 doubles exactly (an index off by one or a
    dropped z0 fails the golden)."""
    zc, zf, za = GOLD["zapply_in"], GOLD["zapply_factor"], GOLD["zapply_out"]
    a = zc.copy()
    orc.orc_zdrift_apply(clibs.fp(a), *zc.shape, 5, clibs.fp(zf))
    assert np.array_equal(a, za)
    one = np.ones(zf.size, np.float32)
    b = zc.copy()
    orc.orc_zdrift_apply(clibs.fp(b), *zc.shape, 5, clibs.fp(one))
    assert np.array_equal(b, zc)
    two = np.ones(zf.size, np.float32)
    two[5 + 3] = 2.0
    c = zc.copy()
    orc.orc_zdrift_apply(clibs.fp(c), *zc.shape, 5, clibs.fp(two))
    assert np.array_equal(c[3], 2 * zc[3]) and np.array_equal(np.delete(c, 3, 0), np.delete(zc, 3, 0))


# ------------------------------------------------------------------ rings (dering.c)
def test_dering_golden(orc):
    vol = GOLD["dering_vol"]
    nz, ny, nx = vol.shape
    d = clibs.Dering()
    t = clibs.Dering()
    assert orc.orc_dering_init(C.byref(d), nz, ny, nx, 47.5, 55.0, 4, 8) == 0
    for slab in range(2):
        assert orc.orc_dering_tile_init(C.byref(t), C.byref(d)) == 0
        sub = np.ascontiguousarray(vol[4 * slab:4 * slab + 4])
        orc.orc_dering_accumulate_u8(C.byref(t), clibs.up(sub), 0, 0, 4, ny, nx, 1)
        orc.orc_dering_merge_tile(C.byref(d), slab, C.byref(t))
        orc.orc_dering_free(C.byref(t))
    nb = d.nslab * d.ns * d.nr
    assert np.array_equal(np.ctypeslib.as_array(d.sum, (nb,)), GOLD["dering_sum"])
    assert np.array_equal(np.ctypeslib.as_array(d.cnt, (nb,)), GOLD["dering_cnt"])
    n = orc.orc_dering_finalize(C.byref(d), 9, 20, 1.0, 6.0)
    assert [n, d.have_rings, d.nr, d.nslab] == list(GOLD["dering_found"])
    ring = np.ctypeslib.as_array(d.ring, (d.nslab * d.nr,))
    # votes and radii are integer decisions: identical; the amplitudes are fp32
    # medians whose last bits depend on the reference build's fast-math division
    assert np.array_equal(ring != 0, GOLD["dering_ring"] != 0)
    assert np.allclose(ring, GOLD["dering_ring"], rtol=1e-5, atol=0)
    f = (vol * np.float32(1 / 255)).astype(np.float32)
    orc.orc_dering_apply(C.byref(d), clibs.fp(f), 0, 0, 0, nz, ny, nx, 1.0 / 255.0)
    assert np.abs(f - GOLD["dering_applied"]).max() < 1e-7
    orc.orc_dering_free(C.byref(d))


def test_dering_counts_bruteforce(orc):
    """Counts per (sector, radius) bin re-derived with numpy's sqrt/arctan2."""
    vol = GOLD["dering_vol"][:2]
    nz, ny, nx = vol.shape
    d = clibs.Dering()
    orc.orc_dering_init(C.byref(d), nz, ny, nx, 47.5, 55.0, 4, 8)
    t = clibs.Dering()
    orc.orc_dering_tile_init(C.byref(t), C.byref(d))
    orc.orc_dering_accumulate_u8(C.byref(t), clibs.up(np.ascontiguousarray(vol)), 0, 0, nz, ny, nx, 2)
    yy, xx = np.meshgrid(np.arange(0, ny, 2) - 47.5, np.arange(0, nx, 2) - 55.0, indexing="ij")
    ri = np.floor(np.sqrt(yy ** 2 + xx ** 2) + 0.5).astype(int)
    q = np.minimum(((np.arctan2(yy, xx) + np.pi) * (8 / (2 * np.pi))).astype(int), 7)
    cnt = np.zeros(8 * d.nr, np.int64)
    sm = np.zeros(8 * d.nr, np.int64)
    for z in range(nz):
        v = vol[z, ::2, ::2].astype(np.int64)
        ok = (v > 0) & (ri < d.nr)
        np.add.at(cnt, (q * d.nr + ri)[ok], 1)
        np.add.at(sm, (q * d.nr + ri)[ok], v[ok])
    assert np.array_equal(np.ctypeslib.as_array(t.cnt, (8 * d.nr,)), cnt)
    assert np.array_equal(np.ctypeslib.as_array(t.sum, (8 * d.nr,)), sm.astype(np.float32))
    orc.orc_dering_free(C.byref(t))
    orc.orc_dering_free(C.byref(d))


# ------------------------------------------------------------------ noise (noise.c)
def test_noise_golden(orc):
    v = GOLD["noise_in"]
    nm = clibs.NoiseModel()
    assert orc.orc_estimate_noise(clibs.fp(v), *v.shape, 5, 10.0, 0.5, C.byref(nm)) == 0
    g = GOLD["noise_model"]
    assert nm.n_bins_used == int(g[4])
    assert np.allclose([nm.g, nm.b, nm.noise_ref], g[:3], rtol=1e-9, atol=1e-15)
    a = GOLD["aniso_vol"]
    rz, rxy = C.c_double(), C.c_double()
    assert orc.orc_noise_aniso(clibs.fp(a), *a.shape, C.byref(rz), C.byref(rxy)) == 0
    assert np.allclose([rz.value, rxy.value], GOLD["aniso_rho"], rtol=1e-9)
    assert orc.orc_guided_eps_for_noise(0.01) == pytest.approx(9e-4)
    assert orc.orc_guided_eps_for_noise(0.0) == 1e-8


# ------------------------------------------------------------------ chain (pipeline.c)
def _cfg(**kw):
    c = clibs.PipelineCfg()
    base = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3, unsharp_sigma=0.8,
                unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2, do_deconv=1, guided_eps=1.2e-3)
    base.update(kw)
    for k, v in base.items():
        setattr(c, k, v)
    return c


def test_chain_golden(orc):
    u8 = GOLD["chain_u8"]
    q = np.empty((20, 22, 24), np.uint8)
    air = C.c_int(0)
    assert orc.orc_process_buffer(C.byref(_cfg()), clibs.up(u8), 10, 20, 30, *u8.shape, 14, 24, 34, 20, 22, 24,
                                  300, 400, clibs.up(q), C.byref(air)) == 0
    diff = np.abs(q.astype(int) - GOLD["chain_out"].astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


CHAIN_CFGS = [dict(), dict(no_dither=1), dict(do_deconv=0), dict(guided_subsample=1),
              dict(use_matched_deconv=1, psf_sigma_vox=1.0, psf_z_ratio=1.2),
              dict(have_norm=1, norm_lo=30, norm_hi=210), dict(have_dec_range=1, dec_lo=0.05, dec_hi=0.8),
              dict(have_eps_r=1, eps_fn_a=0.01, eps_fn_b=0.0001, eps_fn_med=0.012)]


@pytest.mark.parametrize("i", [0, 2, 3, 4, 5, 6])
def test_chain_f64_vs_f32_chain(orc, i):
    """orc_chain_cube_f64 (deconvolution and guided filter in fp64) against the
    fp32 chain, which the goldens and the live reference pin: the float tiles
    agree to the fp32 rounding level (relL2 <= 2e-6, far below what a dropped
    term, a wrong sign or index changes) and the u8 tiles to +-1, with a
    z-drift table and an inner tile offset in the region."""
    from fysics_amd import phantom
    rng = np.random.default_rng(40 + i)
    u8 = phantom.noisy_u8(phantom.forward(phantom.sheets((26, 28, 30), rng, 5), **phantom.CONFIG1_PHYS), rng)
    cfg = _cfg(**CHAIN_CFGS[i])
    P = clibs.CubeParams()
    assert orc.orc_cube_params_from_cfg(C.byref(cfg), 24, 34, 20, 22, 400, 500, C.byref(P)) == 0
    zf = rng.uniform(0.9, 1.1, 64).astype(np.float32)
    P.zdrift_factor = zf.ctypes.data_as(clibs.c_float_p)
    P.vol_z = 64
    args = (10, 20, 30, *u8.shape, 13, 24, 34, 18, 20, 22)
    q32 = np.empty((18, 20, 22), np.uint8)
    f32 = np.empty(u8.shape, np.float32)
    air = C.c_int(0)
    assert orc.orc_chain_cube(C.byref(P), clibs.up(u8), *args, clibs.up(q32), C.byref(air), clibs.fp(f32)) == 0
    q64 = np.empty_like(q32)
    f64 = np.empty(u8.shape, np.float64)
    assert orc.orc_chain_cube_f64(C.byref(P), clibs.up(u8), *args, clibs.up(q64), clibs.dp(f64)) == 0
    assert rel_l2(f32, f64) <= 2e-6
    du = np.abs(q32.astype(int) - q64.astype(int))
    assert du.max() <= 1 and (du > 0).mean() < 2e-3
    # without the z-drift table the tile changes: the factor is applied
    P.zdrift_factor = None
    f0 = np.empty(u8.shape, np.float64)
    assert orc.orc_chain_cube_f64(C.byref(P), clibs.up(u8), *args, clibs.up(q64), clibs.dp(f0)) == 0
    assert rel_l2(f0, f64) > 1e-3


@pytest.mark.parametrize("i", range(len(CHAIN_CFGS)))
def test_chain_vs_live_reference(orc, ref, i):
    from fysics_amd import phantom
    rng = np.random.default_rng(i)
    u8 = phantom.noisy_u8(phantom.forward(phantom.sheets((26, 28, 30), rng, 5), **phantom.CONFIG1_PHYS), rng)
    cfg = _cfg(**CHAIN_CFGS[i])
    args = (10, 20, 30, *u8.shape, 13, 24, 34, 18, 20, 22, 300, 400)
    a = np.empty((18, 20, 22), np.uint8)
    b = np.empty((18, 20, 22), np.uint8)
    air = C.c_int(0)
    assert orc.orc_process_buffer(C.byref(cfg), clibs.up(u8), *args, clibs.up(a), C.byref(air)) == 0
    assert ref.fy_process_buffer(C.byref(cfg), clibs.up(u8), *args, clibs.up(b), C.byref(air)) == 0
    diff = np.abs(a.astype(int) - b.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 2e-3
    for y0, x0 in ((0, 0), (100, 350), (390, 10)):
        assert orc.orc_eps_for_tile(C.byref(cfg), 150.0, 200.0, y0, x0, 32, 32) == \
            ref.ref_eps_for_tile(C.byref(cfg), 150.0, 200.0, y0, x0, 32, 32)


def test_chain_all_air(orc):
    u8 = np.zeros((10, 10, 10), np.uint8)
    q = np.full((8, 8, 8), 7, np.uint8)
    air = C.c_int(0)
    assert orc.orc_process_buffer(C.byref(_cfg()), clibs.up(u8), 0, 0, 0, 10, 10, 10, 1, 1, 1, 8, 8, 8,
                                  10, 10, clibs.up(q), C.byref(air)) == 0
    assert air.value == 1 and not q.any()


@pytest.mark.parametrize("shape", [(17, 13, 29), (8, 32, 6)])
def test_deconv_guided_vs_live_reference(orc, ref, shape):
    rng = np.random.default_rng(9)
    x = rng.random(shape).astype(np.float32)
    p = clibs.physics(**PHYS_B)
    a = np.empty_like(x); b = np.empty_like(x)
    orc.orc_deconvolve(clibs.fp(x), clibs.fp(a), *shape, C.byref(p), 0.02)
    ref.fy_deconvolve(clibs.fp(x), clibs.fp(b), *shape, C.byref(p), 0.02)
    assert rel_l2(a, b) < 1e-6
    ws = np.empty(4 * x.size, np.float32)
    orc.orc_guided_fast_f32(clibs.fp(x), clibs.fp(a), *shape, 2, 1e-3, 3)
    ref.fy_guided_denoise_fast_ws(clibs.fp(x), clibs.fp(b), *shape, 2, 1e-3, 3, clibs.fp(ws))
    assert rel_l2(a, b) < 1e-6


# ------------------------------------------------------------------ pass-1 calibration
def _hist_of(orc, data, chunks):
    hs = clibs.HistState()
    for a, b in chunks:
        part = np.ascontiguousarray(data[a:b])
        orc.orc_hist_accumulate_u8(C.byref(hs), clibs.up(part), part.size)
    return hs


def test_hist_golden_and_bincount(orc):
    u8 = GOLD["hist_u8"]
    hs = _hist_of(orc, u8, [(0, 1000), (1000, 1000), (1000, u8.size)])
    counts = np.array(hs.hist[:], np.int64)
    assert np.array_equal(counts, GOLD["hist_counts"])
    assert np.array_equal(counts, np.bincount(u8, minlength=256))
    assert hs.total == GOLD["hist_total"][0] == u8.size
    for q, want, want_in in zip(GOLD["hist_pcts"], GOLD["hist_pct"], GOLD["hist_pct_inner"]):
        assert orc.orc_hist_percentile_u8(C.byref(hs), q) == want
        assert orc.orc_hist_percentile_u8_inner(C.byref(hs), q) == want_in
        # first bin whose running count reaches (long)(q/100 * total)
        cs = np.cumsum(counts)
        assert want == min(int(np.searchsorted(cs, int(q / 100.0 * u8.size))), 255)
        inner = counts[1:255]
        tgt = int(q / 100.0 * inner.sum())
        assert want_in == min(1 + int(np.searchsorted(np.cumsum(inner), tgt)), 254)


def test_hist_empty_and_inner_empty(orc):
    hs = clibs.HistState()
    assert orc.orc_hist_percentile_u8(C.byref(hs), 50.0) == 0
    assert orc.orc_hist_percentile_u8_inner(C.byref(hs), 50.0) == 0
    z = np.zeros(10, np.uint8)
    orc.orc_hist_accumulate_u8(C.byref(hs), clibs.up(z), 10)
    assert orc.orc_hist_percentile_u8_inner(C.byref(hs), 50.0) == 0      # bins 1..254 empty


def test_norm_apply_golden(orc):
    nin = GOLD["norm_in"]
    for k, (lo, hi) in enumerate(((10, 200), (0, 255), (90, 90), (200, 10))):
        o = np.empty(nin.size, np.float32)
        orc.orc_norm_apply_u8(clibs.up(nin), clibs.fp(o), nin.size, lo, hi)
        assert np.array_equal(o, GOLD["norm_out"][k])
        inv = np.float32(1.0) / np.float32(hi - lo) if hi > lo else np.float32(1.0)
        want = np.clip((nin.astype(np.float32) - np.float32(lo)) * inv, 0, 1)
        assert np.array_equal(o, want)


def test_zdrift_accumulate_finalize_golden(orc):
    zch = GOLD["zdrift_chunk"]
    zsum = np.linspace(0.0, 5.0, 12)
    zcnt = np.arange(12, dtype=np.int64) * 3
    orc.orc_zdrift_accumulate(clibs.fp(zch), 8, 20, 24, 3, clibs.dp(zsum), clibs.lp(zcnt), 0.3)
    assert np.array_equal(zcnt, GOLD["zdrift_counts"])
    assert np.allclose(zsum, GOLD["zdrift_sums"], rtol=1e-13, atol=0)
    # independent: per-slice masked sums
    m = zch > np.float32(0.3)
    base = np.linspace(0.0, 5.0, 12)
    want = base.copy()
    want[3:11] += (zch.astype(np.float64) * m).sum(axis=(1, 2))
    assert np.allclose(zsum, want, rtol=1e-12, atol=0)
    fac = np.empty(12, np.float32)
    orc.orc_zdrift_finalize(clibs.dp(GOLD["zdrift_sums"]), clibs.lp(np.ascontiguousarray(GOLD["zdrift_counts"])),
                            12, clibs.fp(fac))
    assert np.array_equal(fac, GOLD["zdrift_factor"])


def test_zdrift_finalize_closed_form(orc):
    # constant papyrus mean with one empty slice: every factor is 1; a linear
    # ramp is flattened to the global mean by the windowed means
    nz = 30
    counts = np.full(nz, 10, np.int64)
    counts[7] = 0
    sums = 10 * 0.5 * np.ones(nz)
    sums[7] = 0
    fac = np.empty(nz, np.float32)
    orc.orc_zdrift_finalize(clibs.dp(sums), clibs.lp(counts), nz, clibs.fp(fac))
    assert np.array_equal(fac, np.ones(nz, np.float32))
    mean = np.linspace(0.4, 0.6, nz)
    sums = mean * 10
    counts = np.full(nz, 10, np.int64)
    orc.orc_zdrift_finalize(clibs.dp(sums), clibs.lp(counts), nz, clibs.fp(fac))
    win = nz // 20 + 1
    g = sums.sum() / counts.sum()
    sm = np.array([mean[max(0, z - win):z + win + 1].mean() for z in range(nz)])
    assert np.allclose(fac, (g / sm).astype(np.float32), rtol=1e-6)


def test_zdrift_u8_sub_golden_and_exact(orc):
    zu8 = GOLD["zdrift_u8"]
    s = np.zeros(9)
    c = np.zeros(9, np.int64)
    orc.orc_zdrift_u8_sub(clibs.up(zu8), 6, 19, 23, 2, clibs.dp(s), clibs.lp(c))
    assert np.array_equal(c, GOLD["zdrift_u8_counts"])
    assert np.array_equal(s, GOLD["zdrift_u8_sums"])          # sums of multiples of 2^-27: exact
    v = zu8[:, ::4, ::4].astype(np.float32) * np.float32(1.0 / 255.0)
    m = v > np.float32(0.10)
    assert np.array_equal(c[2:8], m.sum(axis=(1, 2)))
    assert np.array_equal(s[2:8], (v.astype(np.float64) * m).sum(axis=(1, 2)))


def _flat_noise_np(v, blk):
    if blk < 4:
        blk = 8
    nz, ny, nx = (d // blk for d in v.shape)
    if nz * ny * nx == 0:
        return 0.0
    b = v[:nz * blk, :ny * blk, :nx * blk].reshape(nz, blk, ny, blk, nx, blk)
    s1 = b.astype(np.float64).sum(axis=(1, 3, 5))
    s2 = (b * b).astype(np.float64).sum(axis=(1, 3, 5))       # fp32 squares
    c = blk ** 3
    mean = s1 / c
    var = s2 / c - mean * mean
    sd = np.sqrt(np.maximum(var, 0))[mean > 0.1]
    return float(np.sort(sd)[sd.size // 10]) if sd.size else 0.0


def test_flat_noise_golden_and_blocks(orc):
    fv, fg = GOLD["flat_u8vol"], GOLD["flat_gvol"]
    got = [orc.orc_flat_noise(clibs.fp(fv), 40, 40, 48, b) for b in (8, 5, 3, 16, 41)] + \
          [orc.orc_flat_noise(clibs.fp(fg), 33, 34, 35, b) for b in (8, 4)]
    # the reference build (-Ofast) divides by the block count through a
    # reciprocal: it agrees to ~1e-14
    assert np.allclose(got, GOLD["flat_noise"], rtol=1e-12, atol=0)
    for b, g in zip((8, 5, 3, 16, 41), got[:5]):
        assert g == _flat_noise_np(fv, b)
    for b, g in zip((8, 4), got[5:]):
        assert abs(g - _flat_noise_np(fg, b)) <= 1e-12 * max(g, 1e-30)


# ------------------------------------------------------------------ air-zero
def test_valley_depth_golden(orc):
    for h, want in zip(GOLD["valley_hists"], GOLD["valley_out"]):
        h = np.ascontiguousarray(h, np.int64)
        dk, lt, vl = C.c_int(-7), C.c_int(-7), C.c_int(-7)
        d = orc.orc_valley_depth(h.ctypes.data_as(C.POINTER(C.c_long)), C.byref(dk), C.byref(lt), C.byref(vl))
        assert [d, dk.value, lt.value, vl.value] == list(want)


def test_valley_depth_closed_form(orc):
    # two separated boxes of counts: peaks at their centres' smoothed maxima, the
    # valley in the empty gap, depth 1 - 0/min
    h = np.zeros(256, np.int64)
    h[50:55] = 100
    h[150:155] = 80
    dk, lt, vl = C.c_int(), C.c_int(), C.c_int()
    d = orc.orc_valley_depth(h.ctypes.data_as(C.POINTER(C.c_long)), C.byref(dk), C.byref(lt), C.byref(vl))
    assert (dk.value, lt.value) == (52, 152)
    assert 55 <= vl.value <= 149 and d == 1.0


def _despeckle_np(f, min_size):
    from scipy import ndimage
    lab, n = ndimage.label(f > 0)                       # 6-connectivity (the default structure)
    out = f.copy()
    nz, ny, nx = f.shape
    for k, sl in enumerate(ndimage.find_objects(lab), start=1):
        m = lab[sl] == k
        touches = sl[0].start == 0 or sl[1].start == 0 or sl[2].start == 0 or sl[0].stop == nz or \
            sl[1].stop == ny or sl[2].stop == nx
        if not touches and m.sum() < min_size:
            out[sl][m] = 0
    return out


@pytest.mark.parametrize("ms", [3, 8, 40])
def test_despeckle_golden_and_scipy(orc, ms):
    f = GOLD["despeckle_in"].copy()
    orc.orc_air_cc_despeckle(clibs.fp(f), 18, 20, 22, ms)
    assert np.array_equal(f, GOLD[f"despeckle_out_{ms}"])
    assert np.array_equal(f, _despeckle_np(GOLD["despeckle_in"], ms))


def air_params(case):
    P = clibs.CubeParams()
    P.air_zero = 1
    P.air_cut_u8, P.air_cut_band = int(case[0]), int(case[1])
    P.air_cut_aggr, P.air_thresh = float(case[2]), float(case[3])
    P.air_min_component, P.scratch_passes = int(case[4]), int(case[5])
    return P


@pytest.mark.parametrize("k", range(5))
def test_air_zero_golden(orc, k):
    """The reference build contracts the upsample blend into FMAs, so a voxel whose
    smoothed value sits on the cut may fall either way: at most 0.1% of the zero
    mask differs, every other voxel is identical."""
    f = GOLD["air_f"].copy()
    P = air_params(GOLD[f"air_case_{k}"])
    orc.orc_air_zero(clibs.fp(f), clibs.fp(GOLD["air_orig"]), 30, 34, 38, C.byref(P))
    want = GOLD[f"air_out_{k}"]
    diff = f != want
    assert diff.mean() <= 1e-3
    assert np.all((f[diff] == 0) | (want[diff] == 0))


def test_air_zero_thin_tile_golden(orc):
    f = np.ascontiguousarray(GOLD["air_f"][:2]).copy()
    P = clibs.CubeParams()
    P.air_zero, P.air_cut_u8 = 1, -1
    orc.orc_air_zero(clibs.fp(f), clibs.fp(np.ascontiguousarray(GOLD["air_orig"][:2])), 2, 34, 38, C.byref(P))
    assert np.array_equal(f, GOLD["air_out_thin"])


# ------------------------------------------------------------------ LOD downscale
def _ds_np(v, method, alpha):
    nz, ny, nx = v.shape
    oz, oy, ox = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    pad = np.pad(v, ((0, 2 * oz - nz), (0, 2 * oy - ny), (0, 2 * ox - nx)), mode="edge")
    c = pad.reshape(oz, 2, oy, 2, ox, 2).transpose(0, 2, 4, 1, 3, 5).reshape(oz, oy, ox, 8).astype(np.int32)
    mean = c.sum(-1).astype(np.float32) * np.float32(0.125)
    if method == 0:
        v = mean
    else:
        dev = np.abs(c.astype(np.float32) - mean[..., None])
        best = np.take_along_axis(c, dev.argmax(-1)[..., None], -1)[..., 0].astype(np.float32)
        v = mean + np.float32(alpha) * (best - mean)
    q = np.where(v < 0, 0, np.where(v > 255, 255, v + np.float32(0.5))).astype(np.uint8)
    return np.where(c.max(-1) == 0, 0, q).astype(np.uint8)


@pytest.mark.parametrize("mth,al", [(0, 0.0), (1, 0.5), (1, 0.8)])
def test_downscale_golden_and_numpy(orc, mth, al):
    dv = GOLD["ds_in"]
    o = np.zeros((5, 7, 9), np.uint8)
    ox, oy, oz = C.c_int(), C.c_int(), C.c_int()
    orc.orc_downscale2x(clibs.up(dv), 17, 14, 9, clibs.up(o), C.byref(ox), C.byref(oy), C.byref(oz), mth,
                        C.c_float(al))
    assert (oz.value, oy.value, ox.value) == (5, 7, 9)
    assert np.array_equal(o, GOLD[f"ds_out_{mth}_{int(al * 10)}"])
    assert np.array_equal(o, _ds_np(dv, mth, al))


# ------------------------------------------------------------------ phase correlation
def _pc(lib, prefix, fx_, mv, win):
    sh = np.zeros(3)
    pk = C.c_double(-1.0)
    rc = getattr(lib, prefix + "phase_correlate")(clibs.fp(fx_), clibs.fp(mv), *fx_.shape, clibs.dp(sh),
                                                  C.byref(pk), win)
    return rc, sh, pk.value


# Unwindowed, zero-padded patches are ill-conditioned: the unit-magnitude
# normalisation divides by cross-power magnitudes that are tiny where the padded
# patch has little energy, so fp32 rounding (the -Ofast reference's or this
# restatement's) moves the estimate by up to ~0.04 voxel around the fp64 value;
# windowed cases agree to ~1e-4.
PC_ATOL = {0: 5e-2, 1: 1e-4, 2: 5e-2, 3: 1e-4, 4: 1e-4, 5: 0.0}


@pytest.mark.parametrize("i", range(6))
def test_phase_correlate_golden(orc, i):
    want = GOLD[f"pc_out_{i}"]
    rc, sh, pk = _pc(orc, "orc_", GOLD[f"pc_fixed_{i}"], GOLD[f"pc_moving_{i}"], int(want[1]))
    assert rc == want[0]
    if rc == 0:
        assert np.allclose(sh, want[2:5], rtol=0, atol=PC_ATOL[i])
        assert abs(pk - want[5]) <= max(PC_ATOL[i], 1e-5) * want[5] * 10


def test_phase_correlate_fourier_shift(orc):
    """A band-limited sub-voxel shift of a periodic volume is recovered to the
    estimator's accuracy, in the right direction on every axis."""
    b = GOLD["pc_fixed_2"]
    rc, sh, pk = _pc(orc, "orc_", b, GOLD["pc_moving_2"], 0)
    assert rc == 0 and np.allclose(sh, (0.3, -1.6, 2.25), atol=0.05)


def test_phase_correlate_circular_shift_exact(orc):
    """A circular integer shift of a power-of-two volume makes the normalised
    cross-power spectrum an exact phase ramp: the surface is a delta of height
    pn at the shift (peak 1, no sub-voxel part)."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((16, 32, 8)).astype(np.float32)
    for s in [(3, -5, 2), (0, 0, 0), (-8, 16, 4)]:
        rc, sh, pk = _pc(orc, "orc_", a, np.roll(a, s, (0, 1, 2)).astype(np.float32), 0)
        want = [((si + n // 2) % n) - n // 2 if ((si % n) > n // 2) else si % n for si, n in zip(s, a.shape)]
        assert rc == 0 and np.allclose(sh, want, atol=1e-4) and abs(pk - 1.0) < 1e-4


# ---- pass-1 commits (pipeline.c:735-740,851-900), pinned by closed forms
def _pcfg(**kw):
    c = clibs.PipelineCfg()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _dp(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(clibs.c_double_p)


def test_commit_eps_fit_recovers_a_line(orc):
    r = np.linspace(10.0, 900.0, 37)
    f = 0.01 + 2.5e-5 * r                      # f(rmax)/f(0) = 3.25
    c = _pcfg(guided_eps=0.0, denoise_k=0.0)
    orc.orc_calib_commit_eps(C.byref(c), _dp(r), _dp(f), len(r))
    assert c.have_eps_r == 1
    assert abs(c.eps_fn_a - 0.01) < 1e-12 and abs(c.eps_fn_b - 2.5e-5) < 1e-15
    med = np.median(f)
    assert abs(c.eps_fn_med - med) < 1e-15
    assert abs(c.noise_floor - 255 * med) < 1e-12
    assert abs(c.guided_eps - (4.2 * med) ** 2) < 1e-15          # k = 4.2 default


@pytest.mark.parametrize("case", ["few", "neg_slope", "low_corr", "flat_span", "neg_intercept"])
def test_commit_eps_gates(orc, case):
    rng = np.random.default_rng(4)
    r = np.linspace(10.0, 900.0, 40)
    f = 0.01 + 2.5e-5 * r
    if case == "few":
        r, f = r[:7], f[:7]
    elif case == "neg_slope":
        f = 0.05 - 2.5e-5 * r
    elif case == "low_corr":
        f = 0.02 + 0.02 * rng.random(r.size) + 1e-6 * r
    elif case == "flat_span":
        f = 0.01 + 2e-6 * r                    # f1/f0 = 1.18 < 1.3
    elif case == "neg_intercept":
        f = -0.005 + 2.5e-5 * r
    c = _pcfg(guided_eps=2e-3, denoise_k=4.0, eps_fn_a=7.0, eps_fn_b=8.0, have_eps_r=1)
    orc.orc_calib_commit_eps(C.byref(c), _dp(r), _dp(f), len(r))
    assert c.have_eps_r == 0 and c.eps_fn_a == 7.0 and c.eps_fn_b == 8.0
    assert c.guided_eps == 2e-3                                   # preset eps kept


def test_commit_eps_medians_and_denoise_off(orc):
    f = np.array([0.3, 0.1, 0.2, 0.4])
    r = np.zeros(4)
    c = _pcfg(guided_eps=0.0, denoise_k=-1.0)
    orc.orc_calib_commit_eps(C.byref(c), _dp(r), _dp(f), 4)
    assert c.eps_fn_med == 0.25 and c.guided_eps == 0.0           # even n: mean of the middle two
    c = _pcfg(guided_eps=0.0, denoise_k=3.0)
    orc.orc_calib_commit_eps(C.byref(c), _dp(r[:3]), _dp(f[:3]), 3)
    assert c.eps_fn_med == 0.2 and abs(c.guided_eps - (3.0 * 0.2) ** 2) < 1e-15
    c = _pcfg(guided_eps=0.0, denoise_k=0.0)
    orc.orc_calib_commit_eps(C.byref(c), None, None, 0)
    assert c.eps_fn_med == 0.015


def test_aniso_ratio_and_commit(orc):
    # rho = exp(-1/(2 s^2)) for a Gaussian correlation of width s: ratio = s_z/s_xy
    sz, sxy = 1.4, 1.0
    rz, rxy = np.exp(-1 / (2 * sz ** 2)), np.exp(-1 / (2 * sxy ** 2))
    assert abs(orc.orc_aniso_ratio(rz, rxy) - sz / sxy) < 1e-12
    assert orc.orc_aniso_ratio(0.99, 0.5) == -1.0 and orc.orc_aniso_ratio(0.5, 0.04) == -1.0
    assert orc.orc_aniso_ratio(0.94, 0.06) == -1.0                # ratio > 3 is rejected
    c = _pcfg(psf_z_ratio=1.0)
    orc.orc_calib_commit_aniso(C.byref(c), _dp([1.2, 1.3, 1.25, 1.1, 1.4]), 5)
    assert c.psf_z_ratio == 1.25
    c = _pcfg(psf_z_ratio=1.0)
    orc.orc_calib_commit_aniso(C.byref(c), _dp([2.0] * 6), 6)
    assert c.psf_z_ratio == 1.6                                   # clamped
    for vals in ([1.2] * 4, [1.0, 1.01, 1.04, 1.02, 1.03]):       # too few / below 1.05
        c = _pcfg(psf_z_ratio=1.0)
        orc.orc_calib_commit_aniso(C.byref(c), _dp(vals), len(vals))
        assert c.psf_z_ratio == 1.0
