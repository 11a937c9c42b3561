"""Pass-1 calibration statistics on the GPU (stream.c:65-200, zdrift.c,
pipeline.c:576-701) against the oracle.

Bars: histograms, counts and the u8 z-drift sums are bit-exact (integers, and
sums of multiples of 2^-27); flat-block noise of u8-derived data is bit-exact
(its double sums are exact in any order); float z-drift sums within 1e-12
relative (summation order differs), their counts exact; normalisation apply
bit-exact.  Full-size (256^3) checks use independent implementations
(torch.bincount on the device, numpy) as the reference.
"""
import ctypes as C
import os

import numpy as np
import pytest
import torch

import clibs
import fysics_amd as F

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def orc_hist(orc, u8):
    hs = clibs.HistState()
    u8 = np.ascontiguousarray(u8)
    orc.orc_hist_accumulate_u8(C.byref(hs), clibs.up(u8), u8.size)
    return hs


def fy_hist(fy, u8, hs=None):
    hs = hs if hs is not None else F.HistState()
    fy.fy_hist_accumulate_u8(C.byref(hs), u8.ctypes.data_as(F.c_u8_p), u8.size)
    return hs


@pytest.mark.parametrize("n", [0, 1, 15, 17, 33, 1000, (1 << 20) + 7])
@pytest.mark.parametrize("offset", [0, 1, 7])
def test_hist_vs_oracle(fy, orc, n, offset):
    print(f"hist n={n} offset={offset}", flush=True)
    rng = np.random.default_rng(n + offset)
    buf = np.clip(rng.normal(100, 60, n + offset), 0, 255).astype(np.uint8)
    u8 = buf[offset:]                      # host pointer at any alignment
    got, want = fy_hist(fy, u8), orc_hist(orc, u8)
    assert list(got.hist) == list(want.hist)
    assert got.total == want.total == n
    for q in (0.0, 0.5, 10.0, 50.0, 99.5, 100.0):
        assert fy.fy_hist_percentile_u8(C.byref(got), q) == orc.orc_hist_percentile_u8(C.byref(want), q)
        assert fy.fy_hist_percentile_u8_inner(C.byref(got), q) == \
            orc.orc_hist_percentile_u8_inner(C.byref(want), q)


def test_hist_golden_and_merge(fy):
    u8 = GOLD["hist_u8"]
    a = fy_hist(fy, np.ascontiguousarray(u8[:3001]))
    b = fy_hist(fy, np.ascontiguousarray(u8[3001:]))
    fy.fy_hist_merge(C.byref(a), C.byref(b))
    assert np.array_equal(np.array(a.hist[:]), GOLD["hist_counts"])
    assert a.total == GOLD["hist_total"][0]
    got = [fy.fy_hist_percentile_u8(C.byref(a), q) for q in GOLD["hist_pcts"]]
    assert got == list(GOLD["hist_pct"])
    lo, hi = C.c_ubyte(), C.c_ubyte()
    fy.fy_norm_finalize(C.byref(a), 0.5, 99.5, C.byref(lo), C.byref(hi))
    assert (lo.value, hi.value) == (got[1], got[6])


@pytest.mark.parametrize("value", [0, 255, 77])
def test_hist_single_value(fy, value):
    """Every byte in one bin: the sub-histograms all hit the same bin."""
    u8 = np.full(3 * (1 << 20) + 5, value, np.uint8)
    hs = fy_hist(fy, u8)
    assert hs.hist[value] == u8.size and hs.total == u8.size


def test_norm_apply_vs_golden(fy, orc):
    nin = GOLD["norm_in"]
    for k, (lo, hi) in enumerate(((10, 200), (0, 255), (90, 90), (200, 10))):
        o = np.empty(nin.size, np.float32)
        fy.fy_norm_apply_u8(nin.ctypes.data_as(F.c_u8_p), o.ctypes.data_as(F.c_float_p), nin.size, lo, hi)
        assert np.array_equal(o, GOLD["norm_out"][k])


def test_zdrift_accumulate_vs_oracle(fy, orc):
    zch = GOLD["zdrift_chunk"]
    zsum = np.linspace(0.0, 5.0, 12)
    zcnt = np.arange(12, dtype=np.int64) * 3
    fy.fy_zdrift_accumulate(zch.ctypes.data_as(F.c_float_p), 8, 20, 24, 3, zsum.ctypes.data_as(F.c_double_p),
                            zcnt.ctypes.data_as(C.POINTER(C.c_long)), 0.3)
    assert np.array_equal(zcnt, GOLD["zdrift_counts"])
    assert np.allclose(zsum, GOLD["zdrift_sums"], rtol=1e-12, atol=0)
    rng = np.random.default_rng(3)
    v = rng.random((37, 45, 61)).astype(np.float32)
    for thresh in (0.0, 0.1, 0.75, 2.0):
        s1, c1 = np.zeros(40), np.zeros(40, np.int64)
        s2, c2 = np.zeros(40), np.zeros(40, np.int64)
        fy.fy_zdrift_accumulate(v.ctypes.data_as(F.c_float_p), 37, 45, 61, 2, s1.ctypes.data_as(F.c_double_p),
                                c1.ctypes.data_as(C.POINTER(C.c_long)), thresh)
        orc.orc_zdrift_accumulate(clibs.fp(v), 37, 45, 61, 2, clibs.dp(s2), clibs.lp(c2), thresh)
        assert np.array_equal(c1, c2)
        assert np.allclose(s1, s2, rtol=1e-12, atol=0)


def test_zdrift_finalize_and_correct(fy, orc):
    fac = np.empty(12, np.float32)
    counts = np.ascontiguousarray(GOLD["zdrift_counts"])
    fy.fy_zdrift_finalize(GOLD["zdrift_sums"].ctypes.data_as(F.c_double_p),
                          counts.ctypes.data_as(C.POINTER(C.c_long)), 12, fac.ctypes.data_as(F.c_float_p))
    assert np.array_equal(fac, GOLD["zdrift_factor"])
    rng = np.random.default_rng(4)
    nz, ny, nx = 40, 33, 29
    vol = (rng.random((nz, ny, nx)) * np.linspace(1.0, 0.7, nz)[:, None, None]).astype(np.float32)
    got = vol.copy()
    assert fy.fy_correct_zdrift(got.ctypes.data_as(F.c_float_p), nz, ny, nx, 0.2) == 0
    s, c = np.zeros(nz), np.zeros(nz, np.int64)
    orc.orc_zdrift_accumulate(clibs.fp(vol), nz, ny, nx, 0, clibs.dp(s), clibs.lp(c), 0.2)
    f = np.empty(nz, np.float32)
    orc.orc_zdrift_finalize(clibs.dp(s), clibs.lp(c), nz, clibs.fp(f))
    want = vol * f[:, None, None]
    assert np.allclose(got, want, rtol=2e-7, atol=0)
    # the correction itself is one fp32 multiply per voxel: with the factors
    # finalized from the device's own sums the result is bit-exact
    sg, cg = np.zeros(nz), np.zeros(nz, np.int64)
    fy.fy_zdrift_accumulate(vol.ctypes.data_as(F.c_float_p), nz, ny, nx, 0, sg.ctypes.data_as(F.c_double_p),
                            cg.ctypes.data_as(C.POINTER(C.c_long)), 0.2)
    assert np.array_equal(cg, c) and np.allclose(sg, s, rtol=1e-12, atol=0)
    fg = np.empty(nz, np.float32)
    orc.orc_zdrift_finalize(clibs.dp(sg), clibs.lp(cg), nz, clibs.fp(fg))
    assert np.array_equal(got, vol * fg[:, None, None])


@pytest.mark.parametrize("blk", [8, 5, 3, 16, 41])
def test_flat_noise_u8_data_exact(fy, orc, blk):
    fv = GOLD["flat_u8vol"]
    got = fy.fy_flat_noise(fv.ctypes.data_as(F.c_float_p), 40, 40, 48, blk)
    assert got == orc.orc_flat_noise(clibs.fp(fv), 40, 40, 48, blk)
    assert abs(got - GOLD["flat_noise"][[8, 5, 3, 16, 41].index(blk)]) <= 1e-12 * max(got, 1e-30)


def test_flat_noise_float_data_and_device_api(fy, orc):
    fg = GOLD["flat_gvol"]
    for k, blk in enumerate((8, 4)):
        got = fy.fy_flat_noise(fg.ctypes.data_as(F.c_float_p), 33, 34, 35, blk)
        assert abs(got - orc.orc_flat_noise(clibs.fp(fg), 33, 34, 35, blk)) <= 1e-12 * got
        assert abs(got - GOLD["flat_noise"][5 + k]) <= 1e-12 * got
        d = torch.from_numpy(fg).cuda()
        out = C.c_double(-5.0)
        assert fy.fy_gpu_flat_noise(C.c_void_p(d.data_ptr()), 33, 34, 35, blk, C.byref(out), None) == 0
        assert out.value == got
    # no bright block -> 0; volume smaller than a block -> 0
    z = np.zeros((16, 16, 16), np.float32)
    assert fy.fy_flat_noise(z.ctypes.data_as(F.c_float_p), 16, 16, 16, 8) == 0.0
    assert fy.fy_flat_noise(z.ctypes.data_as(F.c_float_p), 7, 16, 16, 8) == 0.0


def sweep_oracle(orc, vol, tiles):
    Z = vol.shape[0]
    hs = clibs.HistState()
    zs, zc = np.zeros(Z), np.zeros(Z, np.int64)
    for (z0, y0, x0, dz, dy, dx) in tiles:
        t = np.ascontiguousarray(vol[z0:z0 + dz, y0:y0 + dy, x0:x0 + dx])
        if not t.any():
            continue                                   # pipeline.c:677
        orc.orc_hist_accumulate_u8(C.byref(hs), clibs.up(t), t.size)
        orc.orc_zdrift_u8_sub(clibs.up(t), dz, dy, dx, z0, clibs.dp(zs), clibs.lp(zc))
    return hs, zs, zc


def test_calib_sweep_vs_oracle(fy, orc):
    """Pass-1 sweep over ptiles of a volume, edge tiles ragged, one tile all zero."""
    rng = np.random.default_rng(9)
    Z, Y, X, T = 150, 131, 170, 64
    vol = np.clip(rng.normal(90, 70, (Z, Y, X)), 0, 255).astype(np.uint8)
    vol[64:128, 64:128, 0:64] = 0
    tiles = [(z, y, x, min(T, Z - z), min(T, Y - y), min(T, X - x))
             for z in range(0, Z, T) for y in range(0, Y, T) for x in range(0, X, T)]
    stream = torch.cuda.Stream()
    cal = fy.fy_calib_create(Z, C.c_void_p(stream.cuda_stream))
    assert cal
    keep = []
    for (z0, y0, x0, dz, dy, dx) in tiles:
        t = torch.from_numpy(np.ascontiguousarray(vol[z0:z0 + dz, y0:y0 + dy, x0:x0 + dx])).cuda()
        keep.append(t)
        torch.cuda.synchronize()
        assert fy.fy_calib_accumulate_u8(cal, C.c_void_p(t.data_ptr()), z0, dz, dy, dx, 3) == 0
    hs = F.HistState()
    zs, zc = np.zeros(Z), np.zeros(Z, np.int64)
    assert fy.fy_calib_read(cal, C.byref(hs), zs.ctypes.data_as(F.c_double_p),
                            zc.ctypes.data_as(C.POINTER(C.c_long))) == 0
    whs, wzs, wzc = sweep_oracle(orc, vol, tiles)
    assert list(hs.hist) == list(whs.hist)
    assert hs.total == whs.total == vol.size - 64 ** 3
    assert np.array_equal(zc, wzc)
    assert np.array_equal(zs, wzs)                       # exact sums
    # out-of-range slices are refused; reset clears
    t = keep[0]
    assert fy.fy_calib_accumulate_u8(cal, C.c_void_p(t.data_ptr()), Z - 10, 64, 64, 64, 2) == 1
    assert fy.fy_calib_reset(cal) == 0
    hs2 = F.HistState()
    assert fy.fy_calib_read(cal, C.byref(hs2), None, None) == 0
    assert hs2.total == 0 and sum(hs2.hist) == 0
    fy.fy_calib_destroy(cal)


def test_calib_full_size_256(fy):
    """The 256^3 u8 cube of the chain workload: histogram against torch.bincount on
    the device, z-drift sums against numpy (both exact)."""
    from fysics_amd import phantom
    n = 256
    u8 = phantom.config2_volume((n, n, n), seed=2)
    d = torch.from_numpy(u8).cuda()
    cal = fy.fy_calib_create(n, None)
    assert fy.fy_calib_accumulate_u8(cal, C.c_void_p(d.data_ptr()), 0, n, n, n, 3) == 0
    hs = F.HistState()
    zs, zc = np.zeros(n), np.zeros(n, np.int64)
    assert fy.fy_calib_read(cal, C.byref(hs), zs.ctypes.data_as(F.c_double_p),
                            zc.ctypes.data_as(C.POINTER(C.c_long))) == 0
    fy.fy_calib_destroy(cal)
    want = torch.bincount(d.flatten().long(), minlength=256).cpu().numpy()
    assert np.array_equal(np.array(hs.hist[:]), want)
    v = u8[:, ::4, ::4].astype(np.float32) * np.float32(1.0 / 255.0)
    m = v > np.float32(0.10)
    assert np.array_equal(zc, m.sum(axis=(1, 2)))
    assert np.array_equal(zs, (v.astype(np.float64) * m).sum(axis=(1, 2)))
