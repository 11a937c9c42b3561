"""The §8(f) rank-4 consumers of the same kernels on the GPU: the 2x LOD
downscale (downscale.c) and phase correlation (phasecorr.c), against the oracle.

Bars: the downscale is byte work and bit-exact (vectorised and edge paths,
both methods, volumes whose output rows need more than one launch).  Phase
correlation: a circular integer shift of a power-of-two volume is recovered
exactly with peak 1 (the surface is a delta); otherwise the shift agrees with
the oracle's to 1e-4 voxel and the peak to 1e-3 for windowed patches.  For
unwindowed zero-padded patches the unit-magnitude normalisation divides by
cross-power magnitudes that are tiny where the padded patch has little energy,
and fp32 results -- the reference's included -- scatter by up to 0.04 voxel
around the fp64 value (e.g. 1.790 / 1.787 / 1.777 for oracle / reference / GPU
against 1.764 in fp64): there the bar is 0.05 voxel and 5% of the peak.  A
flat input returns 1.
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


def ds_oracle(orc, v, method, alpha):
    nz, ny, nx = v.shape
    o = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), np.uint8)
    orc.orc_downscale2x(clibs.up(np.ascontiguousarray(v)), nx, ny, nz, clibs.up(o), None, None, None, method,
                        C.c_float(alpha))
    return o


def ds_host(fy, v, method, alpha):
    nz, ny, nx = v.shape
    o = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), np.uint8)
    ox, oy, oz = C.c_int(), C.c_int(), C.c_int()
    fy.fy_downscale2x(np.ascontiguousarray(v).ctypes.data_as(F.c_u8_p), nx, ny, nz, o.ctypes.data_as(F.c_u8_p),
                      C.byref(ox), C.byref(oy), C.byref(oz), method, C.c_float(alpha))
    assert (oz.value, oy.value, ox.value) == o.shape
    return o


@pytest.mark.parametrize("shape", [(9, 14, 17), (8, 48, 64), (7, 33, 24), (1, 5, 40), (16, 16, 1), (2, 64, 256)])
@pytest.mark.parametrize("method,alpha", [(0, 0.0), (1, 0.5), (1, 0.8)])
def test_downscale_vs_oracle(fy, orc, shape, method, alpha):
    print(f"downscale shape={shape} method={method}", flush=True)
    rng = np.random.default_rng(sum(shape))
    v = np.where(rng.random(shape) < 0.3, 0, rng.integers(0, 256, shape)).astype(np.uint8)
    v[: max(1, shape[0] // 2), : max(1, shape[1] // 2), : max(1, shape[2] // 2)] = 0    # all-zero cells
    assert np.array_equal(ds_host(fy, v, method, alpha), ds_oracle(orc, v, method, alpha))


def test_downscale_golden(fy):
    dv = GOLD["ds_in"]
    for mth, al in ((0, 0.0), (1, 0.5), (1, 0.8)):
        assert np.array_equal(ds_host(fy, dv, mth, al), GOLD[f"ds_out_{mth}_{int(al * 10)}"])


def test_downscale_device_256_and_slabs(fy, orc):
    """The 256^3 workload cube through the device entry point, and a volume whose
    oz*oy output rows exceed one launch's grid."""
    from fysics_amd import phantom
    for v in (phantom.config2_volume((256, 256, 256), seed=2),
              np.random.default_rng(5).integers(0, 256, (600, 500, 16)).astype(np.uint8)):
        nz, ny, nx = v.shape
        d = torch.from_numpy(v).cuda()
        o = torch.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        assert fy.fy_gpu_downscale2x(C.c_void_p(d.data_ptr()), nx, ny, nz, C.c_void_p(o.data_ptr()), 1,
                                     C.c_float(0.5), None) == 0
        assert fy.fy_gpu_sync(None) == 0
        assert np.array_equal(o.cpu().numpy(), ds_oracle(orc, v, 1, 0.5))


@pytest.mark.parametrize("shape,slab", [((37, 40, 52), 4096), ((30, 24, 24), 1), ((9, 14, 17), 1000)])
def test_downscale_host_slabs(fy, orc, shape, slab, monkeypatch):
    """fy_downscale2x works through slabs of output planes (FYSICS_DS_SLAB_BYTES
    of input per slab, at least two planes): small slabs -- odd plane counts,
    a last slab with one parent plane -- give the oracle's bytes exactly."""
    monkeypatch.setenv("FYSICS_DS_SLAB_BYTES", str(slab))
    print(f"downscale slabs shape={shape} slab={slab}", flush=True)
    rng = np.random.default_rng(sum(shape) + slab)
    v = np.where(rng.random(shape) < 0.3, 0, rng.integers(0, 256, shape)).astype(np.uint8)
    for method, alpha in ((0, 0.0), (1, 0.5)):
        assert np.array_equal(ds_host(fy, v, method, alpha), ds_oracle(orc, v, method, alpha))


PC_ATOL = {0: 5e-2, 1: 1e-4, 2: 5e-2, 3: 1e-4, 4: 1e-4, 5: 0.0}


def pc_host(fy, a, b, win):
    sh = np.zeros(3)
    pk = C.c_double(-1.0)
    rc = fy.fy_phase_correlate(a.ctypes.data_as(F.c_float_p), b.ctypes.data_as(F.c_float_p), *a.shape,
                               sh.ctypes.data_as(F.c_double_p), C.byref(pk), win)
    return rc, sh, pk.value


@pytest.mark.parametrize("i", range(6))
def test_phase_correlate_vs_oracle_and_golden(fy, orc, i):
    a, b = GOLD[f"pc_fixed_{i}"], GOLD[f"pc_moving_{i}"]
    win = int(GOLD[f"pc_out_{i}"][1])
    rc, sh, pk = pc_host(fy, a, b, win)
    osh = np.zeros(3)
    opk = C.c_double(-1.0)
    orc_rc = orc.orc_phase_correlate(clibs.fp(a), clibs.fp(b), *a.shape, clibs.dp(osh), C.byref(opk), win)
    assert rc == orc_rc == GOLD[f"pc_out_{i}"][0]
    if rc == 0:
        assert np.allclose(sh, osh, atol=PC_ATOL[i])
        assert np.allclose(sh, GOLD[f"pc_out_{i}"][2:5], atol=PC_ATOL[i])
        assert abs(pk - opk.value) <= (5e-2 if PC_ATOL[i] > 1e-3 else 1e-3) * opk.value


def test_phase_correlate_circular_and_device(fy):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((16, 32, 8)).astype(np.float32)
    for s in [(3, -5, 2), (0, 0, 0), (-7, 16, 4)]:
        b = np.roll(a, s, (0, 1, 2)).astype(np.float32)
        want = [(si + n // 2) % n - n // 2 if (si % n) > n // 2 else si % n for si, n in zip(s, a.shape)]
        rc, sh, pk = pc_host(fy, a, b, 0)
        assert rc == 0 and np.allclose(sh, want, atol=1e-4) and abs(pk - 1.0) < 1e-3
        da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        torch.cuda.synchronize()
        sh2 = np.zeros(3)
        pk2 = C.c_double()
        assert fy.fy_gpu_phase_correlate(C.c_void_p(da.data_ptr()), C.c_void_p(db.data_ptr()), *a.shape,
                                         sh2.ctypes.data_as(F.c_double_p), C.byref(pk2), 0, None) == 0
        assert np.array_equal(sh2, sh) and pk2.value == pk
    z = np.zeros((4, 4, 4), np.float32)
    assert pc_host(fy, z, a[:4, :4, :4].copy(), 1)[0] == 1                 # flat fixed
    assert pc_host(fy, a[:1].copy(), a[:1].copy(), 1)[0] == 1              # an axis below 2
