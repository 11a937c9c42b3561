"""Tile I/O feeding through the GPU: fy_process_chunk (pipeline.c:518-538) and
the pass-2 export loop fy_gpu_export (pipeline.c:1000-1070) from a zarr input
to a zarr output.

Bars: fy_process_chunk equals fy_process_buffer on the same region bit for bit
and the oracle's process_buffer within the chain's u8 bar (+-1 on < 0.1 % of
voxels); every chunk the export writes equals fy_process_chunk's tile bit for
bit, whatever the batch size, thread count or tile range; all-air tiles leave
no chunk (read back as 0), as in the reference; a failed read is counted and
fails the call without stopping the other tiles.
"""
import ctypes as C
import os

import numpy as np
import pytest

import clibs
import fysics_amd as F
from fysics_amd import phantom
from test_zarr_host import write_volume

pytestmark = pytest.mark.gpu

SHAPE, IN_CHUNK, TILE, HALO = (72, 80, 88), (24, 40, 32), 32, 8


def chain_cfg(mod, **kw):
    base = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3, unsharp_sigma=0.8,
                unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2, do_deconv=1, guided_eps=1.2e-3,
                guided_subsample=0, halo=HALO)
    base.update(kw)
    c = mod.PipelineCfg()
    for k, v in base.items():
        setattr(c, k, v)
    return c


@pytest.fixture(scope="module")
def volume(tmp_path_factory):
    rng = np.random.default_rng(11)
    vol = phantom.noisy_u8(phantom.forward(phantom.sheets(SHAPE, rng, n_sheets=8), **phantom.CONFIG1_PHYS), rng)
    vol[:, :, 56:] = 0                     # tiles with x0 = 64 (halo from 56) are all air
    root = str(tmp_path_factory.mktemp("exp") / "in.zarr")
    write_volume(root, vol, IN_CHUNK)
    z = F.Zarr()
    assert F.lib().fy_gpu_zarr_open(C.byref(z), root.encode()) == 0
    return vol, z, root


def tiles():
    n = [(s + TILE - 1) // TILE for s in SHAPE]
    for t in range(n[0] * n[1] * n[2]):
        iz, r = divmod(t, n[1] * n[2])
        iy, ix = divmod(r, n[2])
        yield t, (iz, iy, ix)


def process_chunk(z, cfg, org):
    out = np.empty(TILE ** 3, np.uint8)
    tz, ty, tx, air = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = F.lib().fy_gpu_process_chunk(C.byref(z), C.byref(cfg), *org, TILE, out.ctypes.data, C.byref(tz), C.byref(ty),
                                  C.byref(tx), C.byref(air))
    assert rc == 0, F.lib().fy_gpu_last_error()
    shp = (tz.value, ty.value, tx.value)
    return out[:np.prod(shp)].reshape(shp), bool(air.value)


def region(org, shp):
    lo = [max(0, o - HALO) for o in org]
    hi = [min(S, o + s + HALO) for S, o, s in zip(SHAPE, org, shp)]
    return lo, hi


def test_process_chunk_vs_buffer_and_oracle(volume, orc):
    vol, z, _ = volume
    cfg, ocfg = chain_cfg(F), chain_cfg(clibs)
    for org in [(0, 0, 0), (32, 40, 32), (64, 64, 32)]:
        print(f"process_chunk tile at {org}", flush=True)
        got, air = process_chunk(z, cfg, org)
        assert not air
        lo, hi = region(org, got.shape)
        u8 = np.ascontiguousarray(vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]])
        buf, bair = F.process_buffer(cfg, u8, tuple(lo), org, got.shape, SHAPE[1:])
        assert np.array_equal(got, buf) and not bair
        want = np.empty(got.shape, np.uint8)
        oair = C.c_int(0)
        assert orc.orc_process_buffer(C.byref(ocfg), clibs.up(u8), *lo, *u8.shape, *org, *got.shape, *SHAPE[1:],
                                      clibs.up(want), C.byref(oair)) == 0
        diff = np.abs(got.astype(int) - want.astype(int))
        assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


def chunk_bytes(root, idx):
    p = os.path.join(root, "0", *map(str, idx))
    return open(p, "rb").read() if os.path.exists(p) else None


def run_export(z, root_out, cfg, t0=0, t1=0, threads=4, batch=3):
    lib = F.lib()
    zo = F.Zarr()
    assert lib.fy_gpu_zarr_create(C.byref(zo), root_out.encode(), (C.c_long * 3)(*SHAPE), (C.c_long * 3)(TILE, TILE, TILE)) == 0
    st = F.ExportStats()
    rc = lib.fy_gpu_export(C.byref(z), C.byref(zo), C.byref(cfg), TILE, t0, t1, threads, batch, C.byref(st), None)
    return rc, st


@pytest.mark.parametrize("threads,batch", [(4, 3), (1, 8), (8, 1)])
def test_export_matches_process_chunk(volume, tmp_path, threads, batch):
    vol, z, _ = volume
    cfg = chain_cfg(F)
    out = str(tmp_path / "out.zarr")
    print(f"export threads={threads} batch={batch}", flush=True)
    rc, st = run_export(z, out, cfg, threads=threads, batch=batch)
    assert rc == 0, F.lib().fy_gpu_last_error()
    n = len(list(tiles()))
    assert st.tiles == n and st.io_fail == 0 and st.written + st.air == n and st.air == 9
    for t, idx in tiles():
        org = tuple(i * TILE for i in idx)
        want, air = process_chunk(z, cfg, org)
        got = chunk_bytes(out, idx)
        if air:
            assert got is None, idx
        else:
            assert got == want.tobytes(), idx


def test_export_tile_range(volume, tmp_path):
    vol, z, _ = volume
    cfg = chain_cfg(F, no_dither=1)
    out = str(tmp_path / "part.zarr")
    rc, st = run_export(z, out, cfg, t0=2, t1=7, threads=3, batch=2)
    assert rc == 0 and st.tiles == 5
    for t, idx in tiles():
        got = chunk_bytes(out, idx)
        if not 2 <= t < 7:
            assert got is None
            continue
        want, air = process_chunk(z, cfg, tuple(i * TILE for i in idx))
        assert (got is None) if air else (got == want.tobytes())


@pytest.mark.skipif(os.geteuid() == 0, reason="file permissions do not bind root")
def test_export_counts_failed_reads(volume, tmp_path):
    vol, z, root = volume
    bad = os.path.join(root, "0", "1", "1", "1")
    os.chmod(bad, 0)
    try:
        out = str(tmp_path / "fail.zarr")
        rc, st = run_export(z, out, chain_cfg(F))
        assert rc == 1 and st.io_fail >= 1
        assert st.written + st.air + st.io_fail >= st.tiles      # the other tiles still ran
        assert b"zarr" in F.lib().fy_gpu_last_error()
    finally:
        os.chmod(bad, 0o644)
