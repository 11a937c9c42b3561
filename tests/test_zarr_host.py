"""Zarr v2 raw-u8 I/O on the host (zarr_io.c): the format the pass-2 tile
loop reads and writes.

A volume is written chunk by chunk with fy_gpu_zarr_create /
fy_gpu_zarr_write_chunk (the reference's fy_zarr_create / fy_zarr_write_chunk
for local roots; far-face chunks at their true extent, zarr_io.c:280-304), some
chunks are then deleted (sparse: read as fill) or truncated (the tail reads as
fill), and regions read back with fy_gpu_zarr_read must equal the same assembly done with
numpy slicing -- and, where the reference is built (oracle/_ref), the
reference's own fy_zarr_read on the same files, byte for byte.  Errors: a
missing header, an S3 root, compressed chunks.
"""
import ctypes as C
import os

import numpy as np
import pytest

import clibs
import fysics_amd as F

SHAPE, CHUNK = (37, 50, 29), (16, 20, 12)


def write_volume(root, vol, chunk, drop=(), trunc=()):
    lib = F.lib()
    z = F.Zarr()
    shape = (C.c_long * 3)(*vol.shape)
    ch = (C.c_long * 3)(*chunk)
    assert lib.fy_gpu_zarr_create(C.byref(z), root.encode(), shape, ch) == 0
    nc = [(n + c - 1) // c for n, c in zip(vol.shape, chunk)]
    for cz in range(nc[0]):
        for cy in range(nc[1]):
            for cx in range(nc[2]):
                blk = np.ascontiguousarray(vol[cz * chunk[0]:(cz + 1) * chunk[0], cy * chunk[1]:(cy + 1) * chunk[1],
                                               cx * chunk[2]:(cx + 1) * chunk[2]])
                assert lib.fy_gpu_zarr_write_chunk(C.byref(z), cz, cy, cx, blk.ctypes.data, *blk.shape) == 0
    expect = vol.copy()
    for (cz, cy, cx) in drop:
        os.unlink(os.path.join(root, "0", str(cz), str(cy), str(cx)))
        expect[cz * chunk[0]:(cz + 1) * chunk[0], cy * chunk[1]:(cy + 1) * chunk[1],
               cx * chunk[2]:(cx + 1) * chunk[2]] = 0
    for (cz, cy, cx), keep in trunc:
        path = os.path.join(root, "0", str(cz), str(cy), str(cx))
        with open(path, "r+b") as f:
            f.truncate(keep)
        sl = (slice(cz * chunk[0], (cz + 1) * chunk[0]), slice(cy * chunk[1], (cy + 1) * chunk[1]),
              slice(cx * chunk[2], (cx + 1) * chunk[2]))
        blk = expect[sl].copy()
        flat = blk.reshape(-1)
        flat[keep:] = 0
        expect[sl] = flat.reshape(blk.shape)
    return z, expect


def read(lib, z, org, shp, zcls=None):
    out = np.full(shp, 0xAB, np.uint8)
    assert lib.fy_gpu_zarr_read(C.byref(z), *org, *shp, out.ctypes.data) == 0
    return out


def regions(rng, shape, n):
    yield (0, 0, 0), shape
    yield (5, 19, 11), (1, 1, 1)
    yield (15, 19, 11), (3, 3, 3)                        # a chunk corner
    for _ in range(n):
        o = [int(rng.integers(0, s)) for s in shape]
        d = [int(rng.integers(1, s - oo + 1)) for s, oo in zip(shape, o)]
        yield tuple(o), tuple(d)


def test_round_trip_sparse_and_truncated(tmp_path):
    rng = np.random.default_rng(5)
    vol = rng.integers(1, 256, SHAPE, dtype=np.uint8)
    root = str(tmp_path / "vol.zarr")
    z, expect = write_volume(root, vol, CHUNK, drop=[(1, 0, 1), (2, 2, 2)], trunc=[((0, 1, 0), 1000),
                                                                                  ((2, 0, 0), 7)])
    lib = F.lib()
    zz = F.Zarr()
    assert lib.fy_gpu_zarr_open(C.byref(zz), root.encode()) == 0
    assert tuple(zz.shape) == SHAPE and tuple(zz.chunk) == CHUNK and zz.fill == 0
    for org, shp in regions(rng, SHAPE, 40):
        got = read(lib, zz, org, shp)
        sl = tuple(slice(o, o + d) for o, d in zip(org, shp))
        assert np.array_equal(got, expect[sl]), (org, shp)
    # the far-face chunk holds its true extent only
    assert os.path.getsize(os.path.join(root, "0", "2", "0", "2")) == (37 - 32) * 20 * (29 - 24)


def test_matches_reference_reader(tmp_path, ref):
    """The reference's fy_zarr_read (zarr_io.c:197-330) on the same files."""
    rng = np.random.default_rng(6)
    vol = rng.integers(0, 256, SHAPE, dtype=np.uint8)
    root = str(tmp_path / "ref.zarr")
    write_volume(root, vol, CHUNK, drop=[(0, 2, 1)], trunc=[((1, 1, 1), 333)])
    lib = F.lib()
    zo, zr = F.Zarr(), clibs.Zarr()
    assert lib.fy_gpu_zarr_open(C.byref(zo), root.encode()) == 0
    assert ref.fy_zarr_open(C.byref(zr), root.encode()) == 0
    assert tuple(zr.shape) == tuple(zo.shape) and tuple(zr.chunk) == tuple(zo.chunk) and zr.fill == zo.fill
    for org, shp in regions(rng, SHAPE, 60):
        a = read(lib, zo, org, shp)
        b = np.full(shp, 0xAB, np.uint8)
        assert ref.fy_zarr_read(C.byref(zr), *org, *shp, b.ctypes.data) == 0
        assert np.array_equal(a, b), (org, shp)


def test_fill_value_and_outside_volume(tmp_path):
    """A nonzero fill_value fills missing chunks; voxels past the volume read as fill."""
    root = tmp_path / "f.zarr"
    (root / "0" / "0" / "0").mkdir(parents=True)
    (root / "0" / ".zarray").write_text(
        '{"chunks": [4, 4, 4], "compressor": null, "dtype": "|u1", "fill_value": 7, '
        '"order": "C", "shape": [6, 5, 4], "zarr_format": 2}')
    blk = np.arange(64, dtype=np.uint8).reshape(4, 4, 4)
    (root / "0" / "0" / "0" / "0").write_bytes(blk.tobytes())
    lib = F.lib()
    z = F.Zarr()
    assert lib.fy_gpu_zarr_open(C.byref(z), str(root).encode()) == 0
    assert z.fill == 7
    got = read(lib, z, (0, 0, 0), (8, 6, 4))
    want = np.full((8, 6, 4), 7, np.uint8)
    want[:4, :4, :4] = blk
    assert np.array_equal(got, want)


@pytest.mark.parametrize("header,why", [
    (None, "missing header"),
    ('{"chunks":[4,4,4],"compressor":{"id":"blosc"},"dtype":"|u1","fill_value":0,"shape":[8,8,8]}', "blosc"),
    ('{"chunks":[4,4,4],"compressor":null,"dtype":"<u2","fill_value":0,"shape":[8,8,8]}', "u2"),
    ('{"chunks":[4,4,4],"compressor":null,"dtype":"|u1","fill_value":0,"shape":[8,0,8]}', "empty"),
])
def test_open_rejects(tmp_path, header, why):
    root = tmp_path / "bad.zarr"
    (root / "0").mkdir(parents=True)
    if header is not None:
        (root / "0" / ".zarray").write_text(header)
    z = F.Zarr()
    assert F.lib().fy_gpu_zarr_open(C.byref(z), str(root).encode()) == 1, why
    assert F.lib().fy_gpu_zarr_open(C.byref(z), b"s3://bucket/vol.zarr") == 1


def test_header_is_readable_json(tmp_path):
    import json
    z = F.Zarr()
    root = str(tmp_path / "h.zarr")
    assert F.lib().fy_gpu_zarr_create(C.byref(z), root.encode(), (C.c_long * 3)(3, 4, 5), (C.c_long * 3)(2, 2, 2)) == 0
    h = json.load(open(os.path.join(root, "0", ".zarray")))
    assert h["shape"] == [3, 4, 5] and h["chunks"] == [2, 2, 2] and h["dtype"] == "|u1"
    assert h["compressor"] is None and h["fill_value"] == 0 and h["dimension_separator"] == "/"


def test_shared_chunk_mappings_under_threads(tmp_path):
    """The exporter's reads go through one set of chunk mappings shared by its
    I/O threads (mapped once, evicted least recently used past a byte budget).
    Eight threads read overlapping regions with a budget of a few chunks, so
    entries are evicted and remapped while others hold theirs; every region
    must equal the uncached read."""
    import threading
    rng = np.random.default_rng(9)
    vol = rng.integers(0, 256, SHAPE, dtype=np.uint8)
    root = str(tmp_path / "c.zarr")
    write_volume(root, vol, CHUNK, drop=[(1, 1, 1)], trunc=[((0, 0, 2), 100)])
    lib = F.lib()
    z = F.Zarr()
    assert lib.fy_gpu_zarr_open(C.byref(z), root.encode()) == 0
    cache = lib.fy_gpu_zarr_cache_new(3 * 16 * 20 * 12)
    regs = list(regions(rng, SHAPE, 200))
    errors = []

    def work(part):
        for org, shp in regs[part::8]:
            a = np.empty(shp, np.uint8)
            b = np.empty(shp, np.uint8)
            if lib.fy_gpu_zarr_read_cached(C.byref(z), cache, *org, *shp, a.ctypes.data) != 0 or \
                    lib.fy_gpu_zarr_read(C.byref(z), *org, *shp, b.ctypes.data) != 0 or not np.array_equal(a, b):
                errors.append((org, shp))

    ts = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    lib.fy_gpu_zarr_cache_free(cache)
    assert not errors, errors[:3]


def test_zarr_cache_threads_under_tsan(tmp_path):
    """The shared chunk mappings (fy_gpu_zarr_cache) under ThreadSanitizer:
    fy_zarr.c and the driver tests/tsan/zarr_cache_threads.c built with
    -fsanitize=thread; eight threads read overlapping regions through a cache
    that evicts and remaps chunks while others hold them.  Bars: no race
    reported, every cached region equals the uncached read."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    exe = str(tmp_path / "tsan_zarr")
    cmd = ["gcc", "-std=gnu11", "-g", "-O1", "-fsanitize=thread", "-I", os.path.join(root, "include"),
           "-I", os.path.join(root, "fysics_amd", "csrc"), "-I", os.path.join(root, "fysics_amd", "csrc", "host"),
           os.path.join(here, "tsan", "zarr_cache_threads.c"),
           os.path.join(root, "fysics_amd", "csrc", "host", "fy_zarr.c"), "-lpthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("ThreadSanitizer not available: " + r.stderr[:200])
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert r.stdout.startswith("ok"), r.stdout
