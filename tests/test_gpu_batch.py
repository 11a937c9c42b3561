"""Batched chain (fy_gpu_process_cubes over many cubes of mixed shapes and
parameters, one launch per stage per shape group and FYG_BATCH cubes).

Bars: a cube's u8 output is bit-identical whether it runs alone or inside a
mixed batch (result invariance: no state is shared between cubes); against the
oracle's chain (orc_chain_cube, and its fp64 twin orc_chain_cube_f64) the u8
output is within +-1 LSB and the fp32 tile after the guided filter within
relative L2 1e-5.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import clibs
import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu

HALO = 4

# a filtered value within this distance of 0 may round to either side of the
# recenter's air test between two correct fp32 implementations (the float
# tiles agree to ~2e-6 in absolute terms)
AIR_EDGE = 8e-6


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def ring_model(shape, rng):
    """fy_dering detected on a ringed phantom (pass-1 accumulate + finalize)."""
    lib = F.lib()
    vol = phantom.sheets(shape, rng, n_sheets=10)
    u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
    u8, _ = phantom.add_rings(u8, rng, n_rings=8, amp=(5.0, 8.0))
    nz, ny, nx = shape
    d, t = F.Dering(), F.Dering()
    assert lib.fy_dering_init(C.byref(d), nz, ny, nx, -1.0, -1.0, 512, 8) == 0
    assert lib.fy_dering_tile_init(C.byref(t), C.byref(d)) == 0
    lib.fy_dering_accumulate_u8(C.byref(t), u8.ctypes.data_as(F.c_u8_p), 0, 0, nz, ny, nx, 1)
    lib.fy_dering_merge_tile(C.byref(d), 0, C.byref(t))
    lib.fy_dering_finalize(C.byref(d), 15, 20, 0.5, 6.0)
    lib.fy_dering_free(C.byref(t))
    return d, u8


def cube_params(rng, dering=None):
    """Per-cube physics of the batched workload: delta/beta U[100,1000], pixel in
    {2.4, 4.3, 7.9} um, sigma_z/sigma_xy U[1.0, 1.3] (Wiener, cap 8)."""
    P = F.CubeParams()
    P.phys = F.physics(delta_beta=float(rng.uniform(100, 1000)), energy_kev=110.0, distance_mm=1200.0,
                       pixel_um=float(rng.choice([2.4, 4.3, 7.9])), unsharp_sigma=0.8, unsharp_coeff=0.5)
    P.deconv, P.deconv_param, P.deconv_cap = 1, 0.015, 8.0
    P.fz2_scale = float(rng.uniform(1.0, 1.3)) ** 2
    P.window_lo, P.window_hi = -0.1, 1.2
    P.guided_eps, P.guided_radius, P.guided_s = float(rng.uniform(5e-4, 3e-3)), 2, 2
    if dering is not None:
        P.dering = C.pointer(dering)
    return P


def as_oracle(P):
    o = clibs.CubeParams()
    C.memmove(C.byref(o), C.byref(P), C.sizeof(o))
    return o


class Case:
    def __init__(self, u8, P, origin, want_float=False, halo=(HALO, HALO, HALO)):
        self.u8, self.P, self.origin = u8, P, origin
        self.inner_o = tuple(o + h for o, h in zip(origin, halo))
        self.inner = tuple(n - 2 * h for n, h in zip(u8.shape, halo))
        self.d_in = torch.from_numpy(u8).cuda()
        self.d_out = torch.zeros(self.inner, dtype=torch.uint8, device="cuda")
        self.d_float = torch.zeros(u8.shape, dtype=torch.float32, device="cuda") if want_float else None

    def cube(self):
        return F.Cube(self.d_in.data_ptr(), *self.origin, *self.u8.shape, *self.inner_o, *self.inner,
                      self.d_out.data_ptr(), self.d_float.data_ptr() if self.d_float is not None else None,
                      C.pointer(self.P))

    def oracle64(self, orc):
        want = np.empty(self.inner, np.uint8)
        ftile = np.empty(self.u8.shape, np.float64)
        op = as_oracle(self.P)
        assert orc.orc_chain_cube_f64(C.byref(op), clibs.up(self.u8), *self.origin, *self.u8.shape, *self.inner_o,
                                      *self.inner, clibs.up(want), clibs.dp(ftile)) == 0
        return want, ftile

    def oracle(self, orc):
        want = np.empty(self.inner, np.uint8)
        ftile = np.empty(self.u8.shape, np.float32)
        air = C.c_int(0)
        op = as_oracle(self.P)
        assert orc.orc_chain_cube(C.byref(op), clibs.up(self.u8), *self.origin, *self.u8.shape, *self.inner_o,
                                  *self.inner, clibs.up(want), C.byref(air), clibs.fp(ftile)) == 0
        return want, ftile, bool(air.value)


def run(cases):
    arr = (F.Cube * len(cases))(*[c.cube() for c in cases])
    stream = torch.cuda.Stream()
    assert F.lib().fy_gpu_process_cubes(arr, len(cases), C.c_void_p(stream.cuda_stream)) == 0, \
        F.lib().fy_gpu_last_error()
    stream.synchronize()
    return [c.d_out.cpu().numpy().copy() for c in cases]


def make_cases(rng, want_float=False):
    """22 cubes of three shapes, interleaved: shape A (18 cubes, more than one
    launch of FYG_BATCH = 16), shape B (2), shape C with ring removal (2); one
    cube of shape A is all air."""
    ring, _ = ring_model((24, 48, 56), np.random.default_rng(7))
    assert ring.have_rings, "the ringed phantom must exercise the ring-removal group"
    shapes = ["A", "B", "A", "C", "A", "A", "B", "A", "A", "C"] + ["A"] * 12
    dims = {"A": (20, 24, 28), "B": (24, 32, 40), "C": (16, 36, 40)}
    cases = []
    for i, k in enumerate(shapes):
        shp = dims[k]
        vol = phantom.sheets(shp, rng, n_sheets=6)
        u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
        if i == 5:
            u8[:] = 0
        P = cube_params(rng, ring if k == "C" else None)
        origin = (int(rng.integers(0, 8)), int(rng.integers(0, 12)), int(rng.integers(0, 16)))
        cases.append(Case(u8, P, origin, want_float))
    cases[0].ring = ring
    return cases


def test_batch_matches_alone_and_oracle(orc):
    rng = np.random.default_rng(11)
    cases = make_cases(rng)
    got = run(cases)
    for i, c in enumerate(cases):
        alone = run([c])[0]
        assert np.array_equal(got[i], alone), f"cube {i}: batched and alone differ"
        want, _, air = c.oracle(orc)
        if i == 5:
            assert air and not got[i].any()
            continue
        diff = np.abs(got[i].astype(int) - want.astype(int))
        assert diff.max() <= 1, f"cube {i}"
        assert (diff > 0).mean() < 2e-3, f"cube {i}"


def test_batch_float_tiles_vs_oracle(orc):
    rng = np.random.default_rng(12)
    cases = make_cases(rng, want_float=True)
    got = run(cases)
    for i, c in enumerate(cases):
        want, ftile, air = c.oracle(orc)
        if i == 5:
            continue
        gf = c.d_float.cpu().numpy()
        assert rel_l2(gf, ftile) <= 1e-5, f"cube {i}"
        diff = np.abs(got[i].astype(int) - want.astype(int))
        assert diff.max() <= 1, f"cube {i}"
        want64, f64 = c.oracle64(orc)
        assert rel_l2(gf, f64) <= 1e-5, f"cube {i} vs the fp64 oracle"
        d64 = np.abs(got[i].astype(int) - want64.astype(int))
        assert d64.max() <= 1, f"cube {i} vs the fp64 oracle"


def test_batch_order_invariance():
    """Reversing the submission order (which regroups the launches) leaves every
    cube's output bit-identical."""
    rng = np.random.default_rng(13)
    cases = make_cases(rng)
    a = run(cases)
    b = run(cases[::-1])[::-1]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def dithered_u8(v, origin):
    """The u8 quantiser of pipeline.c:53-65 (dither hash of the global voxel
    coordinates, v*255 + d rounded as two float32 operations, truncated and
    clamped to [0, 255]) applied to a float32 tile whose voxel (0,0,0) sits at
    global coordinates origin."""
    gz, gy, gx = np.meshgrid(*[np.arange(n, dtype=np.uint64) + o for n, o in zip(v.shape, origin)],
                             indexing="ij")
    m = np.uint64(0xFFFFFFFF)
    h = ((gz * np.uint64(73856093)) & m) ^ ((gy * np.uint64(19349663)) & m) ^ ((gx * np.uint64(83492791)) & m)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & m
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846CA68B)) & m
    h ^= h >> np.uint64(16)
    d = (h >> np.uint64(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    t = (v.astype(np.float32) * np.float32(255.0)).astype(np.float32) + d
    return np.where(t <= 0, 0, np.minimum(np.trunc(t), 255)).astype(np.uint8)


@pytest.mark.parametrize("halo", [(4, 4, 3), (3, 5, 5), (2, 2, 1), (4, 4, 7)])
def test_odd_inner_origin(orc, halo):
    """Inner tiles at odd x offsets in their region: with decimation s = 2,
    four consecutive voxels from an odd x span three low-res knots.  A cube
    with no float tile requested is filtered and quantised in one pass over its
    inner tile (the case under test); the same cube with its float tile is
    filtered over the whole region from x = 0 and quantised after.  Bars: the
    one-pass u8 equals the dithered quantisation of the float tile voxel for
    voxel; the same bits alone and batched; u8 within +-1 of the oracle's on
    < 0.2 % of voxels and the float tile within relative L2 1e-5 of the fp64
    oracle."""
    rng = np.random.default_rng(sum(halo) * 7 + halo[2])
    cases, fcases = [], []
    for shp in [(20, 26, 37), (18, 40, 64), (22, 24, 29)]:
        vol = phantom.sheets(shp, rng, n_sheets=6)
        u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
        P = cube_params(rng)
        origin = (int(rng.integers(0, 8)), int(rng.integers(0, 12)), int(rng.integers(0, 16)))
        cases.append(Case(u8, P, origin, halo=halo))
        fcases.append(Case(u8, P, origin, want_float=True, halo=halo))
    print(f"odd inner origin: halo {halo}", flush=True)
    got = run(cases)
    run(fcases)
    for i, (c, fc) in enumerate(zip(cases, fcases)):
        gf = fc.d_float.cpu().numpy()
        inner_f = gf[halo[0]:halo[0] + c.inner[0], halo[1]:halo[1] + c.inner[1], halo[2]:halo[2] + c.inner[2]]
        expect = dithered_u8(inner_f, c.inner_o)
        assert np.array_equal(fc.d_out.cpu().numpy(), expect), f"cube {i}: float-tile path"
        bad = got[i] != expect
        print(f"  cube {i} {c.u8.shape}: one-pass u8 differing from its float tile {bad.mean():.2e}", flush=True)
        assert not bad.any(), f"cube {i}: first at {np.argwhere(bad)[0]}"
        assert np.array_equal(got[i], run([c])[0]), f"cube {i}: batched and alone differ"
        want, _, _ = c.oracle(orc)
        diff = np.abs(got[i].astype(int) - want.astype(int))
        assert diff.max() <= 1 and (diff > 0).mean() < 2e-3, f"cube {i}"
        _, f64 = fc.oracle64(orc)
        assert rel_l2(gf, f64) <= 1e-5, f"cube {i}"


def make_mix_cases(rng, rings):
    """Eight cubes of five shapes whose padded lengths the single-launch form
    covers: 64^3 and 96^3 (fused plane passes), 64x128x128, 192x64x96 and
    128^3 with mixed axes (row passes); one cube all air.  rings: every cube
    with ring removal (the mixed launches then carry ring images)."""
    ring = None
    if rings:
        ring, _ = ring_model((24, 48, 56), np.random.default_rng(7))
        assert ring.have_rings
    dims = [(56, 60, 64), (80, 90, 96), (60, 100, 120), (130, 64, 70), (56, 60, 64), (100, 110, 128),
            (80, 90, 96), (56, 60, 64)]
    cases = []
    for i, shp in enumerate(dims):
        vol = phantom.sheets(shp, rng, n_sheets=6)
        u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
        if i == 4:
            u8[:] = 0
        P = cube_params(rng, ring)
        origin = (int(rng.integers(0, 8)), int(rng.integers(0, 12)), int(rng.integers(0, 16)))
        cases.append(Case(u8, P, origin))
    cases[0].ring = ring
    return cases


def kernels_run(fn):
    """names of the kernels (launch sites) that fn() ran"""
    lib = F.lib()
    torch.cuda.synchronize()
    lib.fy_gpu_profile_reset()
    lib.fy_gpu_profile(1)
    try:
        out = fn()
    finally:
        lib.fy_gpu_profile(0)
    prof = F.profile_read()
    lib.fy_gpu_profile_reset()
    return out, set(prof)


@pytest.mark.parametrize("rings", [False, True])
def test_mixed_launch_equals_grouped(orc, rings):
    """A batch of several shapes in one launch per stage (the default) against
    the same batch in grouped launches and each cube alone: the same bits.  The
    u8 output is within +-1 of the oracle's on < 0.2 % of voxels."""
    rng = np.random.default_rng(41 + rings)
    cases = make_mix_cases(rng, rings)
    lib = F.lib()
    print(f"mixed launches (rings {rings}): one launch per stage", flush=True)
    run(cases)                  # decides the radix plans of new lengths (timed with the per-shape kernels)
    mixed, names = kernels_run(lambda: run(cases))
    assert {"mix_xfwd", "mix_z", "mix_xinv", "mix_coef", "mix_up"} <= names, sorted(names)
    assert not any(n.startswith("rfft_x") or n == "zfused" or n == "guided_coef" for n in names), sorted(names)
    old = lib.fy_gpu_batch_mix(0)
    try:
        print("mixed launches: grouped", flush=True)
        grouped, gnames = kernels_run(lambda: run(cases))
        assert not any(n.startswith("mix_") for n in gnames)
    finally:
        lib.fy_gpu_batch_mix(old)
    for i, c in enumerate(cases):
        assert np.array_equal(mixed[i], grouped[i]), f"cube {i} {c.u8.shape}: mixed and grouped differ"
        assert np.array_equal(mixed[i], run([c])[0]), f"cube {i}: mixed and alone differ"
        if i == 4:
            assert not mixed[i].any()
            continue
        assert mixed[i].any()
        if i in (0, 3, 5):
            want, _, air = c.oracle(orc)
            diff = np.abs(mixed[i].astype(int) - want.astype(int))
            assert diff.max() <= 1 and (diff > 0).mean() < 2e-3, f"cube {i}"


def test_mixed_launch_options(orc):
    """The mixed launches with the chain's other per-cube options: z-drift
    factors (origin z0 != 0), recenter, no dither, decimation s = 4, odd inner
    tile origins in x.  Bars: the same bits as the grouped launches and as each
    cube alone; u8 within +-1 of the oracle on < 0.2 % of voxels."""
    rng = np.random.default_rng(57)
    vol_z = 400
    factor = rng.uniform(0.85, 1.15, vol_z).astype(np.float32)
    dims = [(56, 60, 64), (80, 90, 96), (60, 100, 120), (130, 64, 70), (64, 64, 63)]
    halos = [(4, 4, 3), (3, 5, 5), (4, 4, 4), (2, 2, 1), (4, 4, 7)]
    cases = []
    for i, (shp, halo) in enumerate(zip(dims, halos)):
        vol = phantom.sheets(shp, rng, n_sheets=6)
        u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
        P = cube_params(rng)
        P.zdrift_factor = factor.ctypes.data_as(F.c_float_p)
        P.vol_z = vol_z
        if i % 2 == 0:
            P.recenter, P.norm_lo, P.norm_hi = 1, 20, 230
        if i == 1:
            P.no_dither = 1
        if i == 2:
            P.guided_s = 4
        origin = (int(rng.integers(1, 200)), int(rng.integers(0, 12)), int(rng.integers(0, 16)))
        cases.append(Case(u8, P, origin, halo=halo))
    cases[0].factor = factor
    lib = F.lib()
    print("mixed launches with options", flush=True)
    run(cases)
    mixed, names = kernels_run(lambda: run(cases))
    assert {"mix_xfwd", "mix_z", "mix_xinv", "mix_coef", "mix_up"} <= names, sorted(names)
    old = lib.fy_gpu_batch_mix(0)
    try:
        grouped = run(cases)
    finally:
        lib.fy_gpu_batch_mix(old)
    for i, c in enumerate(cases):
        assert mixed[i].any(), f"cube {i}"
        assert np.array_equal(mixed[i], grouped[i]), f"cube {i} {c.u8.shape}: mixed and grouped differ"
        assert np.array_equal(mixed[i], run([c])[0]), f"cube {i}: mixed and alone differ"
        want, ftile, _ = c.oracle(orc)
        diff = np.abs(mixed[i].astype(int) - want.astype(int))
        # recenter keeps a voxel at 0 when its filtered value is <= 0 ("air stays
        # 0", pipeline.c:417) and lifts any positive value to ~0.05 (u8 ~13): where
        # the oracle's value lies within rounding of 0 either side may be air
        h = [o - r for o, r in zip(c.inner_o, c.origin)]
        pre = ftile[h[0]:h[0] + c.inner[0], h[1]:h[1] + c.inner[1], h[2]:h[2] + c.inner[2]]
        edge = (np.abs(pre) < AIR_EDGE) & ((mixed[i] == 0) | (want == 0))
        far = diff[~edge]
        print(f"  cube {i} {c.u8.shape}: u8 off by one {(far > 0).mean():.2e}, at the air edge "
              f"{int((edge & (diff > 1)).sum())}", flush=True)
        assert far.max() <= 1 and (far > 0).mean() < 2e-3, f"cube {i}"
        assert (edge & (diff > 1)).mean() < 1e-3, f"cube {i}"


def test_mixed_launch_many_cubes_and_fallback():
    """140 cubes of three shapes (more than one planning chunk of 128) in every
    launch form, and a batch where one cube's padded length (48) is outside
    the mixed kernels' cover, so the whole batch takes the grouped launches:
    every cube equals its run alone."""
    rng = np.random.default_rng(71)
    lib = F.lib()
    pool = []
    for shp in [(56, 60, 64), (60, 64, 60), (80, 90, 96)]:
        vol = phantom.sheets(shp, rng, n_sheets=4)
        pool.append(phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng))
    cases = []
    for i in range(140):
        u8 = pool[i % 3]
        P = cube_params(rng)
        origin = (int(rng.integers(0, 8)), int(rng.integers(0, 12)), int(rng.integers(0, 16)))
        cases.append(Case(u8, P, origin))
    alone = [run([c])[0] for c in cases]
    old = lib.fy_gpu_batch_mix(1)
    try:
        for mode in (1, 2, 0):
            lib.fy_gpu_batch_mix(mode)
            print(f"140 cubes, launch form {mode}", flush=True)
            got, names = kernels_run(lambda: run(cases))
            assert any(n.startswith("mix_") for n in names) == (mode != 0), sorted(names)
            for i in range(len(cases)):
                assert np.array_equal(got[i], alone[i]), f"mode {mode} cube {i}"
    finally:
        lib.fy_gpu_batch_mix(old)
    # one cube padded to 48 in z: the batch runs grouped
    vol = phantom.sheets((44, 60, 64), rng, n_sheets=4)
    odd = Case(phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng), cube_params(rng), (0, 0, 0))
    batch = cases[:5] + [odd]
    got, names = kernels_run(lambda: run(batch))
    assert not any(n.startswith("mix_") for n in names), sorted(names)
    for i, c in enumerate(batch):
        assert np.array_equal(got[i], run([c])[0]), f"fallback cube {i}"


def test_stream_release_and_thread_stream():
    """fy_gpu_stream_release drops the library's state of a caller's stream (its
    workspaces and the helper streams of the mixed launches) and of a thread's
    own stream (NULL); the calls after it rebuild that state and give the same
    bits."""
    import threading
    rng = np.random.default_rng(43)
    cases = make_mix_cases(rng, False)
    lib = F.lib()
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    arr = (F.Cube * len(cases))(*[c.cube() for c in cases])

    def call(h):
        for c in cases:
            c.d_out.zero_()
        torch.cuda.synchronize()
        assert lib.fy_gpu_process_cubes(arr, len(cases), h) == 0, lib.fy_gpu_last_error()
        assert lib.fy_gpu_sync(h) == 0
        return [c.d_out.cpu().numpy().copy() for c in cases]

    print("stream release: a caller's stream", flush=True)
    first = call(sh)
    assert sum(int(o.any()) for o in first) == len(cases) - 1        # cube 4 is all air
    for _ in range(3):
        assert lib.fy_gpu_stream_release(sh) == 0
        again = call(sh)
        assert all(np.array_equal(a, b) for a, b in zip(first, again))
    assert lib.fy_gpu_stream_release(sh) == 0
    assert lib.fy_gpu_stream_release(sh) == 0                        # nothing left to release
    print("stream release: a worker thread's own stream", flush=True)
    out = {}

    def worker():
        try:
            out["a"] = call(None)
            assert lib.fy_gpu_stream_release(None) == 0
            out["b"] = call(None)
            assert lib.fy_gpu_stream_release(None) == 0
        except BaseException as e:                                    # reported by the main thread
            out["err"] = e

    t = threading.Thread(target=worker)
    t.start()
    t.join()
    assert "err" not in out, out.get("err")
    for k in ("a", "b"):
        assert all(np.array_equal(a, b) for a, b in zip(first, out[k])), k


def test_helper_streams_past_the_queue_limit():
    """Six caller streams each running a grouped batch of five shapes (three
    helper streams apiece): the first twelve helpers get hardware queues of
    their own, the rest are ordinary streams; every stream's results are the
    same bits, and releasing the streams returns their queues."""
    rng = np.random.default_rng(44)
    cases = make_mix_cases(rng, False)
    lib = F.lib()
    streams = [torch.cuda.Stream() for _ in range(6)]
    outs = [[torch.zeros_like(c.d_out) for c in cases] for _ in streams]
    old = lib.fy_gpu_batch_mix(0)
    try:
        for k, (st, out) in enumerate(zip(streams, outs)):
            print(f"helper streams: caller stream {k}", flush=True)
            arr = (F.Cube * len(cases))(*[F.Cube(c.d_in.data_ptr(), *c.origin, *c.u8.shape, *c.inner_o, *c.inner,
                                                 o.data_ptr(), None, C.pointer(c.P)) for c, o in zip(cases, out)])
            assert lib.fy_gpu_process_cubes(arr, len(cases), C.c_void_p(st.cuda_stream)) == 0, \
                lib.fy_gpu_last_error()
        torch.cuda.synchronize()
    finally:
        lib.fy_gpu_batch_mix(old)
        for st in streams:
            assert lib.fy_gpu_stream_release(C.c_void_p(st.cuda_stream)) == 0
    first = [o.cpu().numpy() for o in outs[0]]
    assert sum(int(o.any()) for o in first) == len(cases) - 1
    for k, out in enumerate(outs[1:], 1):
        assert all(np.array_equal(a, o.cpu().numpy()) for a, o in zip(first, out)), f"stream {k}"
