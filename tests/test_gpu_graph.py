"""The batched form captured into a HIP graph (fy_gpu_graph_cubes /
fy_graph_launch) against eager fy_gpu_process_cubes calls.

The graph is replayed several times with new input data written into the same
device buffers between replays (an all-air cube gains data, a cube with data
becomes all air), with eager work and synchronisations in between, and every
replay's u8 output must equal an eager call's on the same data bit for bit.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import fysics_amd as F
from fysics_amd import phantom
from test_gpu_batch import make_cases, run

pytestmark = pytest.mark.gpu


def refill(cases, rng, replay):
    for i, c in enumerate(cases):
        vol = phantom.sheets(c.u8.shape, rng, n_sheets=6)
        u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
        if (replay == 1 and i == 2) or (replay == 2 and i == 5):
            u8[:] = 0                                    # data -> all air, and back
        c.u8 = u8
        c.d_in.copy_(torch.from_numpy(u8))
    torch.cuda.synchronize()


def test_graph_replays_match_eager():
    rng = np.random.default_rng(21)
    cases = make_cases(rng)
    arr = (F.Cube * len(cases))(*[c.cube() for c in cases])
    stream = torch.cuda.Stream()
    lib = F.lib()
    torch.cuda.synchronize()
    g = lib.fy_gpu_graph_cubes(arr, len(cases), C.c_void_p(stream.cuda_stream))
    assert g, lib.fy_gpu_last_error()
    try:
        stream.synchronize()
        first = [c.d_out.cpu().numpy().copy() for c in cases]
        eager = run(cases)
        for a, b in zip(first, eager):
            assert np.array_equal(a, b)
        for replay in range(3):
            print(f"graph replay {replay}", flush=True)
            refill(cases, rng, replay)
            for c in cases:
                c.d_out.zero_()
            torch.cuda.synchronize()
            assert lib.fy_graph_launch(C.c_void_p(g)) == 0
            stream.synchronize()
            got = [c.d_out.cpu().numpy().copy() for c in cases]
            want = run(cases)                             # eager, other stream, same buffers
            for i, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), f"replay {replay} cube {i}"
            if replay == 1:
                assert not got[2].any()
            if replay == 2:
                assert not got[5].any() and got[2].any()
    finally:
        lib.fy_graph_destroy(C.c_void_p(g))


def _cache_stats():
    st = (C.c_long * 7)()
    F.lib().fy_gpu_table_cache(st, 0)
    return list(st)


def test_graph_tables_survive_cache_eviction():
    """Regression test of the table cache: a captured graph keeps the device
    tables its launches read (transfer tables, ring and z-drift blobs) while
    later calls fill the cache far past its limit and evict; every replay still
    equals the eager result bit for bit, and the evicted entries' memory is
    freed once their readers have finished."""
    lib = F.lib()
    rng = np.random.default_rng(31)
    cases = make_cases(rng)[:6]
    zf = [rng.uniform(0.9, 1.1, 64).astype(np.float32) for _ in cases]
    for c, f in zip(cases, zf):                       # z-drift blobs held by the graph too
        c.P.zdrift_factor = f.ctypes.data_as(F.c_float_p)
        c.P.vol_z = 64
    arr = (F.Cube * len(cases))(*[c.cube() for c in cases])
    stream = torch.cuda.Stream()
    old_limit = lib.fy_gpu_table_cache(None, 8)
    g = None
    try:
        torch.cuda.synchronize()
        g = lib.fy_gpu_graph_cubes(arr, len(cases), C.c_void_p(stream.cuda_stream))
        assert g, lib.fy_gpu_last_error()
        stream.synchronize()
        held = _cache_stats()[2]
        assert held >= len(cases)                     # per cube: its transfer table and its z-drift blob
        # 40 more tables through eager calls on a tiny cube: each brings a new z-drift blob
        tiny = phantom.noisy_u8(phantom.forward(phantom.sheets((12, 14, 16), rng, 2), **phantom.CONFIG1_PHYS), rng)
        d_t = torch.from_numpy(tiny).cuda()
        d_o = torch.empty_like(d_t)
        before = _cache_stats()
        for k in range(40):
            P = F.CubeParams()
            P.phys = F.physics(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                               unsharp_sigma=0.8, unsharp_coeff=0.5)
            P.deconv, P.deconv_param, P.deconv_cap = 1, 0.015, 8.0
            P.guided_eps, P.guided_radius, P.guided_s = 1e-3, 2, 2
            f = np.full(12, 1.0 + 0.001 * (k + 1), np.float32)
            P.zdrift_factor = f.ctypes.data_as(F.c_float_p)
            P.vol_z = 12
            cube = F.Cube(d_t.data_ptr(), 0, 0, 0, 12, 14, 16, 0, 0, 0, 12, 14, 16, d_o.data_ptr(), None,
                          C.pointer(P))
            assert lib.fy_gpu_process_cubes(C.byref(cube), 1, None) == 0, lib.fy_gpu_last_error()
        assert lib.fy_gpu_sync(None) == 0
        after = _cache_stats()
        print(f"table cache before {before} after {after}", flush=True)
        assert after[4] - before[4] >= 30              # evictions happened
        assert after[1] <= 8 + held                    # bounded: held entries are kept beyond the limit
        for replay in range(2):
            refill(cases, rng, replay)
            for c in cases:
                c.d_out.zero_()
            torch.cuda.synchronize()
            assert lib.fy_graph_launch(C.c_void_p(g)) == 0
            stream.synchronize()
            got = [c.d_out.cpu().numpy().copy() for c in cases]
            want = run(cases)
            for i, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), f"replay {replay} cube {i}"
    finally:
        if g:
            lib.fy_graph_destroy(C.c_void_p(g))
        torch.cuda.synchronize()
        lib.fy_gpu_table_cache(None, old_limit)
    end = _cache_stats()
    assert end[3] == 0                                 # every evicted entry's memory was freed


@pytest.mark.parametrize("radius,s", [(6, 2), (3, 1)])
def test_graph_wide_guided_radius(radius, s):
    """Guided radii beyond the LDS tiles (r/s > 2 on the fast path, r > 2 on
    the exact path) take the global-memory box passes, whose centring constant
    is read on the device: they capture like the rest, and every replay equals
    an eager call bit for bit."""
    rng = np.random.default_rng(30 + radius)
    cases = make_cases(rng)[:6]
    for c in cases:
        c.P.guided_radius, c.P.guided_s = radius, s
    arr = (F.Cube * len(cases))(*[c.cube() for c in cases])
    stream = torch.cuda.Stream()
    lib = F.lib()
    torch.cuda.synchronize()
    print(f"graph, guided r={radius} s={s}", flush=True)
    g = lib.fy_gpu_graph_cubes(arr, len(cases), C.c_void_p(stream.cuda_stream))
    assert g, lib.fy_gpu_last_error()
    try:
        for replay in range(2):
            refill(cases, rng, 0)
            for c in cases:
                c.d_out.zero_()
            torch.cuda.synchronize()
            assert lib.fy_graph_launch(C.c_void_p(g)) == 0
            stream.synchronize()
            got = [c.d_out.cpu().numpy().copy() for c in cases]
            want = run(cases)
            for i, (a, b) in enumerate(zip(got, want)):
                assert a.any() and np.array_equal(a, b), f"replay {replay} cube {i}"
    finally:
        lib.fy_graph_destroy(C.c_void_p(g))
