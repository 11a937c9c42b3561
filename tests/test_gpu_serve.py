"""Served form (fy_serve_*): requests of mixed sizes and physics pass through a
request loop with HBM-budget admission and continuous batching over several
lanes (streams) whose waves overlap on the device.

Bars: every request's u8 output is bit-identical to the same cube processed
alone (alone-vs-in-mix invariance), for device-resident and host-resident
requests, at small sizes and on the served workload's own mix (64^3-256^3,
per-request physics, 200 GB budget); every ticket retires exactly once; a
budget that admits one request at a time changes nothing but the waves; one
request of each size is within relative L2 1e-5 (float tile) and +-1 u8 of
the fp64 oracle.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu

SIZES = [(16, 16, 16), (24, 28, 32), (32, 32, 32), (20, 24, 28)]


class Req:
    def __init__(self, rng, shape):
        vol = phantom.sheets(shape, rng, n_sheets=5)
        self.u8 = phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)
        P = F.CubeParams()
        P.phys = F.physics(delta_beta=float(rng.uniform(100, 1000)), energy_kev=float(rng.uniform(80, 120)),
                           distance_mm=float(rng.uniform(500, 2000)), pixel_um=float(rng.uniform(2.4, 7.9)),
                           unsharp_sigma=float(rng.uniform(0.6, 1.2)), unsharp_coeff=float(rng.uniform(0.3, 0.7)))
        P.deconv, P.deconv_param, P.deconv_cap = 1, 0.015, 8.0
        P.fz2_scale = float(rng.uniform(1.0, 1.3)) ** 2
        P.window_lo, P.window_hi = -0.1, 1.2
        P.guided_eps, P.guided_radius, P.guided_s = 2e-3, 2, 2
        self.P = P
        self.d_in = torch.from_numpy(self.u8).cuda()
        self.d_out = torch.zeros(shape, dtype=torch.uint8, device="cuda")

    def cube(self):
        z = (0, 0, 0)
        return F.Cube(self.d_in.data_ptr(), *z, *self.u8.shape, *z, *self.u8.shape, self.d_out.data_ptr(), None,
                      C.pointer(self.P))


def alone(reqs):
    lib = F.lib()
    outs = []
    for r in reqs:
        c = r.cube()
        assert lib.fy_gpu_process_cubes(C.byref(c), 1, None) == 0
        torch.cuda.synchronize()
        outs.append(r.d_out.cpu().numpy().copy())
        r.d_out.zero_()
    torch.cuda.synchronize()
    return outs


def serve(reqs, budget, max_batch=8, lanes=2, interleave=False, host=()):
    """Requests through a server; those whose index is in `host` are submitted
    host-resident (pinned input and output buffers).  Returns the retired
    tickets in retirement order and every request's output."""
    lib = F.lib()
    stream = torch.cuda.Stream()
    sv = lib.fy_serve_create(budget, max_batch, lanes, C.c_void_p(stream.cuda_stream))
    assert sv
    cubes = [r.cube() for r in reqs]
    hin = {i: torch.from_numpy(reqs[i].u8).pin_memory() for i in host}
    hout = {i: torch.zeros(reqs[i].u8.shape, dtype=torch.uint8).pin_memory() for i in host}
    done = (C.c_long * 64)()
    order = []
    try:
        i = 0
        while len(order) < len(reqs):
            k = 3 if interleave else len(reqs)
            for _ in range(k):
                if i < len(reqs):
                    if i in hin:
                        rc = lib.fy_serve_submit_host(sv, C.byref(cubes[i]), C.c_void_p(hin[i].data_ptr()),
                                                      C.c_void_p(hout[i].data_ptr()), i)
                    else:
                        rc = lib.fy_serve_submit(sv, C.byref(cubes[i]), i)
                    assert rc == 0
                    i += 1
            n = lib.fy_serve_poll(sv, done, 64)
            assert n >= 0, lib.fy_gpu_last_error()
            order += [done[j] for j in range(n)]
        assert lib.fy_serve_outstanding(sv) == 0
    finally:
        lib.fy_serve_destroy(sv)
    stream.synchronize()
    outs = [hout[i].numpy().copy() if i in hout else r.d_out.cpu().numpy().copy() for i, r in enumerate(reqs)]
    return order, outs


@pytest.fixture(scope="module")
def reqs():
    rng = np.random.default_rng(5)
    return [Req(rng, SIZES[int(rng.integers(0, len(SIZES)))]) for _ in range(30)]


def test_served_matches_alone(reqs):
    want = alone(reqs)
    order, got = serve(reqs, 1 << 34)
    assert sorted(order) == list(range(len(reqs)))
    for i, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), f"request {i}"


def test_served_interleaved_arrivals(reqs):
    want = alone(reqs)
    order, got = serve(reqs, 1 << 34, max_batch=5, lanes=3, interleave=True)
    assert sorted(order) == list(range(len(reqs)))
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_served_tight_budget(reqs):
    lib = F.lib()
    big = max(lib.fy_cube_workspace_bytes(C.byref(r.cube())) for r in reqs)
    want = alone(reqs)
    order, got = serve(reqs, big, max_batch=8, lanes=2)
    assert sorted(order) == list(range(len(reqs)))
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_served_rejects_over_budget(reqs):
    lib = F.lib()
    sv = lib.fy_serve_create(4096, 4, 2, None)
    try:
        c = reqs[0].cube()
        assert lib.fy_cube_workspace_bytes(C.byref(c)) > 4096
        assert lib.fy_serve_submit(sv, C.byref(c), 0) == 1
        assert lib.fy_serve_outstanding(sv) == 0
    finally:
        lib.fy_serve_destroy(sv)


def test_served_host_resident(reqs):
    """Half the requests bring host buffers: the same bits as alone."""
    want = alone(reqs)
    order, got = serve(reqs, 1 << 34, max_batch=4, lanes=3, interleave=True, host=set(range(0, len(reqs), 2)))
    assert sorted(order) == list(range(len(reqs)))
    for i, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), f"request {i}"


@pytest.fixture(scope="module")
def mix5():
    """40 requests of the served workload: sizes drawn 40/30/20/10 % of 64^3,
    128^3, 192^3, 256^3 (each size present), one phantom per size, per-request
    physics (phantom.serve_physics) and sigma_z/sigma_xy U[1,1.3]."""
    rng = np.random.default_rng(55)
    sizes = [n for _, n in phantom.SERVE_SIZES]
    made = phantom.generate_many(phantom.batch_cube, [(n, n, n) for n in sizes], [500 + k for k in range(4)])
    vols = {n: u8 for n, (u8, _, _) in zip(sizes, made)}
    kinds = list(rng.choice(4, size=36, p=[p for p, _ in phantom.SERVE_SIZES])) + [0, 1, 2, 3]
    out = []
    for k in kinds:
        n = sizes[k]
        r = Req.__new__(Req)
        r.u8 = vols[n]
        P = F.CubeParams()
        P.phys = F.physics(**phantom.serve_physics(rng))
        P.deconv, P.deconv_param, P.deconv_cap = 1, 0.015, 8.0
        P.fz2_scale = float(rng.uniform(1.0, 1.3)) ** 2
        P.window_lo, P.window_hi = -0.1, 1.2
        P.guided_eps, P.guided_radius, P.guided_s = 2e-3, 2, 2
        r.P = P
        r.d_in = torch.from_numpy(r.u8).cuda()
        r.d_out = torch.zeros((n, n, n), dtype=torch.uint8, device="cuda")
        out.append(r)
    torch.cuda.synchronize()
    return out


def test_served_config5_mix_matches_alone(mix5):
    print("served config-5 mix: 40 requests, 64^3-256^3", flush=True)
    want = alone(mix5)
    order, got = serve(mix5, 200 << 30, max_batch=16, lanes=3, interleave=True,
                       host=set(range(1, len(mix5), 3)))
    assert sorted(order) == list(range(len(mix5)))
    for i, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), f"request {i} {mix5[i].u8.shape}"


@pytest.mark.parametrize("k", [36, 37, 38, 39])
def test_served_sizes_vs_fp64_oracle(mix5, orc, k):
    import clibs
    r = mix5[k]
    n = r.u8.shape[0]
    d_f = torch.zeros(r.u8.shape, dtype=torch.float32, device="cuda")
    d_q = torch.zeros_like(r.d_out)
    z = (0, 0, 0)
    c = F.Cube(r.d_in.data_ptr(), *z, n, n, n, *z, n, n, n, d_q.data_ptr(), d_f.data_ptr(), C.pointer(r.P))
    print(f"served request {k}: {n}^3 vs the fp64 oracle", flush=True)
    assert F.lib().fy_gpu_process_cubes(C.byref(c), 1, None) == 0
    assert F.lib().fy_gpu_sync(None) == 0
    op = clibs.CubeParams()
    C.memmove(C.byref(op), C.byref(r.P), C.sizeof(op))
    q64 = np.empty(r.u8.shape, np.uint8)
    f64 = np.empty(r.u8.shape, np.float64)
    assert orc.orc_chain_cube_f64(C.byref(op), clibs.up(r.u8), *z, n, n, n, *z, n, n, n, clibs.up(q64),
                                  clibs.dp(f64)) == 0
    gf = d_f.cpu().numpy()
    rel = float(np.linalg.norm(gf - f64) / np.linalg.norm(f64))
    d = np.abs(d_q.cpu().numpy().astype(int) - q64.astype(int))
    print(f"  relL2 {rel:.2e}, u8 off by one {(d > 0).mean():.2e}", flush=True)
    assert rel <= 1e-5
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
