"""Served form (fy_serve_*): requests of mixed sizes and physics pass through a
request loop with HBM-budget admission and continuous batching.

Bars: every request's u8 output is bit-identical to the same cube processed
alone (alone-vs-in-mix invariance), tickets retire in submission order, and a
budget that admits one request at a time changes nothing but the waves.
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


def serve(reqs, budget, max_batch=8, max_waves=2, interleave=False):
    lib = F.lib()
    stream = torch.cuda.Stream()
    sv = lib.fy_serve_create(budget, max_batch, max_waves, C.c_void_p(stream.cuda_stream))
    assert sv
    cubes = [r.cube() for r in reqs]
    done = (C.c_long * 64)()
    order = []
    try:
        i = 0
        while len(order) < len(reqs):
            k = 3 if interleave else len(reqs)
            for _ in range(k):
                if i < len(reqs):
                    assert lib.fy_serve_submit(sv, C.byref(cubes[i]), i) == 0
                    i += 1
            n = lib.fy_serve_poll(sv, done, 64)
            assert n >= 0, lib.fy_gpu_last_error()
            order += [done[j] for j in range(n)]
        assert lib.fy_serve_outstanding(sv) == 0
    finally:
        lib.fy_serve_destroy(sv)
    stream.synchronize()
    return order, [r.d_out.cpu().numpy().copy() for r in reqs]


@pytest.fixture(scope="module")
def reqs():
    rng = np.random.default_rng(5)
    return [Req(rng, SIZES[int(rng.integers(0, len(SIZES)))]) for _ in range(30)]


def test_served_matches_alone(reqs):
    want = alone(reqs)
    order, got = serve(reqs, 1 << 34)
    assert order == list(range(len(reqs)))
    for i, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), f"request {i}"


def test_served_interleaved_arrivals(reqs):
    want = alone(reqs)
    order, got = serve(reqs, 1 << 34, max_batch=5, max_waves=3, interleave=True)
    assert order == list(range(len(reqs)))
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_served_tight_budget(reqs):
    lib = F.lib()
    big = max(lib.fy_cube_workspace_bytes(C.byref(r.cube())) for r in reqs)
    want = alone(reqs)
    order, got = serve(reqs, big, max_batch=8, max_waves=2)
    assert order == list(range(len(reqs)))
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
