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
