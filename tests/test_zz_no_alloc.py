"""fy_gpu_deconvolve_noalloc: fy_gpu_deconvolve with its spectrum, weight
table, frequency axes and twiddles in the library's per-GPU scratch array, for
callers that cannot allocate device memory.

Bars: two calls in a row -- a 200x256x256 volume (padded 256^3, the largest
class that fits the 80 MiB spectrum) and then a smaller 40x52x70 one -- each
within relative L2 1e-5 of fy_gpu_deconvolve on the same input (and of the
fp64 oracle for the small one); an in-place call equals the out-of-place one
bit for bit; the calls leave the device's free memory unchanged (nothing is
allocated); a size whose spectrum exceeds the scratch is refused with an error
that names the limit.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import clibs
import fysics_amd as F

pytestmark = pytest.mark.gpu

PHYS = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3, unsharp_sigma=0.8,
            unsharp_coeff=0.5)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_noalloc_matches_existing_entry_point(orc):
    lib = F.lib()
    p = F.physics(**PHYS)
    rng = np.random.default_rng(17)
    shapes = [(200, 256, 256), (40, 52, 70)]
    xs = [rng.random(s, dtype=np.float32) for s in shapes]
    d_in = [torch.from_numpy(x).cuda() for x in xs]
    d_na = [torch.empty_like(d) for d in d_in]
    d_ref = [torch.empty_like(d) for d in d_in]
    d_io = d_in[1].clone()
    tiny = torch.rand((8, 8, 8), device="cuda")
    tiny_o = torch.empty_like(tiny)
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    torch.cuda.synchronize()
    print("noalloc: warm-up call 8^3", flush=True)
    assert lib.fy_gpu_deconvolve_noalloc(C.c_void_p(tiny.data_ptr()), C.c_void_p(tiny_o.data_ptr()), 8, 8, 8,
                                         C.byref(p), 0.015, sh) == 0, lib.fy_gpu_last_error()
    stream.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for s, a, o in zip(shapes, d_in, d_na):
        print(f"noalloc: {s}", flush=True)
        assert lib.fy_gpu_deconvolve_noalloc(C.c_void_p(a.data_ptr()), C.c_void_p(o.data_ptr()), *s, C.byref(p),
                                             0.015, sh) == 0, lib.fy_gpu_last_error()
    assert lib.fy_gpu_deconvolve_noalloc(C.c_void_p(d_io.data_ptr()), C.c_void_p(d_io.data_ptr()), *shapes[1],
                                         C.byref(p), 0.015, sh) == 0
    stream.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    print(f"noalloc: free device memory before {free0} after {free1}", flush=True)
    assert free0 - free1 < (1 << 20)
    for s, a, o in zip(shapes, d_in, d_ref):
        assert lib.fy_gpu_deconvolve(C.c_void_p(a.data_ptr()), C.c_void_p(o.data_ptr()), *s, C.byref(p), 0.015,
                                     sh) == 0, lib.fy_gpu_last_error()
    stream.synchronize()
    for s, na, rf in zip(shapes, d_na, d_ref):
        a, b = na.cpu().numpy(), rf.cpu().numpy()
        r = rel_l2(a, b)
        print(f"noalloc {s}: relL2 vs fy_gpu_deconvolve {r:.2e}, identical {np.mean(a == b):.6f}", flush=True)
        assert r <= 1e-5
    assert np.array_equal(d_io.cpu().numpy(), d_na[1].cpu().numpy())
    want = np.empty(shapes[1], np.float64)
    op = clibs.physics(**PHYS)
    assert orc.orc_deconv_radial_f64(clibs.fp(xs[1]), clibs.dp(want), *shapes[1], C.byref(op), 0.015, 0, 8.0,
                                     1.0) == 0
    assert rel_l2(d_na[1].cpu().numpy(), want) <= 1e-5


def test_noalloc_refuses_sizes_beyond_the_scratch():
    lib = F.lib()
    p = F.physics(**PHYS)
    d = torch.empty((300, 300, 300), dtype=torch.float32, device="cuda")
    rc = lib.fy_gpu_deconvolve_noalloc(C.c_void_p(d.data_ptr()), C.c_void_p(d.data_ptr()), 300, 300, 300,
                                       C.byref(p), 0.015, None)
    assert rc == 1
    msg = lib.fy_gpu_last_error().decode()
    assert "384x384x384" in msg and "scratch holds" in msg, msg
    assert lib.fy_gpu_deconvolve_noalloc(None, None, 0, 4, 4, C.byref(p), 0.015, None) == 0   # empty: nothing to do
