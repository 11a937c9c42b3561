"""The tile-shape autotuner of the fast guided filter: every candidate (the
coefficient march's z chunk, the upsample's planes per workgroup) gives
bit-identical output, so the tuner's choice cannot change results; the tuned
choice is recorded per shape."""
import ctypes as C

import numpy as np
import pytest
import torch

import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,radius", [((40, 44, 52), 2), ((33, 70, 29), 4), ((256, 256, 256), 2)])
def test_candidates_bit_identical(shape, radius):
    print(f"tuning shape={shape} r={radius}", flush=True)
    lib = F.lib()
    rng = np.random.default_rng(sum(shape))
    vol = phantom.sheets(shape, rng, n_sheets=6)
    f = (vol + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    d_in = torch.from_numpy(f).cuda()
    outs = []
    old = lib.fy_gpu_tuning(0)
    try:
        for mode in (0, 2, 3, 4, 1):                 # first, forced 0/1/2, tuned
            lib.fy_gpu_tuning(mode)
            d_out = torch.empty_like(d_in)
            torch.cuda.synchronize()
            assert lib.fy_gpu_guided(C.c_void_p(d_in.data_ptr()), C.c_void_p(d_out.data_ptr()), *shape, radius,
                                     2e-3, 2, None) == 0
            assert lib.fy_gpu_sync(None) == 0
            outs.append(d_out.cpu().numpy())
    finally:
        lib.fy_gpu_tuning(old)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    rep = F.tuning_report()
    assert f"grid {(shape[0] + 1) // 2}x{(shape[1] + 1) // 2}x{(shape[2] + 1) // 2}" in rep, rep


def test_first_call_in_place():
    """A shape's first filter with the output over the guide: the tuner must not
    race the upsample candidates (each would reread a guide the previous one
    overwrote), and the result equals the out-of-place one."""
    shape, radius = (36, 30, 42), 2
    print(f"tuning in place shape={shape}", flush=True)
    lib = F.lib()
    rng = np.random.default_rng(7)
    f = (phantom.sheets(shape, rng, n_sheets=5) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    old = lib.fy_gpu_tuning(0)
    try:
        d_ref_in = torch.from_numpy(f).cuda()
        d_ref = torch.empty_like(d_ref_in)
        torch.cuda.synchronize()
        assert lib.fy_gpu_guided(C.c_void_p(d_ref_in.data_ptr()), C.c_void_p(d_ref.data_ptr()), *shape, radius,
                                 2e-3, 2, None) == 0
        assert lib.fy_gpu_sync(None) == 0
        lib.fy_gpu_tuning(1)
        d = torch.from_numpy(f).cuda()
        torch.cuda.synchronize()
        assert lib.fy_gpu_guided(C.c_void_p(d.data_ptr()), C.c_void_p(d.data_ptr()), *shape, radius, 2e-3, 2,
                                 None) == 0
        assert lib.fy_gpu_sync(None) == 0
    finally:
        lib.fy_gpu_tuning(old)
    assert np.array_equal(d.cpu().numpy(), d_ref.cpu().numpy())
    assert "region 36x30x42 fp32: coef zchunk" in F.tuning_report()


def test_tuned_shapes_persist(tmp_path):
    """Shapes tuned in one process can be saved and loaded by the next: a
    loaded shape runs its recorded candidates without timing them, with the
    same output."""
    shape = (28, 34, 30)
    print(f"tuning persist shape={shape}", flush=True)
    lib = F.lib()
    rng = np.random.default_rng(3)
    f = (phantom.sheets(shape, rng, n_sheets=4) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    d_in = torch.from_numpy(f).cuda()

    def run():
        d_out = torch.empty_like(d_in)
        torch.cuda.synchronize()
        assert lib.fy_gpu_guided(C.c_void_p(d_in.data_ptr()), C.c_void_p(d_out.data_ptr()), *shape, 2, 2e-3, 2,
                                 None) == 0
        assert lib.fy_gpu_sync(None) == 0
        return d_out.cpu().numpy()

    old = lib.fy_gpu_tuning(1)
    try:
        first = run()                                   # tunes this shape
        path = str(tmp_path / "tune.txt").encode()
        assert lib.fy_gpu_tuning_save(path) == 0
        lib.fy_gpu_tuning_clear()
        assert lib.fy_gpu_tuning_load(path) == 0
        key = "region 28x34x30 fp32"
        assert any(key in l and "from a tuning file" in l for l in F.tuning_report().splitlines())
        again = run()
        assert any(key in l and "from a tuning file" in l for l in F.tuning_report().splitlines())
    finally:
        lib.fy_gpu_tuning(old)
    assert np.array_equal(first, again)
