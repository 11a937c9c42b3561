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


@pytest.mark.parametrize("shape,rings", [((64, 64, 64), False), ((96, 96, 96), False), ((40, 60, 64), True),
                                         ((70, 90, 96), True), ((60, 64, 64), False)])
def test_fft_planes_bit_identical(shape, rings):
    """The fused X+Y plane kernels (px == py in {64, 96}) against the
    separate X and Y passes: the same transforms, compiled into other kernels,
    so fp32 rounding may differ where the compiler fuses a multiply and an add
    differently.  Bars: float tile and fp32-input deconvolution
    (fy_gpu_deconvolve) within relative L2 1e-6 of each other, u8 output
    within +-1 on < 0.1 % of voxels."""
    from test_gpu_batch import ring_model
    lib = F.lib()
    rng = np.random.default_rng(sum(shape))
    if rings:
        ring, u8 = ring_model(shape, rng)
        assert ring.have_rings
    else:
        ring = None
        u8 = phantom.noisy_u8(phantom.forward(phantom.sheets(shape, rng, 6), **phantom.CONFIG1_PHYS), rng)
    P = F.CubeParams()
    P.phys = F.physics(delta_beta=400.0, energy_kev=100.0, distance_mm=900.0, pixel_um=4.3, unsharp_sigma=0.8,
                       unsharp_coeff=0.5)
    P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, 1.2 ** 2
    P.window_lo, P.window_hi = -0.1, 1.2
    P.guided_eps, P.guided_radius, P.guided_s = 2e-3, 2, 2
    if ring is not None:
        P.dering = C.pointer(ring)
    d_in = torch.from_numpy(u8).cuda()
    f_in = (d_in.float() / 255.0).contiguous()
    outs = {}
    old = lib.fy_gpu_fft_planes(1)
    try:
        for mode in (1, 0):
            lib.fy_gpu_fft_planes(mode)
            d_q = torch.zeros_like(d_in)
            d_f = torch.zeros(shape, dtype=torch.float32, device="cuda")
            d_d = torch.zeros(shape, dtype=torch.float32, device="cuda")
            z = (0, 0, 0)
            c = F.Cube(d_in.data_ptr(), *z, *shape, *z, *shape, d_q.data_ptr(), d_f.data_ptr(), C.pointer(P))
            print(f"fft planes mode {mode}: {shape} rings={rings}", flush=True)
            assert lib.fy_gpu_process_cubes(C.byref(c), 1, None) == 0, lib.fy_gpu_last_error()
            assert lib.fy_gpu_deconvolve(C.c_void_p(f_in.data_ptr()), C.c_void_p(d_d.data_ptr()), *shape,
                                         C.byref(P.phys), 0.015, None) == 0
            assert lib.fy_gpu_sync(None) == 0
            outs[mode] = (d_q.cpu().numpy(), d_f.cpu().numpy(), d_d.cpu().numpy())
    finally:
        lib.fy_gpu_fft_planes(old)
    (q1, f1, d1), (q0, f0, d0) = outs[1], outs[0]
    for a, b in ((f1, f0), (d1, d0)):
        r = np.linalg.norm((a - b).astype(np.float64)) / np.linalg.norm(b.astype(np.float64))
        print(f"  relL2 planes vs passes {r:.2e}", flush=True)
        assert r <= 1e-6
    du = np.abs(q1.astype(int) - q0.astype(int))
    assert du.max() <= 1 and (du > 0).mean() < 1e-3
    assert q1.any()


def _deconv(lib, x, shape, phys):
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.empty_like(d_in)
    torch.cuda.synchronize()
    assert lib.fy_gpu_deconvolve(C.c_void_p(d_in.data_ptr()), C.c_void_p(d_out.data_ptr()), *shape, C.byref(phys),
                                 0.015, None) == 0, lib.fy_gpu_last_error()
    assert lib.fy_gpu_sync(None) == 0
    return d_out.cpu().numpy()


RADIX_PHYS = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3, unsharp_sigma=0.8,
                  unsharp_coeff=0.5)


@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 128, 128), (100, 180, 250), (60, 64, 64), (200, 150, 96)])
def test_fft_radix_plans(orc, shape):
    """The two radix plans of the deconvolution passes (radix-16 stages, or
    radix-8 stages) on lengths 64..512: padded X/Y/Z of 64 (fused planes),
    128, 256/192/128 and 256/192/96 (96 has one plan).  Bars: the two results
    within relative L2 1e-6 of each other, each within 1e-5 of the fp64
    oracle (same LUT)."""
    import clibs
    lib = F.lib()
    rng = np.random.default_rng(sum(shape))
    x = rng.random(shape, dtype=np.float32)
    phys = F.physics(**RADIX_PHYS)
    outs = {}
    old = lib.fy_gpu_fft_radix(0)
    try:
        for mode in (0, 2):
            lib.fy_gpu_fft_radix(mode)
            print(f"fft radix mode {mode}: {shape}", flush=True)
            outs[mode] = _deconv(lib, x, shape, phys)
    finally:
        lib.fy_gpu_fft_radix(old)
    want = np.empty(shape, np.float64)
    op = clibs.physics(**RADIX_PHYS)
    assert orc.orc_deconv_radial_f64(clibs.fp(x), clibs.dp(want), *shape, C.byref(op), 0.015, 0, 8.0, 1.0) == 0
    a, b = outs[0].astype(np.float64), outs[2].astype(np.float64)
    r = np.linalg.norm(a - b) / np.linalg.norm(a)
    ra, rb = (np.linalg.norm(v - want) / np.linalg.norm(want) for v in (a, b))
    print(f"  relL2 plan 1 vs plan 0 {r:.2e}; vs fp64 oracle {ra:.2e} / {rb:.2e}", flush=True)
    assert r <= 1e-6
    assert ra <= 1e-5 and rb <= 1e-5


def test_fft_radix_tuned_and_persisted(tmp_path):
    """Tuning mode: the first deconvolution of a length times both plans of each
    pass and records the choice; later calls reuse it (bit-identical output), a
    saved and reloaded choice is used without timing (bit-identical again), and
    where every pass chose the same plan the output equals the run forced to
    that plan."""
    lib = F.lib()
    shape = (128, 128, 128)
    rng = np.random.default_rng(11)
    x = rng.random(shape, dtype=np.float32)
    phys = F.physics(**RADIX_PHYS)
    before = str(tmp_path / "before.txt").encode()
    assert lib.fy_gpu_tuning_save(before) == 0          # this process's decisions, restored below
    old = lib.fy_gpu_fft_radix(1)
    try:
        lib.fy_gpu_tuning_clear()
        print("fft radix tuning: 128^3", flush=True)
        first = _deconv(lib, x, shape, phys)
        lines = [l for l in F.tuning_report().splitlines() if l.startswith("fft ")]
        print("\n".join(lines), flush=True)
        plans = {}
        for kind in ("x", "y", "z"):
            l = [l for l in lines if l.startswith(f"fft {kind} 128:")]
            assert len(l) == 1 and "radix-16" in l[0] and "radix-8" in l[0]
            plans[kind] = int(l[0].split("plan ")[1].split()[0])
        assert np.array_equal(first, _deconv(lib, x, shape, phys))
        path = str(tmp_path / "tune.txt").encode()
        assert lib.fy_gpu_tuning_save(path) == 0
        saved = open(path.decode()).read()
        for kind in ("x", "y", "z"):
            assert f"fft\t{kind}\t128\t{plans[kind]}\n" in saved
        lib.fy_gpu_tuning_clear()
        assert lib.fy_gpu_tuning_load(path) == 0
        assert all("from a tuning file" in l for l in F.tuning_report().splitlines() if l.startswith("fft "))
        assert np.array_equal(first, _deconv(lib, x, shape, phys))
        if len(set(plans.values())) == 1:
            lib.fy_gpu_fft_radix(2 if plans["x"] else 0)
            assert np.array_equal(first, _deconv(lib, x, shape, phys))
    finally:
        lib.fy_gpu_tuning_clear()
        lib.fy_gpu_tuning_load(before)
        lib.fy_gpu_fft_radix(old)
