"""Stage (d) of the tile chain on the GPU: the air-zero value cut
(pipeline.c:302-399) and the connected-component despeckle (pipeline.c:227-250),
through fy_gpu_process_cubes and fy_process_buffer, against the oracle.

Bars: with the deconvolution and guided filter off, the float tile after stage
(d) is bit-identical to the oracle's (the scratch box sums, the blend and the
thresholds run in the reference's order, integer histogram and labels) and so
is the u8 output.  With the whole chain on, the zero mask comes from the
pre-deconvolution tile and is identical; the surviving voxels are within the
chain's usual tolerance (u8 +-1).  fy_valley_depth equals the reference's golden
values.
"""
import ctypes as C
import os

import numpy as np
import pytest
import torch

import clibs
import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
HALO = 4


def as_oracle(P):
    o = clibs.CubeParams()
    C.memmove(C.byref(o), C.byref(P), C.sizeof(o))
    return o


def air_params(cut_u8=-1, band=0, aggr=0.0, thresh=0.0, min_comp=0, passes=0, deconv=False, guided=False):
    P = F.CubeParams()
    P.air_zero = 1
    P.air_cut_u8, P.air_cut_band, P.air_cut_aggr, P.air_thresh = cut_u8, band, aggr, thresh
    P.air_min_component, P.scratch_passes = min_comp, passes
    if deconv:
        P.phys = F.physics(delta_beta=300.0 * 0.35, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                           unsharp_sigma=0.8, unsharp_coeff=0.5)
        P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, 1.17 ** 2
        P.window_lo, P.window_hi = -0.1, 1.2
    if guided:
        P.guided_eps, P.guided_radius, P.guided_s = 1.5e-3, 2, 2
    return P


def run_cube(u8, P, origin=(0, 0, 0), stream=None, halo=(HALO, HALO, HALO)):
    inner = tuple(n - 2 * h for n, h in zip(u8.shape, halo))
    inner_o = tuple(o + h for o, h in zip(origin, halo))
    d_in = torch.from_numpy(u8).cuda()
    d_out = torch.zeros(inner, dtype=torch.uint8, device="cuda")
    d_f = torch.zeros(u8.shape, dtype=torch.float32, device="cuda")
    cube = F.Cube(d_in.data_ptr(), *origin, *u8.shape, *inner_o, *inner, d_out.data_ptr(), d_f.data_ptr(),
                  C.pointer(P))
    st = stream or torch.cuda.Stream()
    torch.cuda.synchronize()
    assert F.lib().fy_gpu_process_cubes(C.byref(cube), 1, C.c_void_p(st.cuda_stream)) == 0, \
        F.lib().fy_gpu_last_error()
    st.synchronize()
    return d_out.cpu().numpy(), d_f.cpu().numpy()


def run_oracle(orc, u8, P, origin=(0, 0, 0), halo=(HALO, HALO, HALO)):
    inner = tuple(n - 2 * h for n, h in zip(u8.shape, halo))
    inner_o = tuple(o + h for o, h in zip(origin, halo))
    want = np.empty(inner, np.uint8)
    ftile = np.empty(u8.shape, np.float32)
    air = C.c_int(0)
    op = as_oracle(P)
    assert orc.orc_chain_cube(C.byref(op), clibs.up(u8), *origin, *u8.shape, *inner_o, *inner, clibs.up(want),
                              C.byref(air), clibs.fp(ftile)) == 0
    return want, ftile


def phantom_u8(shape, seed):
    rng = np.random.default_rng(seed)
    vol = phantom.sheets(shape, rng, n_sheets=6)
    return phantom.noisy_u8(phantom.forward(vol, **phantom.CONFIG1_PHYS), rng)


def test_valley_depth_golden(fy):
    for h, want in zip(GOLD["valley_hists"], GOLD["valley_out"]):
        h = np.ascontiguousarray(h, np.int64)
        dk, lt, vl = C.c_int(-7), C.c_int(-7), C.c_int(-7)
        d = fy.fy_valley_depth(h.ctypes.data_as(C.POINTER(C.c_long)), C.byref(dk), C.byref(lt), C.byref(vl))
        assert [d, dk.value, lt.value, vl.value] == list(want)


CASES = [dict(cut_u8=-1), dict(cut_u8=60, band=12, aggr=0.5), dict(cut_u8=60, band=30, aggr=1.5, thresh=0.12),
         dict(cut_u8=-1, min_comp=40, passes=3), dict(cut_u8=45, band=5, aggr=1.0, min_comp=300)]


@pytest.mark.parametrize("k", range(len(CASES)))
@pytest.mark.parametrize("shape", [(30, 34, 38), (11, 40, 33), (2, 30, 31)])
def test_air_stage_exact_vs_oracle(orc, k, shape):
    """Deconvolution and filter off: the tile after stage (d) and the u8 output
    are bit-identical to the oracle's (shape (2, ...) takes the full-resolution
    scratch branch)."""
    print(f"air stage case={k} shape={shape}", flush=True)
    u8 = phantom_u8(shape, 30 + k)
    P = air_params(**CASES[k])
    halo = (0 if shape[0] < 2 * HALO + 1 else HALO, HALO, HALO)
    got_q, got_f = run_cube(u8, P, halo=halo)
    want_q, want_f = run_oracle(orc, u8, P, halo=halo)
    assert np.array_equal(got_f, want_f)
    assert np.array_equal(got_q, want_q)
    if k == 0 and shape[0] > 2:
        assert (want_f == 0).mean() > 0.02              # the cut does zero the phantom's air


def test_air_chain_vs_oracle(orc):
    """Whole chain (dewindow -> deconv -> guided -> air-zero + despeckle -> u8)."""
    shape = (40, 44, 48)
    u8 = phantom_u8(shape, 41)
    P = air_params(cut_u8=-1, min_comp=60, deconv=True, guided=True)
    got_q, got_f = run_cube(u8, P, origin=(3, 5, 7))
    want_q, want_f = run_oracle(orc, u8, P, origin=(3, 5, 7))
    zero_g, zero_w = got_f == 0, want_f == 0
    assert (zero_g != zero_w).mean() <= 1e-4
    both = ~zero_g & ~zero_w
    rel = np.linalg.norm((got_f - want_f)[both]) / np.linalg.norm(want_f[both])
    assert rel <= 1e-5
    inner = (slice(HALO, -HALO),) * 3
    same = zero_g[inner] == zero_w[inner]
    assert np.abs(got_q.astype(int) - want_q.astype(int))[same].max() <= 1


def test_process_buffer_air_zero_vs_oracle(fy, orc):
    """fy_process_buffer with do_air_zero (host buffers) against the oracle."""
    u8 = phantom_u8((36, 40, 44), 52)
    cfg = F.pipeline_cfg(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                         unsharp_sigma=0.8, unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2,
                         do_deconv=0, guided_eps=0.0, do_air_zero=1, air_cut_u8=50, air_cut_band=20,
                         air_cut_aggr=0.8, air_min_component=25)
    inner = (28, 32, 36)
    got = np.empty(inner, np.uint8)
    air = C.c_int(-1)
    rc = fy.fy_process_buffer(C.byref(cfg), u8.ctypes.data_as(F.c_u8_p), 10, 20, 30, *u8.shape, 14, 24, 34, *inner,
                              300, 400, got.ctypes.data_as(F.c_u8_p), C.byref(air))
    assert rc == 0 and air.value == 0
    ocfg = clibs.PipelineCfg()
    C.memmove(C.byref(ocfg), C.byref(cfg), C.sizeof(ocfg))
    want = np.empty(inner, np.uint8)
    oair = C.c_int(-1)
    assert orc.orc_process_buffer(C.byref(ocfg), clibs.up(u8), 10, 20, 30, *u8.shape, 14, 24, 34, *inner, 300, 400,
                                  clibs.up(want), C.byref(oair)) == 0
    assert np.array_equal(got, want)


def test_despeckle_large_components_256(orc):
    """256^3 workload cube, stage (d) with despeckle: components spanning most of
    the tile (deep union-find trees) and many islands; bit-identical to the
    oracle's breadth-first labelling."""
    u8 = phantom.config2_volume((256, 256, 256), seed=2)
    P = air_params(cut_u8=-1, min_comp=500)
    got_q, got_f = run_cube(u8, P)
    want_q, want_f = run_oracle(orc, u8, P)
    assert np.array_equal(got_f, want_f)
    assert np.array_equal(got_q, want_q)


@pytest.mark.parametrize("shape,min_comp", [((26, 70, 150), 6), ((19, 33, 129), 40)])
def test_despeckle_many_islands(orc, shape, min_comp):
    """Oblique sheets in air plus speckle islands of one to a few voxels (about
    2000-7000 components, most of them small and interior), on rows longer than
    one wave and of odd lengths: the union-find over runs joined along y and z
    gives the oracle's zero mask bit for bit."""
    rng = np.random.default_rng(77)
    z, y, x = np.indices(shape)
    u8 = np.zeros(shape, np.uint8)
    sheet = ((z + 0.3 * y + 0.1 * x) % 12) < 3
    u8[sheet] = rng.integers(190, 220, size=int(sheet.sum()))
    speck = rng.random(shape) < 0.06
    u8[speck] = rng.integers(150, 256, size=int(speck.sum()))
    print(f"despeckle shape={shape} min_comp={min_comp}", flush=True)
    P = air_params(cut_u8=-1, min_comp=min_comp)
    got_q, got_f = run_cube(u8, P)
    want_q, want_f = run_oracle(orc, u8, P)
    assert np.array_equal(got_f, want_f)
    assert np.array_equal(got_q, want_q)
