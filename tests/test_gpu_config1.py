"""The CPU-reference workload (128^3 phantom of 40 sheets, Paganin+unsharp
forward model, Poisson-Gaussian noise var = 0.02 I + 4, u8 through the window
[-0.1, 1.2]) restored end to end through the drop-in host API in the order a
reference user calls it: fy_u8_to_phys (dewindow) -> fy_deconvolve(reg 0.015)
-> fy_estimate_noise (win 5, 10th-percentile floor, ref 0.5) ->
fy_guided_eps_for_noise -> fy_guided_denoise (r = 2).

Bars, stage by stage against the oracle on the same inputs: the dewindowed
volume bit-exact; the deconvolution relL2 <= 1e-5 from the fp64 oracle (LUT
semantics); the noise model on the GPU's deconvolved volume equal to the
oracle's fit of the same volume to 1e-12 and the eps equal; the guided output
relL2 <= 1e-5 from the fp64 oracle filter of the same input; and the whole
chain relL2 <= 1e-5 from the oracle chain run in fp64 throughout.
"""
import ctypes as C

import numpy as np
import pytest

import clibs
import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_config1_restore_vs_oracle(orc):
    print("config 1 restore 128^3", flush=True)
    u8 = phantom.config1_volume(128, seed=1)
    shape, n = u8.shape, u8.size
    lo, hi = phantom.WINDOW
    lib = F.lib()

    f = np.empty(shape, np.float32)
    lib.fy_u8_to_phys(u8.ctypes.data_as(F.c_u8_p), F._fp(f), n, lo, hi)
    fo = np.empty(shape, np.float32)
    orc.orc_u8_to_phys(clibs.up(u8), clibs.fp(fo), n, lo, hi)
    assert np.array_equal(f, fo)

    phys = F.physics(**phantom.CONFIG1_PHYS)
    d = F.deconvolve(f, phys, 0.015)
    pc = clibs.physics(**phantom.CONFIG1_PHYS)
    d64 = np.empty(shape, np.float64)
    assert orc.orc_deconv_radial_f64(clibs.fp(f), clibs.dp(d64), *shape, C.byref(pc), 0.015, 0, 8.0, 1.0) == 0
    r_dec = rel(d, d64)
    assert r_dec <= 1e-5, r_dec

    nm, no = F.NoiseModel(), clibs.NoiseModel()
    assert lib.fy_estimate_noise(F._fp(d), *shape, 5, 10.0, 0.5, C.byref(nm)) == 0
    assert orc.orc_estimate_noise(clibs.fp(d), *shape, 5, 10.0, 0.5, C.byref(no)) == 0
    assert nm.n_bins_used == no.n_bins_used
    assert np.allclose([nm.g, nm.b, nm.noise_ref], [no.g, no.b, no.noise_ref], rtol=1e-12, atol=1e-18)
    eps = lib.fy_guided_eps_for_noise(nm.noise_ref)
    assert eps == orc.orc_guided_eps_for_noise(no.noise_ref)

    g = F.guided(d, radius=2, eps=eps)
    g64 = np.empty(shape, np.float64)
    assert orc.orc_guided_exact_f64(clibs.dp(d.astype(np.float64)), clibs.dp(g64), *shape, 2, eps) == 0
    r_g = rel(g, g64)
    assert r_g <= 1e-5, r_g

    # the whole chain against the oracle in fp64 throughout
    c64 = np.empty(shape, np.float64)
    assert orc.orc_guided_exact_f64(clibs.dp(d64), clibs.dp(c64), *shape, 2, eps) == 0
    r_c = rel(g, c64)
    print(f"config 1: deconv relL2 {r_dec:.2e}, guided {r_g:.2e}, chain {r_c:.2e}, eps {eps:.3e}", flush=True)
    assert r_c <= 1e-5, r_c
