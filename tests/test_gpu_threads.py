"""Concurrent host callers: the reference's pass-2 loop and mca_export call
the library from many threads at once (pipeline.c:1008-1046,
tools/mca_export.c:654).  Every host thread gets its own stream and
workspaces, the table caches and the tile-shape tuner are locked, so results
computed from 8 threads at once must equal the same calls made one after the
other, bit for bit."""
import ctypes as C
import threading

import numpy as np
import pytest

import fysics_amd as F
from fysics_amd import phantom

pytestmark = pytest.mark.gpu


def chain_cfg(k):
    c = F.PipelineCfg()
    for name, v in dict(delta_beta=150.0 + 60 * k, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                        unsharp_sigma=0.8, unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2, do_deconv=1,
                        guided_eps=1.2e-3 * (1 + k % 3), guided_subsample=0, no_dither=k % 2).items():
        setattr(c, name, v)
    return c


def cases():
    out = []
    for k in range(8):
        rng = np.random.default_rng(100 + k)
        shape = [(40, 44, 48), (36, 36, 36), (30, 50, 42)][k % 3]
        u8 = phantom.noisy_u8(phantom.forward(phantom.sheets(shape, rng, n_sheets=6), **phantom.CONFIG1_PHYS), rng)
        out.append((k, u8))
    return out


def run_buffer(k, u8):
    inner = tuple(s - 8 for s in u8.shape)
    got, air = F.process_buffer(chain_cfg(k), u8, (10 * k, 20, 30), (10 * k + 4, 24, 34), inner, (400, 500))
    return got


def run_deconv(k, u8):
    vol = u8.astype(np.float32) / 255.0
    return F.deconvolve(vol, F.physics(delta_beta=200.0 + 50 * k, energy_kev=110.0, distance_mm=1200.0,
                                       pixel_um=4.3, unsharp_sigma=0.8, unsharp_coeff=0.5), 0.015)


def run_guided(k, u8):
    return F.guided_fast(u8.astype(np.float32) / 255.0, radius=2, eps=1e-3 * (k + 1), s=2)


@pytest.mark.parametrize("fn", [run_buffer, run_deconv, run_guided])
def test_threads_match_sequential(fn):
    print(f"threads {fn.__name__}", flush=True)
    cs = cases()
    seq = [fn(k, u8) for k, u8 in cs]
    for rep in range(2):
        got = [None] * len(cs)
        errs = []

        def work(i):
            try:
                got[i] = fn(*cs[i])
            except Exception as e:          # reported below
                errs.append(repr(e))

        ts = [threading.Thread(target=work, args=(i,)) for i in range(len(cs))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errs, errs
        for i in range(len(cs)):
            assert np.array_equal(got[i], seq[i]), (fn.__name__, rep, i)
