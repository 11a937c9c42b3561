"""Synthetic scroll-like volumes for tests and benchmarks.

A seeded generator of the workloads the benchmarks are quoted on:

* curved papyrus-like sheets: sinusoidal height fields
  z = z0 + A*sin(kx*x + p1)*sin(ky*y + p2), 3-6 voxels thick, attenuation
  U[0.3, 1.0] over a 0.05 background (fp32);
* the reconstruction forward model: spectrum times T(f)*U(f) -- the nabu
  Paganin low-pass and the unsharp boost, fysics.h:8-11 -- with an optional
  anisotropic z axis (f_z scaled by sigma_z/sigma_xy);
* Poisson-Gaussian noise var = g*I + b in u8^2 units (I in u8 units),
  optionally rising linearly from the rotation axis to the rim;
* quantisation to u8 through a physical window;
* optional concentric rings (u8 amplitude) about a known axis.

Random numbers come from numpy's PCG64 (numpy.random.default_rng(seed)).
This module is workload generation only: it never calls the product or the
oracle.
"""
import numpy as np

HC_KEV_UM = 1.23984199e-3


def transfer(fr, delta_beta, energy_kev, distance_mm, pixel_um, unsharp_coeff, unsharp_sigma):
    """T(f)*U(f) at radial frequency fr (cycles/voxel), float64 (paganin.c:17-37 without the floor)."""
    lam = HC_KEV_UM / energy_kev
    dist = distance_mm * 1000.0
    fu = fr / pixel_um
    t = 1.0 / (1.0 + delta_beta * lam * dist * np.pi * fu * fu)
    if unsharp_coeff and unsharp_sigma:
        t = t * (1.0 + unsharp_coeff * (1.0 - np.exp(-2.0 * np.pi ** 2 * unsharp_sigma ** 2 * fr * fr)))
    return t


def sheets(shape, rng, n_sheets=40, background=0.05):
    nz, ny, nx = shape
    vol = np.full(shape, background, np.float32)
    z = np.arange(nz, dtype=np.float32)[:, None, None]
    y = np.arange(ny, dtype=np.float32)[None, :, None]
    x = np.arange(nx, dtype=np.float32)[None, None, :]
    for _ in range(n_sheets):
        z0 = rng.uniform(0, nz)
        amp = rng.uniform(1.0, max(2.0, nz / 16))
        kx, ky = rng.uniform(0.02, 0.12, size=2)
        p1, p2 = rng.uniform(0, 2 * np.pi, size=2)
        thick = rng.uniform(3.0, 6.0)
        att = np.float32(rng.uniform(0.3, 1.0))
        h = (z0 + amp * np.sin(kx * x + p1) * np.sin(ky * y + p2)).astype(np.float32)
        mask = np.abs(z - h) <= thick / 2
        np.maximum(vol, np.where(mask, att, np.float32(0)), out=vol)
    return vol


def forward(vol, delta_beta, energy_kev, distance_mm, pixel_um, unsharp_coeff, unsharp_sigma,
            z_ratio=1.0):
    """Apply T*U in the Fourier domain (periodic), f_z scaled by z_ratio."""
    nz, ny, nx = vol.shape
    fz = np.fft.fftfreq(nz)[:, None, None] * z_ratio
    fy = np.fft.fftfreq(ny)[None, :, None]
    fx = np.fft.rfftfreq(nx)[None, None, :]
    fr = np.sqrt(fz * fz + fy * fy + fx * fx)
    h = transfer(fr, delta_beta, energy_kev, distance_mm, pixel_um, unsharp_coeff, unsharp_sigma)
    out = np.fft.irfftn(np.fft.rfftn(vol.astype(np.float64)) * h, s=vol.shape, axes=(0, 1, 2))
    return out.astype(np.float32)


def noisy_u8(phys, rng, window=(-0.1, 1.2), gain=0.02, floor=4.0, radial_rise=1.0, axis=None):
    """Window to u8 units, add N(0, g*I + b) noise (std scaled by a radial ramp), quantise."""
    lo, hi = window
    i_u8 = (phys.astype(np.float64) - lo) / (hi - lo) * 255.0
    var = gain * np.clip(i_u8, 0, None) + floor
    std = np.sqrt(var)
    if radial_rise != 1.0:
        nz, ny, nx = phys.shape
        cy, cx = axis if axis is not None else ((ny - 1) / 2.0, (nx - 1) / 2.0)
        yy, xx = np.meshgrid(np.arange(ny) - cy, np.arange(nx) - cx, indexing="ij")
        r = np.sqrt(yy * yy + xx * xx)
        std = std * (1.0 + (radial_rise - 1.0) * r / max(r.max(), 1.0))[None]
    v = i_u8 + std * rng.standard_normal(phys.shape)
    return np.clip(np.floor(v + 0.5), 0, 255).astype(np.uint8)


def add_rings(u8, rng, n_rings=20, amp=(3.0, 8.0), axis=None, width=0.8):
    nz, ny, nx = u8.shape
    cy, cx = axis if axis is not None else ((ny - 1) / 2.0, (nx - 1) / 2.0)
    yy, xx = np.meshgrid(np.arange(ny) - cy, np.arange(nx) - cx, indexing="ij")
    r = np.sqrt(yy * yy + xx * xx)
    rmax = min(cy, cx, ny - 1 - cy, nx - 1 - cx)
    radii = np.sort(rng.uniform(8.0, max(9.0, rmax - 4.0), size=n_rings))
    prof = np.zeros((ny, nx))
    for rad in radii:
        a = rng.uniform(*amp) * rng.choice([-1.0, 1.0])
        prof += a * np.exp(-(r - rad) ** 2 / (2 * width * width))
    out = np.clip(u8.astype(np.float64) + prof[None], 1, 255)
    out = np.where(u8 == 0, 0, out)
    return out.astype(np.uint8), radii


# --------------------------------------------------------------------------- named workloads
CONFIG1_PHYS = dict(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                    unsharp_coeff=0.5, unsharp_sigma=0.8)
WINDOW = (-0.1, 1.2)


def config1_volume(n=128, seed=1):
    """128^3 phantom: sheets -> Paganin+unsharp forward -> noise var=0.02 I+4 -> u8."""
    rng = np.random.default_rng(seed)
    vol = sheets((n, n, n), rng)
    phys = forward(vol, **CONFIG1_PHYS)
    return noisy_u8(phys, rng, WINDOW, 0.02, 4.0)


def config2_volume(shape=(256, 256, 256), seed=2, z_ratio=1.17, rings=20):
    """256^3 u8: anisotropic forward blur, radial noise rising 3x to the rim, concentric rings."""
    rng = np.random.default_rng(seed)
    vol = sheets(shape, rng)
    phys = forward(vol, z_ratio=z_ratio, **CONFIG1_PHYS)
    u8 = noisy_u8(phys, rng, WINDOW, 0.02, 4.0, radial_rise=3.0)
    if rings:
        u8, _ = add_rings(u8, rng, rings)
    return u8


# batched workload: 96 cubes of five shapes (169.7 M voxels), one seed and one
# physics draw per cube
BATCH_CLASSES = [(32, (64, 64, 64)), (32, (96, 96, 96)), (16, (128, 128, 128)), (12, (160, 160, 160)),
                 (4, (192, 256, 256))]


def batch_cube(shape, seed):
    """One cube of the batched workload: its own sheets, delta/beta U[100,1000],
    pixel in {2.4, 4.3, 7.9} um, sigma_z/sigma_xy U[1.0,1.3]; u8 through the
    window.  Returns (u8, physics dict, z_ratio)."""
    rng = np.random.default_rng(seed)
    phys = dict(CONFIG1_PHYS)
    phys["delta_beta"] = float(rng.uniform(100.0, 1000.0))
    phys["pixel_um"] = float(rng.choice([2.4, 4.3, 7.9]))
    zr = float(rng.uniform(1.0, 1.3))
    vol = sheets(shape, rng, n_sheets=max(4, shape[0] // 6))
    u8 = noisy_u8(forward(vol, z_ratio=zr, **phys), rng, WINDOW, 0.02, 4.0)
    return u8, phys, zr


# served workload: request sizes and their shares
SERVE_SIZES = [(0.4, 64), (0.3, 128), (0.2, 192), (0.1, 256)]


def serve_physics(rng):
    """Per-request physics of the served workload: delta/beta U[100,1000],
    E U[80,120] keV, D U[0.5,2] m, pixel U[2.4,7.9] um, unsharp coeff U[0.3,0.7],
    sigma U[0.6,1.2] px."""
    return dict(delta_beta=float(rng.uniform(100.0, 1000.0)), energy_kev=float(rng.uniform(80.0, 120.0)),
                distance_mm=float(rng.uniform(500.0, 2000.0)), pixel_um=float(rng.uniform(2.4, 7.9)),
                unsharp_coeff=float(rng.uniform(0.3, 0.7)), unsharp_sigma=float(rng.uniform(0.6, 1.2)))


def generate_many(fn, shapes, seeds, workers=8):
    """fn(shape, seed) for each pair, in spawned worker processes (they import
    numpy only), in order."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import os
    workers = max(1, min(workers, (os.cpu_count() or 2) - 1, len(shapes)))
    if workers == 1:
        return [fn(sh, sd) for sh, sd in zip(shapes, seeds)]
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(fn, shapes, seeds))


def batch_set(seed=100):
    """The 96 cubes of the batched workload (BATCH_CLASSES order), cube k with
    seed + k: a list of (shape, u8, physics dict, z_ratio)."""
    jobs = [shape for count, shape in BATCH_CLASSES for _ in range(count)]
    made = generate_many(batch_cube, jobs, [seed + k for k in range(len(jobs))])
    return [(shape, u8, phys, zr) for shape, (u8, phys, zr) in zip(jobs, made)]
