#!/usr/bin/env python3
"""Headline benchmark: Mvoxels/s of the deconvolution + guided-filter chain on a
256^3 u8 cube, inputs and outputs resident in HBM, on one MI355X.

Workload (the plain single-GPU configuration): a 256^3 synthetic scroll volume
(fysics_amd.phantom.config2_volume: curved sheets, Paganin+unsharp forward blur
with sigma_z/sigma_xy = 1.17, Poisson-Gaussian noise rising 3x to the rim, 20
concentric rings, u8 through the [-0.1, 1.2] window).  One step = the whole
cube through the device chain in fy_process_buffer order: u8 -> f01 -> ring
subtraction -> dewindow -> Paganin+unsharp Wiener deconvolution (delta/beta
back-off 0.35, reg 0.015, gain cap 8, anisotropic f_z scale 1.17^2) ->
rewindow/clamp -> guided filter (r=2, s=2, radially scaled eps) -> dithered u8.
Ring detection and the eps calibration are per-volume and run before timing.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported:
this build targets a single GPU.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mvoxels/s deconv+guided on 256³ cubes; p50/p99 cube latency served vs plain"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(name, n, p, s=2):
    """Algorithmic bytes one launch of `name` must move for an n^3 cube padded to
    p^3: each distinct input element read once, each output element written once.
    The half spectrum has H = p*p*(p/2+1) complex fp32 elements."""
    hx = p // 2 + 1
    N = n ** 3
    H = p * p * hx
    l = (n + s - 1) // s
    Nl = l ** 3
    table = {
        "rfft_x_u8": N + 8 * H,                 # u8 in, half spectrum out
        "cfft_y_fwd": 16 * H,                   # spectrum in and out
        "zfused": 16 * H,                       # spectrum in and out (weight fused)
        "cfft_y_inv": 16 * n * p * hx,          # planes z < n
        "irfft_x": 8 * n * n * hx + 4 * N,      # rows y < n of planes z < n in, fp32 tile out
        "guided_moments": 4 * Nl + 8 * Nl,      # decimated samples in, a/b out
        "guided_box_ab": 8 * Nl + 8 * Nl,       # a/b in, mean_a/mean_b out
        "guided_coef": 4 * Nl + 8 * Nl,         # decimated samples in, mean_a/mean_b out
        "guided_upsample": 4 * N + 8 * Nl + N,  # tile + coefficients in, u8 out
    }
    return table.get(name)


# HBM bytes per launch of each kernel on the 256^3 workload: rocprofv3
# FETCH_SIZE and WRITE_SIZE of the same bench.py command, scaled by the same
# counters read over a plain 1 GiB device copy (fetch reports 0.5 of the bytes
# moved, write 1.0), from the run summarised in
# profiles/chain_256cube_kernel_trace.md.  Used for roofline.traffic only.
MEASURED_TRAFFIC_256 = {
    "ring_image": 300186,
    "rfft_x_u8": 86689202,
    "cfft_y_fwd": 144820780,
    "zfused": 145043174,
    "cfft_y_inv": 144820780,
    "irfft_x": 142826081,
    "guided_coef": 42066697,
    "guided_upsample": 105213486,
}


def roofline_of(prof, n, p):
    """Roofline entry of the dominant kernel from per-launch-site HIP-event
    totals {site: (ms, launches)}.  The dominant kernel is the kernel function
    with the largest share of the step, as the kernel trace ranks them: the Y
    pass (k_cfft_cols) runs twice per cube, forward and inverse, so its two
    launch sites count together.  achieved = algorithmic bytes / event time."""
    funcs = {"cfft_y_fwd": "cfft_y", "cfft_y_inv": "cfft_y"}
    agg = {}
    for name, (ms, cnt) in prof.items():
        t = agg.setdefault(funcs.get(name, name), {"ms": 0.0, "cnt": 0, "bytes": 0, "sites": []})
        b = kernel_bytes(name, n, p)
        t["ms"] += ms
        t["cnt"] += cnt
        t["bytes"] = None if b is None or t["bytes"] is None else t["bytes"] + b * cnt
        t["sites"].append(name)
    dom = max(agg, key=lambda k: agg[k]["ms"])
    d = agg[dom]
    dom_bytes = d["bytes"] / max(d["cnt"], 1) if d["bytes"] else None
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["bytes"] else None
    traffic = None
    if n == 256 and all(x in MEASURED_TRAFFIC_256 for x in d["sites"]):
        traffic = round(sum(MEASURED_TRAFFIC_256[x] for x in d["sites"]) / len(d["sites"]))
    return {"kernel": dom, "launch_sites": d["sites"], "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic, "algorithmic_bytes": round(dom_bytes) if dom_bytes else None,
            "avg_us": round(1000.0 * d["ms"] / max(d["cnt"], 1), 2)}


def build_workload(n, seed):
    from fysics_amd import phantom
    t0 = time.time()
    vol = phantom.config2_volume((n, n, n), seed=seed)
    return vol, time.time() - t0


def calibrate(F, vol):
    """Per-volume calibration on the GPU: ring detection (pipeline pass 1:
    accumulate with ss=2 over slabs of 512, finalize(15,100,0.5,6.0))."""
    lib = F.lib()
    nz, ny, nx = vol.shape
    d = F.Dering()
    t = F.Dering()
    assert lib.fy_dering_init(C.byref(d), nz, ny, nx, -1.0, -1.0, 512, 8) == 0
    assert lib.fy_dering_tile_init(C.byref(t), C.byref(d)) == 0
    lib.fy_dering_accumulate_u8(C.byref(t), vol.ctypes.data_as(F.c_u8_p), 0, 0, nz, ny, nx, 2)
    lib.fy_dering_merge_tile(C.byref(d), 0, C.byref(t))
    found = lib.fy_dering_finalize(C.byref(d), 15, 100, 0.5, 6.0)
    lib.fy_dering_free(C.byref(t))
    return d, found


def eps_for_cube(n):
    """Radially scaled guided eps of the generator's noise model: flat-region
    noise std fn(r) = fn0*(1 + 2 r/rmax) (rising 3x to the rim), guided_eps =
    (4.2*fn_med)^2, per-tile scale clamp((fn(r_tile)/fn_med)^2, 0.4, 8)
    (pipeline.c:447-456, 878-891)."""
    fn0 = np.sqrt(0.02 * 110.0 + 4.0) / 255.0
    rmax = np.sqrt(2) * (n - 1) / 2
    fn = lambda r: fn0 * (1 + 2 * r / rmax)
    fn_med = fn(0.5 * rmax)
    guided_eps = (4.2 * fn_med) ** 2
    s2 = min(max((fn(0.0) / fn_med) ** 2, 0.4), 8.0)      # the cube is centred on the axis
    return guided_eps * s2


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--prime-s", type=float, default=0.5,
                    help="seconds of untimed steps before the warm-up steps (clock ramp)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true", help="skip the 96-cube batched measurement")
    ap.add_argument("--no-served", action="store_true", help="skip the served-request measurement")
    ap.add_argument("--no-precision", action="store_true", help="skip the 512^3 fp32 / fp64 timing")
    ap.add_argument("--no-calibration", action="store_true", help="skip the pass-1 statistics timing")
    ap.add_argument("--no-air-zero", action="store_true", help="skip the chain-with-air-zero timing")
    ap.add_argument("--no-lod", action="store_true", help="skip the downscale / phase-correlation timing")
    ap.add_argument("--no-export", action="store_true", help="skip the zarr -> GPU -> zarr export timing")
    ap.add_argument("--served-requests", type=int, default=20000)
    ap.add_argument("--served-rate", type=float, default=400.0)
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the 96-cube batched measurement and print its JSON (for profiling)")
    ap.add_argument("--batched-form", default="",
                    help="with --batched-only: time only this launch form (mixed_parts, mixed_whole, grouped, ...)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference point that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters into HBM bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.batched_only:
        import fysics_amd as F
        out = batched_bench(F, reps=args.steps, only=args.batched_form)
        out["fft_radix_plans"] = [l for l in F.tuning_report().splitlines() if l.startswith("fft ")]
        print(json.dumps(out))
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(cubes are independent; see DESIGN.md for how the work would shard)", file=sys.stderr)
        sys.exit(2)

    import torch
    import fysics_amd as F
    torch.cuda.set_device(0)
    lib = F.lib()

    n = args.size
    vol, gen_s = build_workload(n, args.seed)
    ring, found = calibrate(F, vol)
    eps = eps_for_cube(n)
    phys = F.physics(delta_beta=300.0 * 0.35, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                     unsharp_sigma=0.8, unsharp_coeff=0.5)
    prm = F.CubeParams()
    prm.phys = phys
    prm.deconv, prm.deconv_param, prm.deconv_cap, prm.fz2_scale = 1, 0.015, 8.0, 1.17 ** 2
    prm.window_lo, prm.window_hi = -0.1, 1.2
    prm.guided_eps, prm.guided_radius, prm.guided_s = eps, 2, 2
    prm.dering = C.pointer(ring)

    # a dedicated (non-null) stream: the library enqueues on it, events time it
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    d_in = torch.from_numpy(vol).cuda()
    d_out = torch.empty_like(d_in)
    cube = F.Cube(d_in.data_ptr(), 0, 0, 0, n, n, n, 0, 0, 0, n, n, n, d_out.data_ptr(), None, C.pointer(prm))

    def step():
        rc = lib.fy_gpu_process_cubes(C.byref(cube), 1, sh)
        if rc:
            raise RuntimeError("fy_gpu_process_cubes: " + lib.fy_gpu_last_error().decode())

    # the GPU's clocks ramp up over the first fraction of a second of load: the
    # chain runs untimed for at least --prime-s seconds before the warm-up steps
    t_prime = time.perf_counter()
    primed = 0
    while time.perf_counter() - t_prime < args.prime_s:
        for _ in range(20):
            step()
        torch.cuda.synchronize()
        primed += 20
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: K steps, inputs resident in HBM
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    mvox = n ** 3 / 1e6
    value = mvox * args.steps / wall

    # per-cube latency distribution (events around each step on the same stream)
    lat = []
    for _ in range(200):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step()
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_time(b))
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]

    # per-kernel HIP-event timing on the launch stream -> roofline of the dominant kernel
    lib.fy_gpu_profile_reset()
    lib.fy_gpu_profile(1)
    for _ in range(min(args.steps, 50)):
        step()
    torch.cuda.synchronize()
    lib.fy_gpu_profile(0)
    prof = F.profile_read()
    p = lib.fy_next_fft_size(n)
    kernels = {}
    for name, (ms, cnt) in prof.items():
        avg_us = 1000.0 * ms / max(cnt, 1)
        b = kernel_bytes(name, n, p)
        kernels[name] = {"avg_us": round(avg_us, 2), "launches": cnt,
                         "GBps": round(b / (avg_us * 1e-6) / 1e9, 1) if b else None}
    roofline = roofline_of(prof, n, p)

    # plain drop-in with host buffers (PCIe-inclusive; reported beside, never `value`)
    cfgc = F.pipeline_cfg(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                          unsharp_sigma=0.8, unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2,
                          do_deconv=1, guided_eps=eps, have_dering=1)
    cfgc.dering = C.pointer(ring)
    host_out = np.empty_like(vol)
    air = C.c_int()
    host_t = []
    for i in range(4):
        t1 = time.perf_counter()
        rc = lib.fy_process_buffer(C.byref(cfgc), vol.ctypes.data_as(F.c_u8_p), 0, 0, 0, n, n, n, 0, 0, 0,
                                   n, n, n, n, n, host_out.ctypes.data_as(F.c_u8_p), C.byref(air))
        host_t.append(time.perf_counter() - t1)
        assert rc == 0
    host_rate = mvox / statistics.median(host_t[1:])

    batched = None if args.no_batched else batched_bench(F)
    served = None if args.no_served else served_bench(F, args.served_requests, args.served_rate)
    # capacity: the same request mix offered far above what either loop sustains
    # (throughput then is the loop's capacity), host-resident and HBM-resident
    stress = None
    if not args.no_served:
        stress = {"host": served_bench(F, 6000, 60000.0, seed=8, host=True),
                  "hbm": served_bench(F, 6000, 60000.0, seed=9, host=False)}
    precision = None if args.no_precision else precision_bench(F, vol)
    calib = None if args.no_calibration else calibration_bench(F, vol)
    airz = None if args.no_air_zero else air_zero_bench(F, vol, eps)
    lod = None if args.no_lod else lod_phasecorr_bench(F, vol)
    export = None if args.no_export else export_bench(F, vol, eps)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(vol, phys, eps, ring, args.cpu_threads)
        # the plain CPU reference path of the 128^3 config-1 volume, beside the drop-in
        cpu["config1_128"] = config1_bench(F)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mvoxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "prime": {"seconds": args.prime_s, "steps": primed},
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 u8 cube, device chain: rings -> dewindow -> Paganin+unsharp Wiener "
                               "deconv (sigma_z/sigma_xy 1.17, delta/beta x0.35, reg 0.015) -> guided r=2 s=2 "
                               "radial eps -> dithered u8", "cube": [n, n, n], "fft": [p, p, p],
                   "rings_detected": int(found), "guided_eps": eps, "parallelism": "single GPU"},
        "latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4), "samples": len(lat)},
        "roofline": roofline,
        "kernels": kernels,
        "host_api_mvox_s": round(host_rate, 1),
        "batched": batched,
        "served": served,
        "served_capacity": stress,
        "precision_512": precision,
        "calibration": calib,
        "air_zero": airz,
        "lod_phasecorr": lod,
        "export": export,
        "cpu_baseline": cpu,
        "fft_radix_plans": [l for l in F.tuning_report().splitlines() if l.startswith("fft ")],
        "workload_gen_s": round(gen_s, 1),
    }
    print(json.dumps(line))


def _generate(fn, shapes, seeds):
    """Synthetic cubes in worker processes (spawned: they import numpy only)."""
    from fysics_amd import phantom
    return phantom.generate_many(fn, shapes, seeds)


def _cube_params(F, phys, z_ratio, eps):
    P = F.CubeParams()
    P.phys = F.physics(**phys)
    P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, z_ratio ** 2
    P.window_lo, P.window_hi = -0.1, 1.2
    P.guided_eps, P.guided_radius, P.guided_s = eps, 2, 2
    return P


def batched_bench(F, reps=20, seed=100, only=""):
    """The batched workload: 96 cubes (32x64^3, 32x96^3, 16x128^3, 12x160^3 -> 192^3
    FFT, 4x192x256x256; 169.7 M voxels), each with its own phantom and physics,
    u8 in HBM -> deconv + guided -> u8 in HBM.  Timed as one fy_gpu_process_cubes
    call over all 96 in each launch form (fy_gpu_batch_mix): one launch per
    stage per cache-sized part of mixed shapes, parts on two streams
    ("mixed_parts", the default), one launch per stage for the whole batch
    ("mixed_whole"), one launch per stage per shape group ("grouped"); as the
    replay of the default and the grouped form captured in a HIP graph
    (fy_gpu_graph_cubes / fy_graph_launch); and, beside them, one call per
    cube."""
    import torch
    from fysics_amd import phantom
    lib = F.lib()
    t0 = time.time()
    keep, cubes = [], []
    jobs = [shape for count, shape in phantom.BATCH_CLASSES for _ in range(count)]
    made = _generate(phantom.batch_cube, jobs, [seed + k for k in range(len(jobs))])
    for shape, (u8, phys, zr) in zip(jobs, made):
        P = _cube_params(F, phys, zr, 2e-3)
        d_in = torch.from_numpy(u8).cuda()
        d_out = torch.empty_like(d_in)
        keep += [P, d_in, d_out]
        z = (0, 0, 0)
        cubes.append(F.Cube(d_in.data_ptr(), *z, *shape, *z, *shape, d_out.data_ptr(), None, C.pointer(P)))
    gen_s = time.time() - t0
    nvox = sum(c.hz * c.hy * c.hx for c in cubes)
    arr = (F.Cube * len(cubes))(*cubes)
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)

    def batch():
        if lib.fy_gpu_process_cubes(arr, len(cubes), sh):
            raise RuntimeError(lib.fy_gpu_last_error().decode())

    def one_by_one():
        for c in cubes:
            if lib.fy_gpu_process_cubes(C.byref(c), 1, sh):
                raise RuntimeError(lib.fy_gpu_last_error().decode())

    out = {"workload": "96 cubes: 32x64^3, 32x96^3, 16x128^3, 12x160^3 (FFT 192^3), 4x192x256x256; "
                       "per-cube phantom, delta/beta U[100,1000], pixel {2.4,4.3,7.9} um, sigma_z/sigma_xy "
                       "U[1,1.3]; u8 -> dewindow -> Wiener deconv -> guided r=2 s=2 -> dithered u8",
           "cubes": len(cubes), "voxels": int(nvox), "workload_gen_s": round(gen_s, 1)}
    graphs = {}

    def captured(mode):
        def fn():
            if mode not in graphs:                 # recorded under the form's mode on first use
                graphs[mode] = lib.fy_gpu_graph_cubes(arr, len(cubes), sh)
                if not graphs[mode]:
                    raise RuntimeError(lib.fy_gpu_last_error().decode())
            if lib.fy_graph_launch(C.c_void_p(graphs[mode])):
                raise RuntimeError(lib.fy_gpu_last_error().decode())
        return fn

    runs = (("mixed_parts", 1, batch), ("mixed_parts_graph", 1, captured(1)), ("mixed_whole", 2, batch),
            ("grouped", 0, batch), ("grouped_graph", 0, captured(0)), ("one_call_per_cube", 1, one_by_one))
    old = lib.fy_gpu_batch_mix(1)
    try:
        for name, mode, fn in runs:
            if only and name != only:
                continue
            lib.fy_gpu_batch_mix(mode)
            for _ in range(3):
                fn()
            stream.synchronize()
            t1 = time.perf_counter()
            for _ in range(reps):
                fn()
            stream.synchronize()
            dt = (time.perf_counter() - t1) / reps
            out[name] = {"ms_per_batch": round(1000 * dt, 3), "mvox_s": round(nvox / dt / 1e6, 1)}
    finally:
        lib.fy_gpu_batch_mix(old)
        for g in graphs.values():
            lib.fy_graph_destroy(C.c_void_p(g))
    return out


_SERVE_POOL = {}


def served_bench(F, n_req=20000, rate=400.0, seed=7, pool=4, eps=2e-3, host=True, slots=48, modes=None):
    """The served workload: open-loop Poisson arrivals at `rate` req/s, sizes 40%
    64^3, 30% 128^3, 20% 192^3, 10% 256^3, per-request physics (delta/beta, E, D,
    pixel, unsharp, sigma_z/sigma_xy) so every request brings its own transfer
    table; inputs from a pool of `pool` phantoms per size.  host=True: each
    request's input and output are pinned host buffers, so its latency includes
    both PCIe transfers; host=False: they are resident in HBM.  Latency =
    completion seen by the polling thread - scheduled arrival.  'served' is the
    request loop with continuous batching (waves of up to 16 requests on 3
    lanes -- streams whose waves overlap -- 200 GB budget); 'plain' is one cube
    at a time (waves of one on one lane), the plain one-cube-at-a-time path."""
    import torch
    from fysics_amd import phantom
    lib = F.lib()
    rng = np.random.default_rng(seed)
    sizes = [n for _, n in phantom.SERVE_SIZES]
    probs = [p for p, _ in phantom.SERVE_SIZES]
    jobs = [(n, n, n) for n in sizes for _ in range(pool)]
    if pool not in _SERVE_POOL:             # the same phantoms for every served measurement
        _SERVE_POOL[pool] = _generate(phantom.batch_cube, jobs, [700 + k for k in range(len(jobs))])
    made = _SERVE_POOL[pool]
    inputs = {n: [] for n in sizes}
    for shape, (u8, _, _) in zip(jobs, made):
        t = torch.from_numpy(u8)
        inputs[shape[0]].append(t.pin_memory() if host else t.cuda())
    if host:
        outs = {n: torch.empty((slots, n, n, n), dtype=torch.uint8).pin_memory() for n in sizes}
    else:
        outs = {n: torch.empty((slots, n, n, n), dtype=torch.uint8, device="cuda") for n in sizes}
    kinds = rng.choice(len(sizes), size=n_req, p=probs)
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    params = []
    for i in range(n_req):
        P = _cube_params(F, phantom.serve_physics(rng), float(rng.uniform(1.0, 1.3)), eps)
        params.append(P)
    cubes, hin, hout = [], [], []
    for i in range(n_req):
        n = sizes[kinds[i]]
        src = inputs[n][i % pool]
        dst = outs[n][i % slots]
        z = (0, 0, 0)
        if host:
            cubes.append(F.Cube(None, *z, n, n, n, *z, n, n, n, None, None, C.pointer(params[i])))
            hin.append(C.c_void_p(src.data_ptr()))
            hout.append(C.c_void_p(dst.data_ptr()))
        else:
            cubes.append(F.Cube(src.data_ptr(), *z, n, n, n, *z, n, n, n, dst.data_ptr(), None, C.pointer(params[i])))
    nvox = float(sum(sizes[k] ** 3 for k in kinds))
    torch.cuda.synchronize()

    def submit(sv, i, ticket):
        if host:
            return lib.fy_serve_submit_host(sv, C.byref(cubes[i]), hin[i], hout[i], ticket)
        return lib.fy_serve_submit(sv, C.byref(cubes[i]), ticket)

    def run(max_batch, lanes):
        stream = torch.cuda.Stream()
        sv = lib.fy_serve_create(200 << 30, max_batch, lanes, C.c_void_p(stream.cuda_stream))
        if not sv:
            raise RuntimeError("fy_serve_create")
        buf = (C.c_long * 4096)()
        # warm every lane's workspaces with a full wave of the largest size
        big = [j for j in range(n_req) if kinds[j] == len(sizes) - 1][:max_batch * lanes]
        for j in big + [int(np.flatnonzero(kinds == k)[0]) for k in range(len(sizes))]:
            submit(sv, j, -1)
        while lib.fy_serve_outstanding(sv):
            if lib.fy_serve_poll(sv, buf, 4096) < 0:
                raise RuntimeError(lib.fy_gpu_last_error().decode())
        done_t = np.zeros(n_req)
        t0 = time.perf_counter()
        i = nd = 0
        while nd < n_req:
            now = time.perf_counter() - t0
            while i < n_req and arrive[i] <= now:
                if i >= slots and done_t[i - slots] == 0:
                    break                   # the output slot is still in flight
                if submit(sv, i, i):
                    raise RuntimeError("fy_serve_submit")
                i += 1
            m = lib.fy_serve_poll(sv, buf, 4096)
            if m < 0:
                raise RuntimeError(lib.fy_gpu_last_error().decode())
            t = time.perf_counter() - t0
            for j in range(m):
                done_t[buf[j]] = t
            nd += m
        lib.fy_serve_destroy(sv)
        lat = np.sort(done_t - arrive) * 1000.0
        span = done_t.max() - arrive[0]
        return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3), "mvox_s": round(nvox / span / 1e6, 1),
                "wall_s": round(float(span), 2)}

    res = {"workload": f"{n_req} requests, Poisson {rate:g} req/s, sizes 40% 64^3 / 30% 128^3 / 20% 192^3 / "
                       "10% 256^3, per-request physics; " +
                       ("inputs and outputs in pinned host memory (latency includes H2D and D2H)" if host
                        else "inputs and outputs resident in HBM"),
           "requests": n_req, "rate_req_s": rate, "host_resident": host,
           "offered_mvox_s": round(nvox / (arrive[-1] - arrive[0]) / 1e6, 1)}
    for name, (mb, lanes) in (modes or {"served": (16, 3), "plain": (1, 1)}).items():
        res[name] = run(mb, lanes)
    if "served" in res and "plain" in res:
        res["served_over_plain"] = round(res["served"]["mvox_s"] / res["plain"]["mvox_s"], 3)
    return res


def precision_bench(F, vol, reps=5):
    """The 512^3 precision configuration (the workload cube tiled 2x2x2): the
    fp32 chain (the product path) and its all-FP64 variant (deconvolution
    spectrum, weights and the fast guided filter in double), both u8 in HBM ->
    u8 in HBM, timed; and, outside the timed region, the fp32 results against
    the all-FP64 ones: the deconvolved tile (fy_gpu_deconv_radial vs
    fy_gpu_deconv_radial_f64 on the dewindowed cube), the guided tile and the
    u8 output of the chain.  The FP64 path itself agrees with the fp64 CPU
    oracle to ~1e-12 (deconvolution) and 3.5e-8 (its fp32-emitted guided tile),
    which tests/test_gpu_precision.py checks at this size."""
    import torch
    lib = F.lib()
    big = np.tile(vol, (2, 2, 2))
    n = big.shape[0]
    d_in = torch.from_numpy(big).cuda()
    del big
    d_out = torch.empty_like(d_in)
    d_out64 = torch.empty_like(d_in)
    phys = F.physics(delta_beta=300.0 * 0.35, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                     unsharp_sigma=0.8, unsharp_coeff=0.5)
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    out = {"workload": f"{n}^3 u8 cube (the workload cube tiled 2x2x2): dewindow -> Wiener deconv (sigma ratio "
                       "1.17) -> guided r=2 s=2 -> u8"}
    prm = {}
    for name, fp64 in (("fp32", 0), ("fp64", 1)):
        P = F.CubeParams()
        P.phys = phys
        P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, 1.17 ** 2
        P.window_lo, P.window_hi = -0.1, 1.2
        P.guided_eps, P.guided_radius, P.guided_s = 2e-3, 2, 2
        P.fp64 = fp64
        prm[name] = P
        z = (0, 0, 0)
        dst = d_out64 if fp64 else d_out
        cube = F.Cube(d_in.data_ptr(), *z, n, n, n, *z, n, n, n, dst.data_ptr(), None, C.pointer(P))
        if lib.fy_gpu_process_cubes(C.byref(cube), 1, sh):
            raise RuntimeError(lib.fy_gpu_last_error().decode())
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            lib.fy_gpu_process_cubes(C.byref(cube), 1, sh)
        stream.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"ms_per_cube": round(1000 * dt, 3), "mvox_s": round(n ** 3 / dt / 1e6, 1)}

    def rel_max(a, b):
        a = a.double()
        b = b.double()
        return (float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)),
                float((a - b).abs().max()))

    # errors, untimed: deconvolution alone (the chain's dewindowed input, fp32 vs fp64 spectrum)
    f = d_in.float() * np.float32(1.0 / 255.0)
    f = f * np.float32(1.2 - (-0.1)) + np.float32(-0.1)
    d32 = torch.empty_like(f)
    d64 = torch.empty(f.shape, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = lib.fy_gpu_deconv_radial(C.c_void_p(f.data_ptr()), C.c_void_p(d32.data_ptr()), n, n, n, C.byref(phys),
                                  0.015, 0, 8.0, 1.17 ** 2, sh)
    rc |= lib.fy_gpu_deconv_radial_f64(C.c_void_p(f.data_ptr()), C.c_void_p(d64.data_ptr()), n, n, n,
                                       C.byref(phys), 0.015, 0, 8.0, 1.17 ** 2, sh)
    stream.synchronize()
    if rc:
        raise RuntimeError(lib.fy_gpu_last_error().decode())
    e_dec = rel_max(d32, d64)
    del f, d32, d64
    # the guided tile of the chain (fp32 tile out) and the u8 output
    gt = {}
    for name in ("fp32", "fp64"):
        t = torch.empty(d_in.shape, dtype=torch.float32, device="cuda")
        z = (0, 0, 0)
        cube = F.Cube(d_in.data_ptr(), *z, n, n, n, *z, n, n, n, d_out.data_ptr() if name == "fp32" else
                      d_out64.data_ptr(), t.data_ptr(), C.pointer(prm[name]))
        if lib.fy_gpu_process_cubes(C.byref(cube), 1, sh):
            raise RuntimeError(lib.fy_gpu_last_error().decode())
        gt[name] = t
    stream.synchronize()
    e_gui = rel_max(gt["fp32"], gt["fp64"])
    du = (d_out.int() - d_out64.int()).abs()
    out["errors_vs_fp64_gpu"] = {
        "deconv": {"rel_l2": e_dec[0], "max_abs": e_dec[1]},
        "guided": {"rel_l2": e_gui[0], "max_abs": e_gui[1]},
        "u8": {"max_abs": int(du.max()), "frac_differing": float((du > 0).float().mean())},
        "reference": "the all-FP64 GPU path (within ~1e-12 / 3.5e-8 of the fp64 CPU oracle, "
                     "tests/test_gpu_precision.py)"}
    out["fp32_over_fp64_speed"] = round(out["fp32"]["mvox_s"] / out["fp64"]["mvox_s"], 3)
    return out


def calibration_bench(F, vol, reps=50):
    """Pass-1 calibration statistics of the workload cube, all on the device:
    the u8 histogram plus the per-slice z-drift sums of the 4x4 subsample
    (fy_calib_accumulate_u8, one 256^3 ptile per call), and the flat-block noise
    floor of the central 128^3 sub-block as f = u8/255 (fy_gpu_flat_noise,
    including its read-back and the host selection)."""
    import torch
    lib = F.lib()
    n = vol.shape[0]
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    d = torch.from_numpy(vol).cuda()
    torch.cuda.synchronize()
    cal = lib.fy_calib_create(n, sh)
    if not cal:
        raise RuntimeError(lib.fy_gpu_last_error().decode())
    flags = 3                                   # FY_CALIB_HIST | FY_CALIB_ZDRIFT
    for _ in range(3):
        lib.fy_calib_accumulate_u8(cal, C.c_void_p(d.data_ptr()), 0, n, n, n, flags)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        if lib.fy_calib_accumulate_u8(cal, C.c_void_p(d.data_ptr()), 0, n, n, n, flags):
            raise RuntimeError(lib.fy_gpu_last_error().decode())
    stream.synchronize()
    dt = (time.perf_counter() - t0) / reps
    lib.fy_gpu_profile_reset()
    lib.fy_gpu_profile(1)
    for _ in range(20):
        lib.fy_calib_accumulate_u8(cal, C.c_void_p(d.data_ptr()), 0, n, n, n, flags)
    stream.synchronize()
    lib.fy_gpu_profile(0)
    prof = F.profile_read()
    lib.fy_calib_destroy(cal)
    kern = {k: round(1000.0 * ms / max(c, 1), 2) for k, (ms, c) in prof.items()}
    nbytes = vol.size                           # the tile is read once: N bytes
    hist_us = kern.get("hist_u8")
    h = n // 2 - 64
    sub = (d[h:h + 128, h:h + 128, h:h + 128].float() * np.float32(1.0 / 255.0)).contiguous()
    torch.cuda.synchronize()
    nf = C.c_double()
    lib.fy_gpu_flat_noise(C.c_void_p(sub.data_ptr()), 128, 128, 128, 8, C.byref(nf), sh)
    t1 = time.perf_counter()
    for _ in range(10):
        lib.fy_gpu_flat_noise(C.c_void_p(sub.data_ptr()), 128, 128, 128, 8, C.byref(nf), sh)
    fdt = (time.perf_counter() - t1) / 10
    return {"workload": f"pass-1 statistics of the {n}^3 u8 cube: histogram + 4x4 z-drift sums per call",
            "ms_per_tile": round(1000 * dt, 4), "gvox_s": round(vol.size / dt / 1e9, 1),
            "kernels_us": kern,
            "hist_roofline": {"bound": "hbm", "algorithmic_bytes": nbytes,
                              "achieved": round(nbytes / (hist_us * 1e-6) / 1e9, 1) if hist_us else None,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(nbytes / (hist_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if hist_us else None},
            "flat_noise_128": {"ms_per_call": round(1000 * fdt, 3), "noise_floor": nf.value}}


def air_zero_bench(F, vol, eps, reps=10):
    """The workload cube through the chain with stage (d) added: air-zero value
    cut (valley rule) and despeckle of islands under 500 voxels, u8 in HBM -> u8
    in HBM, with its per-kernel times."""
    import torch
    lib = F.lib()
    n = vol.shape[0]
    P = F.CubeParams()
    P.phys = F.physics(delta_beta=300.0 * 0.35, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                       unsharp_sigma=0.8, unsharp_coeff=0.5)
    P.deconv, P.deconv_param, P.deconv_cap, P.fz2_scale = 1, 0.015, 8.0, 1.17 ** 2
    P.window_lo, P.window_hi = -0.1, 1.2
    P.guided_eps, P.guided_radius, P.guided_s = eps, 2, 2
    P.air_zero, P.air_cut_u8, P.air_min_component = 1, -1, 500
    d_in = torch.from_numpy(vol).cuda()
    d_out = torch.empty_like(d_in)
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    z = (0, 0, 0)
    cube = F.Cube(d_in.data_ptr(), *z, n, n, n, *z, n, n, n, d_out.data_ptr(), None, C.pointer(P))
    torch.cuda.synchronize()
    if lib.fy_gpu_process_cubes(C.byref(cube), 1, sh):
        raise RuntimeError(lib.fy_gpu_last_error().decode())
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.fy_gpu_process_cubes(C.byref(cube), 1, sh)
    stream.synchronize()
    dt = (time.perf_counter() - t0) / reps
    lib.fy_gpu_profile_reset()
    lib.fy_gpu_profile(1)
    for _ in range(5):
        lib.fy_gpu_process_cubes(C.byref(cube), 1, sh)
    stream.synchronize()
    lib.fy_gpu_profile(0)
    prof = F.profile_read()
    stage = {k: round(1000.0 * ms / max(c, 1), 2) for k, (ms, c) in prof.items()
             if k.startswith("air_") or k.startswith("cc_") or k in ("prologue", "emit", "guided_coef",
                                                                     "guided_upsample")}
    zeroed = float((d_out == 0).float().mean().item())
    # stage (d) and the emit: kernel time per cube against the 6N bytes the
    # stage needs at least (float tile in, u8 input in, u8 tile out)
    d_ms = sum(ms for k, (ms, c) in prof.items()
               if k.startswith("air_") or k.startswith("cc_") or k in ("prologue", "emit")) / 5
    d_bytes = 6 * n ** 3
    roof = {"bound": "hbm", "kernels_us": round(1000 * d_ms, 1), "algorithmic_bytes": d_bytes,
            "achieved": round(d_bytes / (d_ms * 1e-3) / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(d_bytes / (d_ms * 1e-3) / 8e12, 4)}
    return {"workload": f"{n}^3 u8 cube, the chain plus air-zero (valley rule) and despeckle (< 500 voxels)",
            "ms_per_cube": round(1000 * dt, 3), "mvox_s": round(n ** 3 / dt / 1e6, 1),
            "zeroed_fraction": round(zeroed, 4), "kernels_us": stage, "stage_d_roofline": roof}


def lod_phasecorr_bench(F, vol, reps=20):
    """The adjacent consumers: the 2x LOD downscale of the workload cube
    (contrast-maintaining box, device to device; algorithmic bytes N + N/8) and
    phase correlation of its central 128^3 block against a copy shifted by
    (3, -2, 5) voxels (windowed; two forward and one inverse 128^3 FFT)."""
    import torch
    lib = F.lib()
    n = vol.shape[0]
    d = torch.from_numpy(vol).cuda()
    o = torch.empty((n + 1) // 2, (n + 1) // 2, (n + 1) // 2, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    torch.cuda.synchronize()
    for _ in range(3):
        lib.fy_gpu_downscale2x(C.c_void_p(d.data_ptr()), n, n, n, C.c_void_p(o.data_ptr()), 1, C.c_float(0.5), sh)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.fy_gpu_downscale2x(C.c_void_p(d.data_ptr()), n, n, n, C.c_void_p(o.data_ptr()), 1, C.c_float(0.5), sh)
    stream.synchronize()
    dt = (time.perf_counter() - t0) / reps
    nbytes = vol.size + o.numel()
    h = n // 2 - 64
    a = (d[h:h + 128, h:h + 128, h:h + 128].float() / 255.0).contiguous()
    b = torch.roll(a, shifts=(3, -2, 5), dims=(0, 1, 2)).contiguous()
    torch.cuda.synchronize()
    s3 = np.zeros(3)
    pk = C.c_double()
    lib.fy_gpu_phase_correlate(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), 128, 128, 128,
                               s3.ctypes.data_as(F.c_double_p), C.byref(pk), 1, sh)
    t1 = time.perf_counter()
    for _ in range(5):
        lib.fy_gpu_phase_correlate(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), 128, 128, 128,
                                   s3.ctypes.data_as(F.c_double_p), C.byref(pk), 1, sh)
    pdt = (time.perf_counter() - t1) / 5
    return {"downscale2x": {"workload": f"{n}^3 u8 -> {(n + 1) // 2}^3, contrast-maintaining box",
                            "ms": round(1000 * dt, 4), "GBps": round(nbytes / dt / 1e9, 1),
                            "frac_of_hbm_peak": round(nbytes / dt / 1e9 / HBM_PEAK_GBS, 4)},
            "phase_correlate": {"workload": "128^3 patch vs itself rolled by (3, -2, 5), Hann window",
                                "ms_per_call": round(1000 * pdt, 3), "shift": [round(v, 4) for v in s3],
                                "peak": round(pk.value, 4)}}


def export_bench(F, vol, eps, tile=128, threads=16, batch=8):
    """Pass-2 export through the GPU (fy_gpu_export): a 512^3 u8 volume (the
    workload cube tiled 2x2x2) stored as zarr v2 with 128^3 chunks in a
    temporary directory, restored tile by tile (tile 128, halo 8, the chain of
    the headline workload without rings) into an output zarr.  Inputs are read
    from files the page cache holds and outputs written to files, so the rate
    includes host I/O and both PCIe directions; it is set beside the GPU-only
    rate of the same tiles with their regions already in HBM."""
    import shutil
    import tempfile
    import torch
    lib = F.lib()
    big = np.tile(vol, (2, 2, 2))
    Z = big.shape[0]
    work = tempfile.mkdtemp(prefix="fy_export_")
    try:
        zi = F.Zarr()
        src = os.path.join(work, "in.zarr").encode()
        shp = (C.c_long * 3)(Z, Z, Z)
        ch = (C.c_long * 3)(tile, tile, tile)
        assert lib.fy_gpu_zarr_create(C.byref(zi), src, shp, ch) == 0
        nt = Z // tile
        for cz in range(nt):
            for cy in range(nt):
                for cx in range(nt):
                    blk = np.ascontiguousarray(big[cz * tile:(cz + 1) * tile, cy * tile:(cy + 1) * tile,
                                                   cx * tile:(cx + 1) * tile])
                    assert lib.fy_gpu_zarr_write_chunk(C.byref(zi), cz, cy, cx, blk.ctypes.data, tile, tile, tile) == 0
        cfg = F.pipeline_cfg(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3,
                             unsharp_sigma=0.8, unsharp_coeff=0.5, window_lo=-0.1, window_hi=1.2,
                             do_deconv=1, guided_eps=eps, halo=8)
        stats = F.ExportStats()

        def export(name, t1=0):
            zo = F.Zarr()
            assert lib.fy_gpu_zarr_create(C.byref(zo), os.path.join(work, name).encode(), shp, ch) == 0
            t = time.perf_counter()
            rc = lib.fy_gpu_export(C.byref(zi), C.byref(zo), C.byref(cfg), tile, 0, t1, threads, batch,
                                   C.byref(stats), None)
            if rc:
                raise RuntimeError("fy_gpu_export: " + lib.fy_gpu_last_error().decode())
            return time.perf_counter() - t

        export("warm.zarr", t1=2 * batch)
        walls = [export(f"out{i}.zarr") for i in range(3)]
        wall = statistics.median(walls)
        # the same tiles with their regions resident in HBM: one fy_gpu_process_cubes per batch
        h = 8
        reg = tile + 2 * h
        d_in = torch.from_numpy(np.ascontiguousarray(big[:reg, :reg, :reg])).cuda()
        d_out = torch.empty((batch, tile, tile, tile), dtype=torch.uint8, device="cuda")
        prm = F.CubeParams()
        assert lib.fy_cube_params_from_cfg(C.byref(cfg), h, h, tile, tile, Z, Z, C.byref(prm)) == 0
        cubes = (F.Cube * batch)(*[F.Cube(d_in.data_ptr(), 0, 0, 0, reg, reg, reg, h, h, h, tile, tile, tile,
                                          d_out[i].data_ptr(), None, C.pointer(prm)) for i in range(batch)])
        torch.cuda.synchronize()
        for _ in range(2):
            assert lib.fy_gpu_process_cubes(cubes, batch, None) == 0
        assert lib.fy_gpu_sync(None) == 0
        t = time.perf_counter()
        reps = 8
        for _ in range(reps):
            assert lib.fy_gpu_process_cubes(cubes, batch, None) == 0
        assert lib.fy_gpu_sync(None) == 0
        gpu_tile_s = (time.perf_counter() - t) / (reps * batch)
        ntiles = nt ** 3
        return {"workload": f"{Z}^3 u8 zarr (128^3 chunks) -> restored zarr, tile {tile} halo {h}, "
                            f"{threads} I/O threads, batches of {batch}; inputs in the page cache",
                "tiles": ntiles, "wall_s": round(wall, 4), "mvox_s": round(Z ** 3 / wall / 1e6, 1),
                "read_thread_s": round(stats.read_s, 3), "write_thread_s": round(stats.write_s, 3),
                "gpu_only_mvox_s": round(tile ** 3 / gpu_tile_s / 1e6, 1),
                "gpu_only_wall_s": round(gpu_tile_s * ntiles, 4)}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def copy_calibration(nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size: a profiler's FETCH_SIZE and
    WRITE_SIZE for this kernel, divided by nbytes, is the factor that turns the
    same counters of the chain's kernels into HBM bytes."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def _host_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def _ref_library():
    """The reference compiled from its sources (oracle/_ref): the AVX-512 build
    where the host CPU has AVX-512 F/BW/CD/DQ/VL (what the reference's own
    -march=native would use there), else the AVX2+FMA build.  Returns
    (CDLL, march) or (None, reason)."""
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    cand = [("libfysics_ref_v4.so", "x86-64-v4 (AVX-512)")] if v4 <= _host_flags() else []
    cand.append(("libfysics_ref.so", "x86-64-v3 (AVX2+FMA)"))
    for so, march in cand:
        path = os.path.join(ROOT, "oracle", "_ref", so)
        if os.path.exists(path):
            return C.CDLL(path), march
    return None, "oracle/_ref not built on this host"


def cpu_baseline(vol, phys, eps, ring, threads):
    """The reference's own chain (compiled from its sources, oracle/_ref, with its
    -Ofast flags) on the host cores: one 256^3 cube per OpenMP thread, one warm-up
    round, median of 3."""
    ref, march = _ref_library()
    if ref is None:
        return {"value": None, "note": march}
    import fysics_amd as F
    ref.ref_chain_many.restype = C.c_double
    ref.ref_chain_many.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_long, C.c_long,
                                   C.c_long, C.c_long, C.POINTER(F.Physics), C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.POINTER(F.Dering), C.c_int]
    if threads <= 0:
        env = os.environ.get("OMP_NUM_THREADS")
        threads = int(env) if env else len(os.sched_getaffinity(0))
    n = vol.shape[0]
    ins = [vol for _ in range(threads)]
    outs = [np.empty_like(vol) for _ in range(threads)]
    pin = (C.c_void_p * threads)(*[a.ctypes.data for a in ins])
    pout = (C.c_void_p * threads)(*[a.ctypes.data for a in outs])
    cpu_name = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_name = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    times = []
    for _ in range(4):
        t = ref.ref_chain_many(threads, pin, pout, n, n, n, 0, C.byref(phys), 1, 0.015, 8.0, 1.17 ** 2,
                               -0.1, 1.2, eps, 2, C.byref(ring), threads)
        if t < 0:
            return {"value": None, "note": "reference chain failed"}
        times.append(t)
    med = statistics.median(times[1:])
    return {"value": round(threads * n ** 3 / 1e6 / med, 1), "unit": "Mvoxels/s", "cores": threads,
            "kind": "reference", "march": march,
            "sample": f"{threads} x {n}^3 cubes per round (one per OpenMP thread), warm-up + median of 3 "
                      f"rounds, {med:.2f} s/round; CPU: {cpu_name}; built -Ofast -ffast-math -funroll-loops "
                      f"-march={march.split()[0]} -fopenmp"}


def config1_bench(F):
    """The plain CPU reference path on one 128^3 volume (the config-1 phantom:
    40 sheets, Paganin+unsharp forward model, noise var = 0.02 I + 4, u8 through
    [-0.1, 1.2]) restored in the order a reference user calls it: fy_u8_to_phys
    -> fy_deconvolve(reg 0.015) -> fy_estimate_noise(win 5, 10th percentile,
    ref 0.5) -> fy_guided_eps_for_noise -> fy_guided_denoise(r 2), host buffers
    in and out.  Timed for the reference compiled from its sources (its own
    OpenMP where it has any) and for this library's drop-in entry points of the
    same names (PCIe included); median of 3 after a warm-up."""
    from fysics_amd import phantom
    ref, march = _ref_library()
    u8 = phantom.config1_volume(128, seed=1)
    shape, n = u8.shape, u8.size
    lo, hi = phantom.WINDOW
    ph = F.physics(**phantom.CONFIG1_PHYS)
    libs = [("gpu", F.lib())]
    if ref is not None:
        for name, (res, args) in F.SIGNATURES.items():
            if name in ("fy_u8_to_phys", "fy_deconvolve", "fy_estimate_noise", "fy_guided_eps_for_noise",
                        "fy_guided_denoise") and hasattr(ref, name):
                fn = getattr(ref, name)
                fn.restype, fn.argtypes = res, args
        libs.append(("reference_cpu", ref))

    def chain(lib):
        f = np.empty(shape, np.float32)
        d = np.empty(shape, np.float32)
        g = np.empty(shape, np.float32)
        t0 = time.perf_counter()
        lib.fy_u8_to_phys(u8.ctypes.data_as(F.c_u8_p), F._fp(f), n, lo, hi)
        if lib.fy_deconvolve(F._fp(f), F._fp(d), *shape, C.byref(ph), 0.015):
            raise RuntimeError("fy_deconvolve")
        nm = F.NoiseModel()
        if lib.fy_estimate_noise(F._fp(d), *shape, 5, 10.0, 0.5, C.byref(nm)):
            raise RuntimeError("fy_estimate_noise")
        eps = lib.fy_guided_eps_for_noise(nm.noise_ref)
        if lib.fy_guided_denoise(F._fp(d), F._fp(g), *shape, 2, eps):
            raise RuntimeError("fy_guided_denoise")
        return time.perf_counter() - t0, g, eps

    out = {"workload": "128^3 config-1 phantom: dewindow -> fy_deconvolve(0.015) -> fy_estimate_noise -> "
                       "guided r=2 exact, host buffers"}
    res = {}
    for name, lib in libs:
        chain(lib)
        runs = [chain(lib) for _ in range(3)]
        t = statistics.median(r[0] for r in runs)
        res[name] = runs[0][1]
        out[name] = {"ms": round(1000 * t, 2), "mvox_s": round(n / t / 1e6, 1), "eps": runs[0][2]}
    if "reference_cpu" in out:
        out["reference_cpu"]["march"] = march
        a, b = res["gpu"].astype(np.float64), res["reference_cpu"].astype(np.float64)
        out["gpu_vs_reference_rel_l2"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out["speedup"] = round(out["reference_cpu"]["ms"] / out["gpu"]["ms"], 1)
    return out


if __name__ == "__main__":
    main()
