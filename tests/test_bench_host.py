"""bench.py's measurement contract on the host: algorithmic bytes per launch,
the dominant-kernel roofline entry, and the one-GPU limit.  No GPU needed."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

H256 = 256 * 256 * 129          # half-spectrum elements of a 256^3 cube


def test_kernel_bytes_256():
    # each pass moves the complex fp32 half spectrum in and out once
    assert bench.kernel_bytes("cfft_y_fwd", 256, 256) == 16 * H256
    assert bench.kernel_bytes("cfft_y_inv", 256, 256) == 16 * H256
    assert bench.kernel_bytes("zfused", 256, 256) == 16 * H256
    # u8 in, spectrum out / spectrum in, fp32 tile out
    assert bench.kernel_bytes("rfft_x_u8", 256, 256) == 256 ** 3 + 8 * H256
    assert bench.kernel_bytes("irfft_x", 256, 256) == 8 * H256 + 4 * 256 ** 3
    # guided filter at s = 2: 128^3 low-res grid
    assert bench.kernel_bytes("guided_coef", 256, 256) == 12 * 128 ** 3
    assert bench.kernel_bytes("guided_upsample", 256, 256) == 5 * 256 ** 3 + 8 * 128 ** 3
    assert bench.kernel_bytes("ring_image", 256, 256) is None


def test_roofline_counts_both_y_sites():
    # zfused has the longest single launch, but the Y pass runs twice per cube
    prof = {"cfft_y_fwd": (2.5, 100), "cfft_y_inv": (2.5, 100), "zfused": (3.2, 100),
            "guided_upsample": (3.1, 100), "ring_image": (0.6, 100)}
    r = bench.roofline_of(prof, 256, 256)
    assert r["kernel"] == "cfft_y"
    assert sorted(r["launch_sites"]) == ["cfft_y_fwd", "cfft_y_inv"]
    assert r["avg_us"] == pytest.approx(25.0)
    assert r["algorithmic_bytes"] == 16 * H256
    assert r["achieved"] == pytest.approx(16 * H256 / 25e-6 / 1e9, rel=1e-4)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, rel=1e-3)
    assert r["traffic"] == bench.MEASURED_TRAFFIC_256["cfft_y_fwd"]


def test_roofline_single_site_dominant():
    prof = {"cfft_y_fwd": (1.0, 100), "cfft_y_inv": (1.0, 100), "zfused": (3.2, 100)}
    r = bench.roofline_of(prof, 256, 256)
    assert r["kernel"] == "zfused" and r["launch_sites"] == ["zfused"]
    assert r["avg_us"] == pytest.approx(32.0)
    assert r["traffic"] == bench.MEASURED_TRAFFIC_256["zfused"]
    # no measured traffic for another cube size
    assert bench.roofline_of(prof, 128, 128)["traffic"] is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
