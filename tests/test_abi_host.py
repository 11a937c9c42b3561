"""CPU-only checks of the product library: it loads, exports every function the
headers declare with the declared struct layouts, and its host-side logic
(sizes, transfer formulas, ring-vote finalize, cfg -> cube parameters) matches
the oracle.  No device work is issued here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import clibs
import fysics_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("fysics.h", "fysics_gpu.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"typedef struct \{.*?\}\s*\w+;", "", src, flags=re.S)
        for m in re.finditer(r"\b(fy_\w+)\s*\(", src):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = F.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds every one of them
    unbound = [n for n in names if n not in F.SIGNATURES]
    assert not unbound, unbound


def _declared_in(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"typedef struct \{.*?\}\s*\w+;", "", src, flags=re.S)
    return {m.group(1) for m in re.finditer(r"\b(fy_\w+)\s*\(", src)}


def _interposed_list():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = doc.split("<!-- interposed-symbols -->", 1)[1].split("<!-- /interposed-symbols -->", 1)[0]
    return {w for w in body.split() if w}


def test_exports_are_the_declared_set():
    """The shared library exports the functions the headers declare and nothing
    else; the ones with reference names (include/fysics.h) are exactly the list
    INTEGRATION.md documents, and, where the reference header is present, every
    exported name the reference also declares is on that list -- so no
    narrowed variant takes a reference symbol's place by accident."""
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-2] in ("T", "W")}
    assert exported == set(declared_functions())
    interposed = _declared_in(HEADERS[0])
    assert _interposed_list() == interposed
    assert not (interposed & _declared_in(HEADERS[1]))
    ref_h = "/root/reference/fysics.h"
    if os.path.exists(ref_h):
        assert exported & _declared_in(ref_h) == interposed


def test_no_oracle_or_reference_symbols_in_product():
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in out and "ref_chain" not in out
    deps = subprocess.run(["ldd", F.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in deps and "libfysics_ref" not in deps


def _c_sizeof(types, header_dir):
    prog = "#include <stdio.h>\n#include \"fysics.h\"\nint main(void){" + "".join(
        f'printf("%zu\\n", sizeof({t}));' for t in types) + "return 0;}"
    exe = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"fy_sizeof_{os.getpid()}")
    r = subprocess.run(["gcc", "-x", "c", "-I", header_dir, "-o", exe, "-"], input=prog, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr
    vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    os.unlink(exe)
    return vals


STRUCTS = [("fy_physics", F.Physics), ("fy_noise_model", F.NoiseModel), ("fy_dering", F.Dering),
           ("fy_pipeline_cfg", F.PipelineCfg), ("fy_zarr", F.Zarr)]


def test_struct_layouts_match_ctypes():
    sizes = _c_sizeof([s for s, _ in STRUCTS], os.path.join(ROOT, "include"))
    assert sizes == [C.sizeof(t) for _, t in STRUCTS]


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference header absent")
def test_struct_layouts_match_reference_header():
    ours = _c_sizeof([s for s, _ in STRUCTS], os.path.join(ROOT, "include"))
    theirs = _c_sizeof([s for s, _ in STRUCTS], "/root/reference")
    assert ours == theirs


def test_sizes_and_transfers_match_oracle(fy, orc):
    for n in list(range(1, 600)) + [1000, 2049, 4096]:
        assert fy.fy_next_fft_size(n) == orc.orc_next_fft_size(n)
    assert fy.fy_is_pow2(64) and not fy.fy_is_pow2(96) and fy.fy_next_pow2(97) == 128
    for kw in (dict(), dict(delta_beta=300, energy_kev=110, distance_mm=1200, pixel_um=4.3,
                            unsharp_sigma=0.8, unsharp_coeff=0.5),
               dict(pixel_um=0.0024), dict(unsharp_coeff=0.0)):
        p, q = F.physics(**kw), clibs.physics(**kw)
        for f in np.linspace(0, 0.5, 23):
            assert fy.fy_paganin_transfer(f, C.byref(p)) == orc.orc_paganin_transfer(f, C.byref(q))
            assert fy.fy_unsharp_transfer(f, C.byref(p)) == orc.orc_unsharp_transfer(f, C.byref(q))
            assert fy.fy_recon_transfer(f, C.byref(p)) == orc.orc_recon_transfer(f, C.byref(q))
        assert fy.fy_auto_deltabeta_scale(C.byref(p)) == orc.orc_auto_deltabeta_scale(C.byref(q))
    assert fy.fy_auto_reg(C.byref(F.physics()), -1.0) == 0.015
    assert fy.fy_guided_eps_for_noise(0.01) == orc.orc_guided_eps_for_noise(0.01)
    assert fy.fy_guided_eps_for_noise(float("nan")) == 1e-8
    assert fy.fy_phys_to_u8_level(0.0, -0.03, 0.21) == orc.orc_phys_to_u8_level(0.0, -0.03, 0.21)
    assert fy.fy_phys_to_u8_level(0.3, 0.1, 0.1) == 0.0


def test_dering_finalize_host_logic_matches_oracle(fy, orc):
    """Sector vote, median and clamps over the golden (sector, radius) sums."""
    gold = np.load(os.path.join(ROOT, "tests", "golden", "reference_vectors.npz"))
    nz, ny, nx = gold["dering_vol"].shape
    for args in ((9, 20, 1.0, 6.0), (15, 50, 2.0, 3.0), (3, 1, 0.2, 1.0)):
        d = F.Dering()
        o = clibs.Dering()
        assert fy.fy_dering_init(C.byref(d), nz, ny, nx, 47.5, 55.0, 4, 8) == 0
        assert orc.orc_dering_init(C.byref(o), nz, ny, nx, 47.5, 55.0, 4, 8) == 0
        nb = d.nslab * d.ns * d.nr
        for st in (d, o):
            np.ctypeslib.as_array(st.sum, (nb,))[:] = gold["dering_sum"]
            np.ctypeslib.as_array(st.cnt, (nb,))[:] = gold["dering_cnt"]
        n1 = fy.fy_dering_finalize(C.byref(d), *args)
        n2 = orc.orc_dering_finalize(C.byref(o), *args)
        assert n1 == n2 and d.have_rings == o.have_rings
        assert np.array_equal(np.ctypeslib.as_array(d.ring, (d.nslab * d.nr,)),
                              np.ctypeslib.as_array(o.ring, (o.nslab * o.nr,)))
        fy.fy_dering_free(C.byref(d))
        orc.orc_dering_free(C.byref(o))


def test_dering_tile_merge_host(fy):
    d = F.Dering()
    t = F.Dering()
    assert fy.fy_dering_init(C.byref(d), 1000, 300, 200, -1, -1, 256, 6) == 0
    assert d.nslab == 4 and d.ns == 6 and d.cy == 149.5 and d.cx == 99.5
    assert d.nr == int(np.sqrt(149.5 ** 2 + 99.5 ** 2)) + 2
    assert fy.fy_dering_tile_init(C.byref(t), C.byref(d)) == 0 and t.nslab == 1
    nb = t.ns * t.nr
    np.ctypeslib.as_array(t.sum, (nb,))[:] = 3.0
    np.ctypeslib.as_array(t.cnt, (nb,))[:] = 2
    fy.fy_dering_merge_tile(C.byref(d), 2, C.byref(t))
    fy.fy_dering_merge_tile(C.byref(d), 7, C.byref(t))        # out of range: ignored
    s = np.ctypeslib.as_array(d.sum, (d.nslab * nb,)).reshape(d.nslab, nb)
    assert (s[2] == 3.0).all() and s[[0, 1, 3]].sum() == 0
    fy.fy_dering_tile_reset(C.byref(t))
    assert not np.ctypeslib.as_array(t.cnt, (nb,)).any()
    fy.fy_dering_free(C.byref(t))
    fy.fy_dering_free(C.byref(d))


def test_tuning_file_round_trip(tmp_path):
    """Tile shapes of the guided filter persist in a text file: loaded shapes
    are reported as such (and used without timing); bad lines are rejected,
    good ones kept; save writes back what was loaded."""
    lib = F.lib()
    lib.fy_gpu_tuning_clear()
    src = tmp_path / "tune.txt"
    src.write_text("# comment\n"
                   "fp32 r1 s2 grid 8x8x8 region 16x16x16 u8\t2\t1\n"
                   "fp32 r2 s2 grid 9x9x9 region 18x18x18 fp32\t0\t-1\n"
                   "broken line without tabs\n"
                   "fp32 r1 s2 grid 4x4x4 region 8x8x8 u8\t7\t0\n")
    assert lib.fy_gpu_tuning_load(str(src).encode()) == 1          # two bad lines
    rep = F.tuning_report()
    assert "grid 8x8x8 region 16x16x16 u8: coef zchunk #2, upsample planes #1 (from a tuning file)" in rep
    assert "grid 9x9x9 region 18x18x18 fp32: coef zchunk #0, upsample planes #0 (untuned) (from a tuning file)" \
        in rep
    assert "grid 4x4x4" not in rep
    out = tmp_path / "saved.txt"
    assert lib.fy_gpu_tuning_save(str(out).encode()) == 0
    lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert sorted(lines) == sorted(["fp32 r1 s2 grid 8x8x8 region 16x16x16 u8\t2\t1",
                                    "fp32 r2 s2 grid 9x9x9 region 18x18x18 fp32\t0\t-1"])
    assert lib.fy_gpu_tuning_load(str(tmp_path / "missing.txt").encode()) == 1
    lib.fy_gpu_tuning_clear()
    assert F.tuning_report() == ""


def test_pass1_commits_match_oracle(fy, orc):
    """fy_calib_commit_eps / fy_aniso_ratio / fy_calib_commit_aniso are host
    fp64: bit-identical to the oracle on random samples and on the gate cases."""
    rng = np.random.default_rng(8)
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    for trial in range(40):
        n = int(rng.integers(0, 60))
        r = rng.uniform(0, 1500, n)
        f = np.abs(rng.uniform(0.005, 0.02, n) + rng.uniform(-1e-5, 4e-5) * r)
        kw = dict(guided_eps=float(rng.choice([0.0, 1e-3])), denoise_k=float(rng.choice([-1.0, 0.0, 3.5])))
        a, b = F.PipelineCfg(), clibs.PipelineCfg()
        for k, v in kw.items():
            setattr(a, k, v)
            setattr(b, k, v)
        fy.fy_calib_commit_eps(C.byref(a), dp(r), dp(f), n)
        orc.orc_calib_commit_eps(C.byref(b), dp(r), dp(f), n)
        for k in ("have_eps_r", "eps_fn_a", "eps_fn_b", "eps_fn_med", "noise_floor", "guided_eps"):
            assert getattr(a, k) == getattr(b, k), (trial, k)
        ratios = rng.uniform(0.9, 1.8, int(rng.integers(0, 9)))
        fy.fy_calib_commit_aniso(C.byref(a), dp(ratios), len(ratios))
        orc.orc_calib_commit_aniso(C.byref(b), dp(ratios), len(ratios))
        assert a.psf_z_ratio == b.psf_z_ratio
        rz, rxy = rng.uniform(0, 1, 2)
        assert fy.fy_aniso_ratio(rz, rxy) == orc.orc_aniso_ratio(rz, rxy)


LINK_PROG = r"""
#include <stdio.h>
#include "fysics.h"
#include "fysics_gpu.h"
int main(void) {
    fy_physics p = {300.0, 110.0, 1200.0, 4.3, 0.8, 0.5, 0.0, 0.0};
    printf("%d %d %d %.17g %.17g\n", fy_next_fft_size(160), fy_next_fft_size(96), fy_is_pow2(64),
           fy_paganin_transfer(0.25, &p), fy_guided_eps_for_noise(0.01));
    return 0;
}
"""


@pytest.mark.parametrize("kind", ["shared", "static"])
def test_c_program_links_unchanged(tmp_path, kind):
    """A C caller of the reference API compiles against include/ and links the
    shared library or the static archive (the reference ships both,
    CMakeLists.txt:55-74) with no source change."""
    src = tmp_path / "prog.c"
    src.write_text(LINK_PROG)
    exe = tmp_path / "prog"
    libdir = os.path.join(ROOT, "fysics_amd")
    if kind == "shared":
        link = ["-L" + libdir, "-lfysics_amd", "-Wl,-rpath," + libdir]
    else:
        link = [os.path.join(libdir, "libfysics_amd.a"), "-L/opt/rocm/lib", "-lamdhip64",
                "-Wl,-rpath,/opt/rocm/lib", "-lstdc++", "-lpthread", "-lm"]
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)] + link,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    v = out.stdout.split()
    p = F.physics(delta_beta=300.0, energy_kev=110.0, distance_mm=1200.0, pixel_um=4.3, unsharp_sigma=0.8,
                  unsharp_coeff=0.5)
    assert v[:3] == ["192", "96", "1"]
    assert float(v[3]) == F.lib().fy_paganin_transfer(0.25, C.byref(p))
    assert float(v[4]) == F.lib().fy_guided_eps_for_noise(0.01)
