"""Build libfysics_amd.so in-tree: hipcc for the HIP kernels (gfx950), gcc for
the C host code, linked into one shared library next to this file.

    python -m fysics_amd.build [--force] [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfysics_amd.so")
ARCHIVE = os.path.join(HERE, "libfysics_amd.a")
ARCH = os.environ.get("FYSICS_OFFLOAD_ARCH", "gfx950")

HIP_SRC = ["fyg_runtime.hip", "fyg_fft.hip", "fyg_fft64.hip", "fyg_mix.hip", "fyg_guided.hip", "fyg_elem.hip",
           "fyg_noise.hip", "fyg_calib.hip", "fyg_air.hip",
           "fyg_phasecorr.hip", "fyg_noalloc.hip"]
C_SRC = ["host/fy_core.c", "host/fy_deconv.c", "host/fy_guided.c", "host/fy_dering.c",
         "host/fy_chain.c", "host/fy_noise.c", "host/fy_serve.c", "host/fy_calib.c",
         "host/fy_zarr.c", "host/fy_export.c"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             # packed fp32 VALU forms are kept out of every kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # multiply-adds fuse only within one source expression, so a kernel
             # body gives the same bits in every kernel it is inlined into (the
             # per-shape kernels and the mixed-shape stage kernels)
             "-ffp-contract=on",
             "-Wno-unused-command-line-argument"]
C_FLAGS = ["-std=gnu11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


HEADERS = [os.path.join(ROOT, "include", "fysics.h"), os.path.join(ROOT, "include", "fysics_gpu.h")]


def exported_names():
    """Every function the public headers declare: the shared library exports
    these and nothing else (the kernels, the fyg_* runtime and the fyh_* host
    helpers stay internal)."""
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"typedef struct \{.*?\}\s*\w+;", "", src, flags=re.S)
        names.update(m.group(1) for m in re.finditer(r"\b(fy_\w+)\s*\(", src))
    return sorted(names)


def _version_script():
    path = os.path.join(OBJ, "exports.map")
    body = "{\n  global:\n" + "".join(f"    {n};\n" for n in exported_names()) + "  local:\n    *;\n};\n"
    with open(path, "w") as f:
        f.write(body)
    return path


def _deps(src):
    hdrs = [os.path.join(CSRC, "fyg.h"), os.path.join(CSRC, "fyg_internal.hpp"),
            os.path.join(CSRC, "fyg_fft_engine.hpp"),
            os.path.join(CSRC, "host", "fyh.h"), os.path.join(ROOT, "include", "fysics.h"),
            os.path.join(ROOT, "include", "fysics_gpu.h")]
    return [src] + [h for h in hdrs if os.path.exists(h)]


def _stale(obj, src):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in _deps(src))


def _compile(job):
    kind, src, obj = job
    if kind == "hip":
        cmd = [hipcc()] + HIP_FLAGS + ["-I", CSRC, "-c", src, "-o", obj]
    else:
        cmd = ["gcc"] + C_FLAGS + ["-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stderr or "").splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l.lower())
    return src, r.returncode, err


def build(force=False, jobs=None, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    work = []
    objs = []
    for s in HIP_SRC:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(obj)
        if force or _stale(obj, src):
            work.append(("hip", src, obj))
    for s in C_SRC:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(obj)
        if force or _stale(obj, src):
            work.append(("c", src, obj))
    jobs = jobs or min(8, os.cpu_count() or 4)
    # the FFT unit is the long pole: start it first
    work.sort(key=lambda j: 0 if ("fft" in j[1] or "mix" in j[1]) else 1)
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, err in ex.map(_compile, work):
            if err.strip() and verbose:
                print(f"[build] {os.path.relpath(src, ROOT)}:\n{err}", file=sys.stderr)
            if rc != 0:
                failed.append(src)
    if failed:
        raise RuntimeError("compile failed: " + ", ".join(os.path.relpath(f, ROOT) for f in failed))
    if force or work or not os.path.exists(LIB) or not os.path.exists(ARCHIVE):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + \
              ["-Wl,-Bsymbolic", "-Wl,--version-script=" + _version_script(), "-lpthread", "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr)
        os.replace(tmp, LIB)
        # the static archive of the same objects, for callers that link the
        # reference's static libfysics.a (CMakeLists.txt:55-62)
        atmp = ARCHIVE + ".tmp"
        if os.path.exists(atmp):
            os.remove(atmp)
        r = subprocess.run(["ar", "rcs", atmp] + objs, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("ar failed:\n" + r.stderr)
        os.replace(atmp, ARCHIVE)
    if verbose:
        print(f"[build] {os.path.relpath(LIB, ROOT)} ({len(work)} units compiled)")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build(force=a.force, jobs=a.j)
