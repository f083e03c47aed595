"""Builds libbm3d_mi355x.so (gfx950) in-tree with hipcc.

    python vapoursynth-bm3dcuda_amd/build_lib.py [--jobs N]

Sources: csrc/*.hip (device code) and csrc/*.cpp (launch layer, host
filters). Output: lib/libbm3d_mi355x.so next to this file. Floating-point
contraction is off so that every FMA in the kernels is an explicit one, and
packed FP32 instructions are disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libbm3d_mi355x.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def compile_one(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(compile_one, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for (_, err), src in zip(results, srcs):
            warn = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l)
            if warn.strip():
                print(f"[{os.path.basename(src)}]\n{warn}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_plugins()
    return LIB


VSPLUGIN = os.path.join(HERE, "vsplugin")
CXX = os.environ.get("CXX", "g++")
# VapourSynth plugin entries over the library: (output, extra flags). The
# first serves the bm3dcuda namespace, the second bm3dcuda_rtc (the NVRTC
# variant's string options), the third a namespace of its own (bm3dmi355x)
# that can be installed beside either.
PLUGINS = [
    ("libbm3dcuda_mi355x.so", []),
    ("libbm3dcuda_rtc_mi355x.so", ["-DBM3D_VS_RTC"]),
    ("libbm3dmi355x.so", ["-DBM3D_VS_OWN_NAMESPACE"]),
]


def plugin_cmd(out, extra, compiler=None, sanitize=False):
    """Compile line of one plugin entry (host code only: it calls the C ABI)."""
    lib_dir = os.path.dirname(out)
    san = SANITIZE if sanitize else []
    return [compiler or CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall", *san, *extra,
            f"-I{os.path.join(ROOT, 'include')}", f"-I{VSPLUGIN}", os.path.join(VSPLUGIN, "bm3d_vsplugin.cpp"),
            f"-L{lib_dir}", "-lbm3d_mi355x", "-Wl,-rpath,$ORIGIN", "-o", out]


CLANGXX = "/opt/rocm/llvm/bin/clang++"
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def build_sanitized(out_dir, jobs=4):
    """AddressSanitizer + UBSan build of the host code, into out_dir: the
    library with its host-only sources instrumented (device code and the
    kernel launchers stay as built by build()), both plugin entries, and
    the mock VapourSynth host of the tests. Every sanitizer flag applies to
    the host pass only. Returns the path of the mock host executable."""
    if not os.path.exists(LIB):
        build(jobs)
    os.makedirs(out_dir, exist_ok=True)
    host_flags = [f for flag in SANITIZE for f in ("-Xarch_host", flag)]

    def compile_san(src):
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        cmd = [HIPCC, *COMMON, *DEVICE, *host_flags, "-x", "hip", "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return obj

    host_srcs = [s for s in sources() if s.endswith(".cpp")]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_san, host_srcs))
    objs += [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in sources() if s.endswith(".hip")]
    lib = os.path.join(out_dir, "libbm3d_mi355x.so")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    for name, extra in PLUGINS:
        cmd = plugin_cmd(os.path.join(out_dir, name), extra, compiler=CLANGXX, sanitize=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"plugin build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    host = os.path.join(out_dir, "mock_vs_host")
    cmd = [CLANGXX, "-O1", "-std=c++17", *SANITIZE, f"-I{VSPLUGIN}",
           os.path.join(ROOT, "tests", "vs_host", "mock_vs_host.cpp"), "-o", host, "-ldl", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mock host build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return host


def build_plugins():
    for name, extra in PLUGINS:
        out = os.path.join(OUT_DIR, name)
        cmd = plugin_cmd(out + ".tmp", extra)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"plugin build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose))
    sys.exit(0)
