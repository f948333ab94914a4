"""Build libafd.so (the gfx950 HIP library behind include/afd.h), in-tree.

    python -m astrofd_amd.build [-j N] [--verbose]

Each translation unit under astrofd_amd/csrc compiles with hipcc for gfx950 in
parallel; the RMHD face kernels compile twice (FP64 and the FP32 mixed path).
Output: astrofd_amd/_lib/libafd.so.
"""
import argparse
import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libafd.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
# Packed FP32 VALU ops are disabled (-packed-fp32-ops): gfx950 code is written for scalar FP32/FP64.
# MFMAs take their accumulators in ordinary VGPRs (-amdgpu-mfma-vgpr-form): with the accumulation
# registers the network's forward kernel copied every result out with v_accvgpr_read (146 a wave)
# and ran at 3 waves per SIMD instead of 5.
HIPFLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-Wno-unused-function",
            "-Wno-unused-variable", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form=1", "-I", INCLUDE, "-I", CSRC]

UNITS = [
    ("afd_capi.hip", []),
    ("afd_common.hip", []),
    ("afd_srhd.hip", []),
    ("afd_mlp.hip", []),
    ("afd_fused.hip", []),
    ("afd_rmhd.hip", ["-DAFD_RMHD_T=double"], "afd_rmhd_f64"),
    ("afd_rmhd.hip", ["-DAFD_RMHD_T=float"], "afd_rmhd_f32"),
]


def _compile(unit, verbose):
    src, extra = unit[0], unit[1]
    name = unit[2] if len(unit) > 2 else os.path.splitext(src)[0]
    obj = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC] + HIPFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src} ({name})")
    return obj, time.time() - t0


def build(jobs=None, verbose=False):
    """Compile every HIP translation unit for gfx950 and link libafd.so.  Always rebuilds."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, (os.cpu_count() or 2)))
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, verbose), UNITS))
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("link of libafd.so failed")
    os.replace(tmp, LIB)
    if verbose:
        for (o, dt) in results:
            print(f"  {os.path.basename(o)}: {dt:.1f}s")
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    build(a.j, a.verbose)


if __name__ == "__main__":
    main()
