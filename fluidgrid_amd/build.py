"""Builds libfluidgrid_amd.so in-tree with hipcc for gfx950.

Usage: python -m fluidgrid_amd.build [--jobs N]

Every .hip file under csrc/ and every .cpp file under csrc/host/ is compiled
to an object in fluidgrid_amd/_build/ and linked into
fluidgrid_amd/libfluidgrid_amd.so.  Device code is compiled with the packed
FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfluidgrid_amd.so")
ARCH = os.environ.get("FG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero",
]


def sources():
    hip = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    host = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    return hip, host


def _compile(src, is_hip):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    # Host sources include the HIP runtime headers, so they go through the
    # same HIP compile (their device pass is empty).
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the device-only feature.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
    return obj, err


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hip, host = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, True) for s in hip] + [ex.submit(_compile, s, False) for s in host]
        for f in futs:
            obj, err = f.result()
            objs.append(obj)
            if verbose and err:
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-Wl,-z,defs"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_app()
    return LIB


APP_SRC = os.path.join(HERE, "app", "AppMain.cpp")
APP = os.path.join(HERE, "fg_appmain")


def build_app():
    """The C++ host program (fluidgrid_amd/app/AppMain.cpp) linked against the
    library, found next to it at run time."""
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), f"--offload-arch={ARCH}",
           APP_SRC, "-o", APP + ".tmp", "-L", HERE, "-lfluidgrid_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"app build failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(APP + ".tmp", APP)
    return APP


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
