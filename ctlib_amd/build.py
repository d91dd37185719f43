"""Build libctlib_amd.so (HIP kernels + C ABI) and the torch extension _C.

Both land in-tree next to this file so that they travel with the source tree:
    ctlib_amd/libctlib_amd.so   gfx950 kernels, FP64 plan builder, C ABI
    ctlib_amd/_C.so             PyTorch-ROCm extension linking the library

Usage:  python -m ctlib_amd.build [--arch gfx950] [--verbose]
"""
import argparse
import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libctlib_amd.so")
EXT = os.path.join(HERE, "_C.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# Packed FP32 VALU ops are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-pass-failed"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    return r


def _torch_flags():
    import torch
    from torch.utils import cpp_extension as ce
    inc = ce.include_paths()
    libdirs = ce.library_paths()
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cflags = [f"-I{p}" for p in inc] + [f"-I{sysconfig.get_paths()['include']}", f"-I{ROCM}/include",
                                        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
                                        "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H",
                                        f"-D_GLIBCXX_USE_CXX11_ABI={abi}"]
    ldflags = [f"-L{p}" for p in libdirs] + [f"-Wl,-rpath,{p}" for p in libdirs] + \
        ["-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python"]
    return cflags, ldflags


def build(arch="gfx950", verbose=False, with_torch=True):
    """Compile every source; always rebuilds (a few tens of seconds)."""
    os.makedirs(BUILD, exist_ok=True)
    objs = {
        "plan.o": ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
                   "-c", os.path.join(CSRC, "plan.cpp")],
        "abi.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "abi.cpp")],
        "dd_kernels.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "dd_kernels.hip")],
        "dd_adj.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "dd_adj.hip")],
        "dd_fwd_w.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "dd_fwd_w.hip")],
        "ramp_mfma.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "ramp_mfma.hip")],
        "batch.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "batch.hip")],
        "laplacian.o": [HIPCC, f"--offload-arch={arch}", *HIP_FLAGS, "-c", os.path.join(CSRC, "laplacian.hip")],
    }
    jobs = []
    for name, cmd in objs.items():
        jobs.append(cmd + ["-o", os.path.join(BUILD, name)])
    ext_cmd = None
    if with_torch:
        cflags, ldflags = _torch_flags()
        ext_obj = os.path.join(BUILD, "torch_ext.o")
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-w", *cflags, "-c",
                     os.path.join(CSRC, "torch_ext.cpp"), "-o", ext_obj])
        ext_cmd = ["g++", "-shared", "-o", EXT + ".tmp", ext_obj, f"-L{HERE}", "-lctlib_amd",
                   "-Wl,-rpath,$ORIGIN", *ldflags, f"-L{ROCM}/lib", "-lamdhip64"]
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    _run([HIPCC, f"--offload-arch={arch}", "-shared", "-fPIC", "-o", LIB + ".tmp",
          *[os.path.join(BUILD, n) for n in objs], f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx",
          f"-Wl,-rpath,{ROCM}/lib"], verbose)
    shutil.move(LIB + ".tmp", LIB)
    if ext_cmd:
        _run(ext_cmd, verbose)
        shutil.move(EXT + ".tmp", EXT)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-torch", action="store_true")
    a = ap.parse_args()
    build(a.arch, a.verbose, not a.no_torch)
    print("built", LIB, "" if a.no_torch else EXT)


if __name__ == "__main__":
    main()
