"""Build every native part of the package in-tree.

  libmonsel_gpu.so   hipcc, gfx950: kernels + the C-ABI of include/monsel_gpu.h
  libmonsel_host.so  g++: the reference-mirror host API (host/monsel_host.cpp)
  _monsel*.so        g++ + pybind11: Python bindings of the host API
  lines_on_layers    g++: a LinesOnLayers-style C++ driver on the host API
                     (host/lines_on_layers.cpp)

Usage: python -m cudamonsel_amd.build   (or build() from Python)
       python -m cudamonsel_amd.build --phase-timing
           the same with the transport kernel's per-phase wave timing
           compiled in (csrc/phase_timing.h); for profiling only
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("MONSEL_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 instructions stay off: some of their forms have returned
    # wrong results on gfx950 without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 divide and sqrt to the hardware's few-ulp sequences instead of the
    # correctly rounded expansions (transport tolerances are statistical)
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # FP32 denormals flush to zero: no transport quantity is that small (a
    # rate below 1e-38 /nm is none), and sqrt, divide, exp and log then need
    # no rescaling around the hardware instructions (8 % fewer VALU
    # instructions in the transport kernel); FP64 keeps its denormals
    "-fgpu-flush-denormals-to-zero",
]
CXX_FLAGS = ["-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build(verbose=True, phase_timing=False):
    gpu = os.path.join(HERE, "libmonsel_gpu.so")
    extra = ["-DMONSEL_PHASE_TIMING"] if phase_timing else []
    host = os.path.join(HERE, "libmonsel_host.so")
    mod = os.path.join(HERE, "_monsel" + ext_suffix())
    _run([os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, *extra,
          os.path.join(HERE, "csrc", "monsel_gpu.hip"), "-o", gpu], verbose)
    _run(["g++", *CXX_FLAGS, os.path.join(HERE, "host", "monsel_host.cpp"),
          "-L" + HERE, "-lmonsel_gpu", "-Wl,-rpath,$ORIGIN", "-ldl", "-o", host], verbose)
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", *CXX_FLAGS, "-I" + pybind11.get_include(), "-I" + py_inc,
          os.path.join(HERE, "host", "pybind_monsel.cpp"),
          "-L" + HERE, "-lmonsel_host", "-lmonsel_gpu", "-Wl,-rpath,$ORIGIN", "-o", mod], verbose)
    driver = os.path.join(HERE, "lines_on_layers")
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "host", "lines_on_layers.cpp"),
          "-L" + HERE, "-lmonsel_host", "-lmonsel_gpu", "-Wl,-rpath,$ORIGIN", "-o", driver], verbose)
    return [gpu, host, mod, driver]


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv, phase_timing="--phase-timing" in sys.argv)
