"""Build libmchip.so (HIP kernels for gfx950 + host C++) in-tree.

Every translation unit is compiled with hipcc for gfx950; packed-FP32 VALU ops
are disabled for device code.  Objects go to chem_amd/_build/, the library to
chem_amd/lib/libmchip.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmchip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("MCHIP_CXX", "/opt/rocm/llvm/bin/clang++")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
LINK = ["-shared", "-fopenmp", "--offload-arch=gfx950", "-L/opt/rocm/lib",
        "-lrocblas", "-lrocsolver", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]


def sources():
    out = []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".cpp", ".hip")):
                out.append(os.path.join(d, f))
    return sorted(out)


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(BUILD, rel + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip"]
    else:
        cmd = [CXX] + COMMON
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


PROBE_SRC = os.path.join(ROOT, "tests", "native", "probe.hip")
PROBE_LIB = os.path.join(ROOT, "tests", "native", "libmchip_probe.so")


def build_probe():
    """Test-only library of builtin-layout probes (tests/native/probe.hip),
    loaded by the tests alone."""
    if not os.path.exists(PROBE_SRC):
        return None
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", PROBE_SRC, "-shared", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("probe build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "native", "fock_driver.cpp")
DRIVER_EXE = os.path.join(ROOT, "tests", "native", "fock_driver")


def build_driver():
    """Test driver of the C++ builder layer (tests/native/fock_driver.cpp),
    linked against the in-tree libmchip.so with a run path relative to itself."""
    if not os.path.exists(DRIVER_SRC):
        return None
    tmp = DRIVER_EXE + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), DRIVER_SRC,
           "-L", LIBDIR, "-lmchip", "-Wl,-rpath,$ORIGIN/../../chem_amd/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, DRIVER_EXE)
    return DRIVER_EXE


WFN_SRC = os.path.join(ROOT, "chem_amd", "wfn", "wfn_h5.cpp")
WFN_LIB = os.path.join(LIBDIR, "libmchip_wfn.so")
HDF5_ROOT = os.environ.get("MCHIP_HDF5_ROOT", "/opt/conda")


def build_wfn():
    """libmchip_wfn.so: the HDF5 wavefunction files (host only), against the
    HDF5 C library of HDF5_ROOT (headers searched after the system ones, run
    path to its lib).  Skipped when that HDF5 is absent."""
    if not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        return None
    tmp = WFN_LIB + ".tmp"
    # libstdc++ linked in statically: the run path must not pull HDF5_ROOT's
    # C++ runtime into processes that load the system one
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Werror", "-static-libstdc++",
           "-static-libgcc", "-idirafter", os.path.join(HDF5_ROOT, "include"), WFN_SRC,
           "-L", os.path.join(HDF5_ROOT, "lib"), "-lhdf5",
           "-Wl,-rpath," + os.path.join(HDF5_ROOT, "lib"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("wfn build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, WFN_LIB)
    return WFN_LIB


WFN_DRIVER_SRC = os.path.join(ROOT, "tests", "native", "wfn_driver.cpp")
WFN_DRIVER_EXE = os.path.join(ROOT, "tests", "native", "wfn_driver")


def build_wfn_driver():
    """Test driver of the C++ wavefunction-file layer (tests/native/wfn_driver.cpp)."""
    if not (os.path.exists(WFN_DRIVER_SRC) and os.path.exists(WFN_LIB)):
        return None
    tmp = WFN_DRIVER_EXE + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), WFN_DRIVER_SRC,
           "-L", LIBDIR, "-lmchip_wfn", "-Wl,-rpath,$ORIGIN/../../chem_amd/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("wfn driver build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, WFN_DRIVER_EXE)
    return WFN_DRIVER_EXE


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = min(8, len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + objs + LINK + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_probe()
    build_driver()
    build_wfn()
    build_wfn_driver()
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
