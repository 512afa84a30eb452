"""Build libdolfin_amd.so (HIP kernels + C ABI + host structure phase) for gfx950.

Usage: python -m dolfin_amd.build   (or dolfin_amd.build.build())
Objects go to dolfin_amd/_build/, the library to dolfin_amd/libdolfin_amd.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdolfin_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["api.hip", "p1_kernels.hip", "p2_kernels.hip", "adjoint_kernels.hip",
           "batch_kernels.hip", "facet_kernels.hip", "la_kernels.hip", "csr_build.hip", "plan_device.hip", "rcm_device.hip", "color_device.hip", "plan_vp2_device.hip",
           "structure.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 instructions are disabled on purpose (all arithmetic here is FP64,
          # and the packed forms are not trusted on this platform)
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result", "-pthread"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, src + ".o")
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "c++"] + [c for c in COMMON if not c.startswith("--offload")
                                      and c not in ("-munsafe-fp-atomics",)] + ["-c", path, "-o", obj]
        # the packed-fp32 flag pair only concerns device code
        cmd = [c for c in cmd]
        i = cmd.index("-Xclang")
        del cmd[i:i + 4]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compilation of {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


EXAMPLE_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "system_assembler_example.cpp")
EXAMPLE_BIN = os.path.join(os.path.dirname(HERE), "tests", "_build", "cpp_system_assembler")


def build_example():
    """C++ host program against the header-only adapter, linked to the in-tree library."""
    os.makedirs(os.path.dirname(EXAMPLE_BIN), exist_ok=True)
    # host-only C++ against the HIP runtime API (the adapter launches nothing itself)
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           EXAMPLE_SRC, "-o", EXAMPLE_BIN, f"-L{HERE}", "-ldolfin_amd", "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../dolfin_amd", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return EXAMPLE_BIN


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_example()
    if verbose:
        print("built", LIB, "and", EXAMPLE_BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
