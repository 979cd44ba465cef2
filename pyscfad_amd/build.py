"""Build libpyscfad_amd.so (gfx950) in-tree with hipcc.

    python -m pyscfad_amd.build

All HIP code is compiled with packed FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host-pass warning
that hipcc prints about that feature is expected.
"""
import concurrent.futures
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpyscfad_amd.so")
ARCH = os.environ.get("PYSCFAD_AMD_ARCH", "gfx950")

HIP_SOURCES = ["t3_gemm.hip", "t3_epilogue.hip", "t3_prep.hip", "t3_noalloc.hip", "t3_lno.hip"]
CPP_SOURCES = ["t3_driver.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    return os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    # hipcc compiles .cpp as HIP too (host code here, but its device pass must
    # target gfx950 like the rest, not the compiler's default architecture)
    cmd = [_hipcc()] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build():
    os.makedirs(OBJDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES + CPP_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
