"""Build recipe for the oracle's C restatement (test infrastructure only).

The reference's own C for this path (pyscfadlib/pyscfadlib/vjp/cc/*.c) is not
built here: it includes a CMake-generated ``config.h`` and links a Fortran-ABI
BLAS (``dgemm_``/``dgemv_``), neither of which this image provides.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ccsd_t_oracle.c")
LIB = os.path.join(HERE, "liboracle_ccsd_t.so")


def build(force=True):
    """Compile ccsd_t_oracle.c into liboracle_ccsd_t.so (gcc, OpenMP).

    -mavx2 -mfma rather than -march=native: the library is built in one
    container and run on the host CPUs of another machine.
    """
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    cmd = ["gcc", "-O3", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared",
           "-Wall", "-o", LIB + ".tmp", SRC]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
