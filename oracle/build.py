"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/ostr_oracle.cpp`` into ``oracle/libostr_oracle.so`` with
g++ and OpenMP. Contraction is switched off so that only the explicit fma()
calls of the restatement fuse; x86-64-v3 makes those calls single FMA
instructions instead of libm calls.

The reference's own CUDA/C++ path cannot be built here (it needs the CUDA
runtime, nvcc and the headers Tomo_OSTR.h / readdata.h, none of which exist),
so there is no oracle/_ref build.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ostr_oracle.cpp")
LIB = os.path.join(HERE, "libostr_oracle.so")

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
    "-ffp-contract=off", "-march=x86-64-v3", "-fno-math-errno",
    "-Wall", "-Wno-unknown-pragmas",
]


def build(verbose: bool = False) -> str:
    """Compile the oracle library; returns its path."""
    cmd = ["g++", *CXXFLAGS, SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
