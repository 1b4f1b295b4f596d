"""Build recipe of the CPU oracle (test infrastructure only).

Compiles oracle/oracle.c and oracle/ws_oracle.c into oracle/liboracle.so with gcc, floating-point
contraction off so that FP32/FP64 operation order is exactly what the source
says, and OpenMP for the all-cores CPU baseline (ref_particles_omp).  The reference implementation itself is Julia and cannot be built here
(no Julia toolchain), so there is no oracle/_ref build.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "oracle.c"), os.path.join(HERE, "ws_oracle.c")]
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
           "-fno-fast-math", "-fopenmp", "-Wall", "-Wno-unused-function", "-o", tmp] + SRC + [
               "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
