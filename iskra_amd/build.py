"""Build recipe of libiskra_hip.so (gfx950), of libiskra_pmd.so (the openPMD
writer) and of the CPU oracle.

`python -m iskra_amd.build` compiles every HIP source under iskra_amd/csrc into
iskra_amd/libiskra_hip.so with hipcc, and csrc/openpmd.c into
iskra_amd/libiskra_pmd.so with the C compiler against libhdf5 (found under
$ISKRA_HDF5_DIR, else /opt/conda, else /usr; the library records that lib
directory as its run path).  Without HDF5 headers the writer is not built
and the diagnostics module raises when asked to write.  The oracle (test
infrastructure) is built by oracle/build.py.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libiskra_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["api.hip", "particles.hip", "poisson.hip", "mcc.hip", "dsmc.hip", "electrode.hip",
           "fluid.hip", "warmstart.hip", "batch.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


PMD_LIB = os.path.join(HERE, "libiskra_pmd.so")
CC = os.environ.get("CC", "gcc")


def hdf5_dir():
    """Prefix holding include/hdf5.h and lib/libhdf5.so, or None."""
    for d in (os.environ.get("ISKRA_HDF5_DIR"), "/opt/conda", "/usr"):
        if d and os.path.exists(os.path.join(d, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(d, "lib", "libhdf5.so")):
            return d
    return None


def build_pmd(verbose=False):
    """libiskra_pmd.so from csrc/openpmd.c; None when HDF5 is absent."""
    d = hdf5_dir()
    if d is None:
        if verbose:
            print("libiskra_pmd.so not built: no HDF5 headers and library found", flush=True)
        return None
    tmp = PMD_LIB + ".tmp"
    libdir = os.path.join(d, "lib")
    cmd = [CC, "-O2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I" + os.path.join(d, "include"), os.path.join(CSRC, "openpmd.c"), "-o", tmp,
           "-L" + libdir, "-lhdf5", "-lm", "-Wl,--enable-new-dtags,-rpath," + libdir]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr + "\n")
        raise RuntimeError("the C compiler failed building libiskra_pmd.so")
    if verbose and (r.stdout + r.stderr).strip():
        print(r.stdout + r.stderr)
    os.replace(tmp, PMD_LIB)
    return PMD_LIB


def build(verbose=False):
    build_pmd(verbose)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("hipcc failed building libiskra_hip.so")
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
