"""Build the pypde_amd libraries in-tree.

    python pypde_amd/build.py        (or __graft_entry__.build())

pypde_amd/libpypde_hip.so (gfx950): each .hip source is compiled to an object
with hipcc, then linked into one shared library exporting the C-ABI declared in
include/pypde_amd.h.

pypde_amd/libpypde_h5.so (host): csrc/h5state.c with gcc against the HDF5 C
library (headers and libhdf5 under HDF5_DIR, default /opt/conda), the state
files of include/pypde_amd_h5.h.  Skipped with a message when HDF5 is absent;
the IO functions then raise.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpypde_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
H5LIB = os.path.join(HERE, "libpypde_h5.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=true",
    "-I" + os.path.join(REPO, "include"),
    "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return obj


def build_h5(verbose=True):
    """Compile the HDF5 state-file library; returns its path or None without HDF5."""
    inc = os.path.join(HDF5_DIR, "include")
    lib = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print("HDF5 headers not found under %s: libpypde_h5.so not built" % HDF5_DIR)
        return None
    tmp = H5LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-shared", "-fPIC", "-Wall", "-I" + inc,
           "-I" + os.path.join(REPO, "include"), os.path.join(CSRC, "h5state.c"),
           "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed for h5state.c:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, H5LIB)
    if verbose:
        print("built", H5LIB)
    return H5LIB


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
