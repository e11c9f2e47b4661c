"""Compiles the HIP library of the nonbonded cutoff force for gfx950.

The shared object is written in-tree (``protomoladdon_amd/libpmnb.so``) so it
travels with the source tree; nothing is installed.
"""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libpmnb.so")
SOURCES = ["pmnb.hip"]
HEADERS = ["pmnb_common.h", "pmnb_cells.h", "pmnb_pair.h", "pmnb_list.h",
           "pmnb_tile.h", "pmnb_queue.h", "pmnb_adjoint.h",
           "pmnb_tilemask.h", "pmnb_leapfrog.h", "pmnb_ewald.h", "pmnb_pme.h"]

# -ffp-contract=off: FP64 expressions round exactly as the reference's x86-64
# build (no FMA); FP32 kernels ask for fused multiply-adds explicitly.
# -packed-fp32-ops: packed FP32 instructions are kept out of the code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# hipFFT for the PME transforms (a library FFT, as rocBLAS is for GEMMs)
LINK_FLAGS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lhipfft"]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False):
    """Compiles libpmnb.so (always; callers decide when to rebuild)."""
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", LIB_PATH + ".tmp"] + [
        os.path.join(CSRC, s) for s in SOURCES] + LINK_FLAGS
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
