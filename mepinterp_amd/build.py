"""Build the gfx950 HIP library (libmepinterp_amd.so) in-tree.

    python -m mepinterp_amd.build            # library only
The Fortran host is built by mepinterp_amd/fortran/Makefile.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmepinterp_amd.so")
SOURCES = ["mep_ft.hip", "mep_eigh.hip", "mep_refine.hip", "mep_velo.hip", "mep_bandsum.hip",
           "mep_mlp.hip",
           "mep_extsum.hip", "mep_capi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -amdgpu-mfma-vgpr-form: MFMA accumulators in VGPRs.  Otherwise the
# compiler may place them in AGPRs and copy them to and from VGPRs around
# every product of a loop (24 + 24 v_accvgpr moves per three FP64 MFMAs in
# the wide velocity and band-sum kernels), and VALU issue serialises with
# FP64 MFMA on gfx950.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wno-unused-result"]


def build_lib(verbose: bool = False) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    errs = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errs.append(f"--- {src} ---\n{out.decode(errors='replace')}")
        elif verbose and out:
            print(out.decode(errors="replace"))
    if errs:
        raise RuntimeError("HIP build failed:\n" + "\n".join(errs))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp],
                   check=True)
    os.replace(tmp, LIB)
    # calibration copy kernel for the HBM-traffic counters (profiling only)
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", os.path.join(CSRC, "copy_calib.hip"),
                    "-o", os.path.join(bdir, "copy_calib")], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return LIB


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
