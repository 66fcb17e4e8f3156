"""Ahead-of-time build of the gfx950 HIP library ``libgpuocean_amd.so``.

Run ``python -m gpuocean_amd.build`` (or ``__graft_entry__.build()``).  Each
translation unit is compiled with hipcc for gfx950 only; the step kernel is
compiled twice (fast and exact arithmetic) into one shared library.
"""

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgpuocean_amd.so")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for every HIP translation unit.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-command-line-argument",
]

UNITS = [
    ("cdklm_step_fast.hip", ["-ffp-contract=fast"]),
    ("cdklm_step_exact.hip", ["-ffp-contract=off"]),
    ("aux_kernels.hip", ["-ffp-contract=off"]),
    ("noise_drift.hip", ["-ffp-contract=off"]),
    ("ensemble.hip", ["-ffp-contract=off"]),
    ("adjoint.hip", ["-ffp-contract=off"]),
    ("adjoint_rev32.hip", ["-ffp-contract=on"]),
    ("observe.hip", ["-ffp-contract=off"]),
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return path


def build(verbose=False, jobs=4):
    cc = hipcc()
    units = [(u, f) for (u, f) in UNITS if os.path.exists(os.path.join(CSRC, u))]
    with tempfile.TemporaryDirectory(prefix="gpo_build_") as tmp:
        def compile_one(item):
            src, flags = item
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            cmd = [cc, "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + flags
            if verbose:
                print(" ".join(cmd), flush=True)
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n{res.stdout}\n{res.stderr}")
            return obj

        with ThreadPoolExecutor(max_workers=jobs) as ex:
            objs = list(ex.map(compile_one, units))
        tmp_lib = LIB + ".tmp"
        cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp_lib] + objs
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
