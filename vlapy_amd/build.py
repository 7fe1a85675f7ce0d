"""Build libvpfp.so in-tree with hipcc for gfx950.

    python -m vlapy_amd.build

The row and drift kernels are compiled once per storage type (and row
variant) from vpfp_row.hip and vpfp_drift.hip, the rest from vpfp.hip and
vpfc.hip; the objects compile in parallel into vlapy_amd/_build/ and link
into vlapy_amd/libvpfp.so.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libvpfp.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# (object name, source, defines)
UNITS = [
    ("vpfp", "vpfp.hip", []),
    ("vpfc", "vpfc.hip", []),
    ("row_f64", "vpfp_row.hip", ["VPFP_ROW_T=double", "VPFP_ROW_BG=0"]),
    ("row_f32", "vpfp_row.hip", ["VPFP_ROW_T=float", "VPFP_ROW_BG=0"]),
    ("rowbg_f64", "vpfp_row.hip", ["VPFP_ROW_T=double", "VPFP_ROW_BG=1"]),
    ("rowbg_f32", "vpfp_row.hip", ["VPFP_ROW_T=float", "VPFP_ROW_BG=1"]),
    ("drift_f64", "vpfp_drift.hip", ["VPFP_DRIFT_T=double"]),
    ("drift_f32", "vpfp_drift.hip", ["VPFP_DRIFT_T=float"]),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _clean(stderr):
    return "\n".join(l for l in stderr.splitlines() if "packed-fp32-ops" not in l)


def _compile(unit, verbose):
    name, src, defines = unit
    obj = os.path.join(OBJDIR, name + ".o")
    cmd = ([hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-I" + os.path.join(ROOT, "include"),
                                                                   "-I" + CSRC, "-c", os.path.join(CSRC, src),
                                                                   "-o", obj])
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (name, _clean(res.stderr)))
    return obj


def build(verbose=True, jobs=None):
    """Compile vlapy_amd/csrc into vlapy_amd/libvpfp.so (always recompiles)."""
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose), UNITS))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + _clean(res.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
