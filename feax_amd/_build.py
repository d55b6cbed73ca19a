"""Build libfeaxhip.so (gfx950) in-tree with hipcc.

Usage: ``python -m feax_amd._build`` or ``feax_amd._build.build()``.
The library lands in ``feax_amd/_lib/libfeaxhip.so``; it is git-ignored but
travels with the source tree to the GPU host.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libfeaxhip.so")

SOURCES = ["pattern.cpp", "krylov.hip", "hex8.hip", "adjoint.hip", "batch.hip", "gmg.hip", "bicgstab.hip", "gmres.hip",
           "direct.hip", "tet4.hip", "quad4.hip", "hex8s.hip",
           "hex20.hip", "hex8ut.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # No packed FP32 instructions: hand the compiler scalar FP32 only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# rocSOLVER / rocBLAS: the potrf / trsm / syrk blocks of the banded Cholesky
# (direct.hip).
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}): {cmd[-1]}")


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel, then link the shared
    library (one hipcc per source: each carries its own gfx950 code object)."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    compile_flags = [fl for fl in HIPCC_FLAGS if fl != "-shared"]
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([cc, *compile_flags, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
                     os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, *LIBS, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
