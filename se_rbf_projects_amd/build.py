"""Build libserbf.so (hand-written gfx950 HIP kernels + C-ABI) in-tree with hipcc.

    python -m se_rbf_projects_amd.build      # or se_rbf_projects_amd.build.build()

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libserbf.so")
SOURCES = ["engine.hip", "rk4.hip", "knn.hip", "weights.hip", "rowlu_96.hip", "rowlu_88.hip", "rowlu_80.hip", "rowlu_72.hip", "rowlu_48_64.hip", "rowhalf.hip", "nodes.cpp",
           "nodefile.cpp", "dropin.hip", "tile.hip", "closure.hip", "mtile.hip"]
ARCH = "gfx950"


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flags(src):
    f = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]
    if src.endswith(".cpp"):
        f += ["-x", "hip"]
    return f


def _stale(src, obj):
    """True when obj is missing or older than its source or any header it may include."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(d, f) for d in (CSRC, os.path.join(HERE, "..", "include"))
                                        for f in os.listdir(d) if f.endswith(".h")]
    return max(os.path.getmtime(d) for d in deps) > os.path.getmtime(obj)


def _compile(src, incremental=False):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [hipcc()] + _flags(src) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile {src} failed:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True, incremental=False):
    """Compile every source (incremental=True: only those older than their inputs) and link."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    try:
        build(incremental="--incremental" in sys.argv)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
