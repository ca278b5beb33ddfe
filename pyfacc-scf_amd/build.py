"""Build libpyfacc.so for gfx950 with hipcc, in-tree (``pyfacc-scf_amd/lib``).

Every source is compiled to an object in parallel, then linked into one
shared library exporting the C-ABI of ``include/pyfacc.h``.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(HERE, "lib", "libpyfacc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYFACC_ARCH", "gfx950")

FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for every device compile
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result",
]
ERI_CLASSES = ["ssss", "psss", "ssps", "psps", "ppss", "sspp", "ppps", "pspp", "pppp"]
SOURCES = ["eri_%s.hip" % c for c in ERI_CLASSES] + ["grad_%s.hip" % c for c in ERI_CLASSES] + ["grad.hip", "eri.hip", "onee.hip", "linalg.hip", "engine.hip", "batch.hip", "eri_batch.hip", "pairs.hip", "eri_export.hip", "generic.hip", "testkern.hip", "noalloc.hip", "basis.cpp", "pool.cpp", "abi.cpp"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(int(os.environ.get('MAX_JOBS', '8')), len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
