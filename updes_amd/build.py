"""Build the HIP library ``updes_amd/_lib/librbfx.so`` for gfx950.

Usage: ``python -m updes_amd.build`` (or ``updes_amd.build.build()``).
Every translation unit under ``updes_amd/csrc`` is compiled with hipcc into
``_build/`` (object files, not shipped) and linked into one shared library
that exports the C ABI declared in ``include/rbfx.h``.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(ROOT, "_build", "rbfx")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "librbfx.so")
ARCH = os.environ.get("RBFX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.cpp", "host_logic.cpp", "assembly.hip", "eval.hip", "lu.hip", "trsv.hip", "krylov.hip", "cardinal.hip"]

# Packed FP32 instructions are disabled for the device code (the host pass
# warns that it does not know the feature; that warning is harmless).
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build updes_amd)")


def _compile(src, verbose):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [_hipcc()] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = (res.stdout or "") + (res.stderr or "")
    out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError("compiling %s failed:\n%s" % (src, out))
    return obj, out


def build(verbose=False, jobs=None):
    """Compile every source and link librbfx.so.  Always rebuilds."""
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4) // 2))
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = {ex.submit(_compile, s, verbose): s for s in SOURCES}
        for f in cf.as_completed(futs):
            obj, out = f.result()
            if out and verbose:
                print(out)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + sorted(objs)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("linking librbfx.so failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
