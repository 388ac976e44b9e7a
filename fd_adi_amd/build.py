"""Build libfdadi.so (HIP, gfx950) in-tree.

    python fd_adi_amd/build.py      (run as a script: importing the package
                                     would load the library being rebuilt)

The library is a single shared object with a C ABI (include/fdadi.h); the
Python layer loads it with ctypes.  Sources are compiled separately and
linked, so a change to one file recompiles only that file.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfdadi.so")
SOURCES = ["array.hip", "tri.hip", "cross.hip", "heston.hip", "mlp.hip", "cf.hip", "serve.hip"]
ARCH = os.environ.get("FDADI_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def build(verbose=False, jobs=None):
    """Compile every HIP source for gfx950 and link libfdadi.so."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs = []
    objs = []
    for s in srcs:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc] + HIPCC_FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                          stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # The host pass does not know the packed-fp32 feature; that warning is expected.
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l
                 and "is not a recognized feature" not in l and l.strip()]
        if p.returncode != 0:
            failed.append(s)
            sys.stderr.write(text)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for: %s" % ", ".join(failed))
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
