"""Compile libfbp.so (gfx950) in-tree with hipcc.

Every source is compiled to an object in csrc/_build (in parallel), then the
objects are linked.  incremental=True keeps an object newer than its source
and every header of csrc/ and include/ (for edit-compile loops); the default
recompiles everything.

Usage: python -m spikesorting_fullpursuit_amd.build [--incremental]
"""
import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libfbp.so")
SOURCES = ["kernels.hip", "runtime.hip", "clips.hip", "overlap_templates.hip", "separability.hip", "server.hip"]
ARCH = os.environ.get("FBP_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: only explicit fused multiply-adds, in the reference's order.
# -packed-fp32-ops off: no v_pk_*_f32 (scalar float32 throughout).
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-ffp-contract=off", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    return res.returncode, res.stdout, out


def build(verbose=False, incremental=False):
    os.makedirs(OBJ, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(HERE, "..", "include", "*.h"))
    newest_header = max((os.path.getmtime(h) for h in headers), default=0.0)
    jobs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header):
            continue
        jobs.append([hipcc()] + FLAGS + ["-c", "-o", obj, src])
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, min(len(jobs), 6))) as ex:
        for cmd, (rc, full, out) in zip(jobs, ex.map(_run, jobs)):
            if verbose:
                print(" ".join(cmd), flush=True)
            if rc != 0:
                raise RuntimeError("hipcc failed:\n" + full)
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-shared", "-o", tmp] + objs
    rc, full, out = _run(cmd)
    if verbose:
        print(" ".join(cmd), flush=True)
    if rc != 0:
        raise RuntimeError("hipcc link failed:\n" + full)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv[1:])
    print("built", LIB)
    sys.exit(0)
