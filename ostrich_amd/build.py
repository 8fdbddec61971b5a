"""Builds the gfx950 solver library ostrich_amd/_lib/libostrich_fb.so with hipcc.

Each translation unit compiles to an object under ostrich_amd/_build/ in parallel; the
fused kernel's instances are split over five units (fb_fused_classes.hip compiled with
OFB_BJ x OFB_RSET, and the Jacobi wide-layout set), then hipcc links the shared library.
Packed-FP32 instructions are disabled for every device pass (-Xclang -target-feature -Xclang -packed-fp32-ops).
Extra hipcc flags may be given in OFB_HIPCC_FLAGS (e.g. -DOFB_PHASE_TIMING).
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libostrich_fb.so")
# (object name, source, extra defines)
UNITS = [
    ("fb_api", "fb_api.hip", []),
    ("fb_fused", "fb_fused.hip", []),
    ("fb_plain", "fb_plain.hip", []),
    ("fb_warmstart", "fb_warmstart.hip", []),
    ("fb_mlp", "fb_mlp.hip", []),
    ("fb_reduce", "fb_reduce.hip", []),
    ("fb_trajopt", "fb_trajopt.hip", []),
    ("fb_collide", "fb_collide.hip", []),
    ("fb_log", "fb_log.hip", []),
    ("fb_fused_j0", "fb_fused_classes.hip", ["-DOFB_BJ=0", "-DOFB_RSET=0"]),
    ("fb_fused_j1", "fb_fused_classes.hip", ["-DOFB_BJ=0", "-DOFB_RSET=1"]),
    ("fb_fused_j2", "fb_fused_classes.hip", ["-DOFB_BJ=0", "-DOFB_RSET=2"]),
    ("fb_fused_b0", "fb_fused_classes.hip", ["-DOFB_BJ=1", "-DOFB_RSET=0"]),
    ("fb_fused_b1", "fb_fused_classes.hip", ["-DOFB_BJ=1", "-DOFB_RSET=1"]),
    ("fb_adj_j0", "fb_adjoint_classes.hip", ["-DOFB_BJ=0", "-DOFB_RSET=0"]),
    ("fb_adj_j1", "fb_adjoint_classes.hip", ["-DOFB_BJ=0", "-DOFB_RSET=1"]),
    ("fb_adj_b0", "fb_adjoint_classes.hip", ["-DOFB_BJ=1", "-DOFB_RSET=0"]),
    ("fb_adj_b1", "fb_adjoint_classes.hip", ["-DOFB_BJ=1", "-DOFB_RSET=1"]),
]
ARCH = os.environ.get("OFB_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _compile(unit, verbose):
    name, src, defs = unit
    obj = os.path.join(OBJ_DIR, name + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")] + defs
    cmd += shlex.split(os.environ.get("OFB_HIPCC_FLAGS", ""))
    cmd += ["-o", obj, os.path.join(CSRC, src)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return obj, r


def build(verbose=False, jobs=None, out_dir=None):
    """Builds the library into ostrich_amd/_lib (or `out_dir`, for a variant build such as
    the phase-timing one: OFB_HIPCC_FLAGS=-DOFB_PHASE_TIMING)."""
    global OBJ_DIR, LIB
    if out_dir is not None:
        out_dir = os.path.abspath(out_dir)
        OBJ_DIR, LIB = os.path.join(out_dir, "obj"), os.path.join(out_dir, "libostrich_fb.so")
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, verbose), UNITS))
    failed = [(o, r) for o, r in results if r.returncode != 0]
    for o, r in results:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l
                        and "is not a recognized feature" not in l and l.strip())
        if msg and (verbose or r.returncode != 0):
            print(msg, file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(os.path.basename(o) for o, _ in failed))
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp"] + [o for o, _ in results]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    out = sys.argv[sys.argv.index("--out") + 1] if "--out" in sys.argv else None
    build(verbose="-v" in sys.argv, out_dir=out)
