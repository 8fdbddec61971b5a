"""Per-phase cycle counts of the fused Newton loop (a profiling aid).

Build a phase-timing variant of the library beside the product one, then run:

    OFB_HIPCC_FLAGS=-DOFB_PHASE_TIMING python -m ostrich_amd.build --out ostrich_amd/_phase
    python -m ostrich_amd.phase_timing --lib ostrich_amd/_phase/libostrich_fb.so

The variant's fused kernels (Jacobi instances of the paired and dense classes and R <= 5)
add clock64() deltas per phase of every Newton iteration into device counters, summed
over worlds; this prints cycles per Newton iteration per world for the skid-steer
workload.  Cycles are wall cycles of a wave, so they include the issue slots other waves
of the SIMD take.
"""
import argparse
import ctypes as C
import os
import sys

import numpy as np

PHASES = ("friction lag + dynamics", "assembly", "RHS + Jacobi", "PCR", "update", "residual", "bookkeeping",
          "dense Schur build")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", required=True, help="library built with -DOFB_PHASE_TIMING")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--variant", default="fused:jacobi:prev")
    args = ap.parse_args()
    from . import _lib
    _lib.LIB_PATH = os.path.abspath(args.lib)
    lib = _lib.load()
    if not hasattr(lib, "ofb_debug_phase_cycles"):
        sys.exit(f"{args.lib} was not built with -DOFB_PHASE_TIMING")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    from . import collide, scenes
    model, state, control = scenes.skid_steer(args.worlds, seed=1)
    col = collide.HipCollider(model, device="cuda:0", gap=0.02, max_contacts_per_world=128)
    f = lib.ofb_debug_phase_cycles
    f.argtypes = [C.c_void_p, C.c_int]
    buf = (C.c_ulonglong * 8)()
    bench.run_path(args.variant, model, state, control, 1, 1, "cuda:0", col)  # load and warm the code
    f(buf, 1)
    r = bench.run_path(args.variant, model, state, control, args.steps, 0, "cuda:0", col)
    f(buf, 1)
    ph = np.array(list(buf), dtype=np.float64) / (args.steps * args.worlds) / r["nr_mean"]
    print(f"{args.variant}: {r['nr_mean']:.2f} Newton and {r['pcr_per_step_mean']:.1f} PCR iterations per "
          f"world-step; cycles per Newton iteration per world:")
    for name, v in zip(PHASES, ph):
        print(f"  {name:24s} {v:9.0f}")


if __name__ == "__main__":
    main()
