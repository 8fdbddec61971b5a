// fb_fused_classes.hip -- instances of the fused kernel for one preconditioner and one
// set of row classes.  Compiled once per (OFB_BJ, OFB_RSET) pair so that the instances
// build in parallel; OFB_BJ = 1 selects per-body-pair block-Jacobi, OFB_RSET = 0 the
// dense classes (code 0: two worlds per wave, code 1: one), R = 2 and, for block-Jacobi,
// R = 3..5; OFB_RSET = 1 the block-Jacobi classes R = 6, 7, 12 and 16, or the Jacobi
// worlds of 2, 4 and 8 waves (Jacobi solvers run worlds of more than 128 rows on several
// waves, block-Jacobi ones on one wave with more row slots per lane); OFB_RSET = 2 (Jacobi
// only) the worlds of 2 waves with three or four row slots per lane and of 4 waves with
// four (the wide layout).

#include "fb_fused_kernel.h"

#if !defined(OFB_BJ) || !defined(OFB_RSET)
#error "compile with -DOFB_BJ=0|1 -DOFB_RSET=0|1|2"
#endif

#define OFB_CAT_(a, b, c, d) a##b##c##d
#define OFB_CAT(a, b, c, d) OFB_CAT_(a, b, c, d)

namespace ofb {

int OFB_CAT(fused_class_op_bj, OFB_BJ, _set, OFB_RSET)(int code, int op, const FusedParams& P, int blocks, int threads,
                                                     size_t lds, hipStream_t st) {
  switch (code) {
#define CASE(RR) \
  case RR: return fused_class_op<RR, OFB_BJ != 0, 64, false>(op, P, blocks, threads, lds, st);
#if OFB_RSET == 0
    case 0: return fused_class_op<1, OFB_BJ != 0, 32, true>(op, P, blocks, threads, lds, st);
    case 1: return fused_class_op<1, OFB_BJ != 0, 64, true>(op, P, blocks, threads, lds, st);
    CASE(2)
#if OFB_BJ != 0
    CASE(3) CASE(4) CASE(5)
#endif
#elif OFB_RSET == 1 && OFB_BJ != 0
    CASE(6) CASE(7) CASE(12) CASE(16)
#elif OFB_RSET == 1
    case multiwave_code(2, 2): return fused_class_op<2, false, 128, false>(op, P, blocks, threads, lds, st);
    case multiwave_code(4, 2): return fused_class_op<2, false, 256, false>(op, P, blocks, threads, lds, st);
    case multiwave_code(8, 2): return fused_class_op<2, false, 512, false>(op, P, blocks, threads, lds, st);
#elif OFB_BJ == 0
    case multiwave_code(2, 3): return fused_class_op<3, false, 128, false>(op, P, blocks, threads, lds, st);
    case multiwave_code(2, 4): return fused_class_op<4, false, 128, false>(op, P, blocks, threads, lds, st);
    case multiwave_code(4, 4): return fused_class_op<4, false, 256, false>(op, P, blocks, threads, lds, st);
#endif
#undef CASE
  }
  return -1;
}

// Profiling build only: the phase counters [17][10] of the instances of this unit (each
// translation unit holds its own copy).
#if defined(OFB_PHASE_TIMING)
extern "C" int OFB_CAT(ofb_debug_phase_cycles_bj, OFB_BJ, _set, OFB_RSET)(unsigned long long* out, int reset) {
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase_cycles), sizeof(unsigned long long) * 170) != hipSuccess) return -4;
  if (reset) {
    static const unsigned long long z[170] = {};
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_phase_cycles), z, sizeof(z)) != hipSuccess) return -4;
  }
  return 0;
}
#endif

}  // namespace ofb
