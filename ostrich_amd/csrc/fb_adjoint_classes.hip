// fb_adjoint_classes.hip -- instances of the adjoint kernel for one preconditioner and
// one set of row classes (OFB_BJ, OFB_RSET as in fb_fused_classes.hip), compiled once
// per pair so that the instances build in parallel.

#include "fb_adjoint_kernel.h"

#if !defined(OFB_BJ) || !defined(OFB_RSET)
#error "compile with -DOFB_BJ=0|1 -DOFB_RSET=0|1"
#endif

#define OFB_CAT_(a, b, c, d) a##b##c##d
#define OFB_CAT(a, b, c, d) OFB_CAT_(a, b, c, d)

namespace ofb {

int OFB_CAT(adjoint_class_op_bj, OFB_BJ, _set, OFB_RSET)(int R, int op, const AdjParams& P, int blocks, int threads,
                                                       size_t lds, hipStream_t st) {
  switch (R) {
#define CASE(RR) \
  case RR: return adjoint_class_op<RR, OFB_BJ != 0>(op, P, blocks, threads, lds, st);
#if OFB_RSET == 0
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#else
    CASE(6) CASE(7) CASE(12) CASE(16)
#endif
#undef CASE
  }
  return -1;
}

}  // namespace ofb
