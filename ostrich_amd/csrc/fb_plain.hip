// fb_plain.hip -- plain per-kernel execution path (filled in below).
#include "fb_solver.h"

namespace ofb {
struct PlainBuffers {};
int plain_create(ofb_solver* s) { (void)s; set_error("plain path not built"); return -2; }
void plain_destroy(ofb_solver* s) { (void)s; }
int plain_step(ofb_solver*, const ofb_state_in*, const ofb_state_out*, const ofb_control*, float, hipStream_t) {
  set_error("plain path not built");
  return -2;
}
}  // namespace ofb
