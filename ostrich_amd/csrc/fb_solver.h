// fb_solver.h -- solver handle shared by the fused and the plain execution paths.
//
// HBM layout (all world-major; W worlds, B bodies, J joints, D dofs, SC shape columns,
// K = max contacts per world, R = N_c constraint rows per world):
//   model      : copies of the ofb_model arrays (read-only during a step)
//   contacts   : packed per world in component-major blocks so that lanes that walk
//                consecutive contacts issue coalesced loads:
//                  cf[(w*CF + k)*K + c]   k in {p0x,p0y,p0z,p1x,p1y,p1z,nx,ny,nz,th0,th1}
//                  ci[(w*2 + k)*K + c]    k in {shape0, shape1} (world-local columns)
//   lambda     : [W*R] picked iterate (reference row layout, core/engine_dims.py:68-107)
//   stats      : [W] Newton iterations, PCR iterations, merit, picked index, flags
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ostrich_fb.h"

namespace ofb {

enum { CF_P0X = 0, CF_P0Y, CF_P0Z, CF_P1X, CF_P1Y, CF_P1Z, CF_NX, CF_NY, CF_NZ, CF_TH0, CF_TH1, CF };

enum {
  FLAG_NR_CAP = 1,
  FLAG_CONTACT_OVERFLOW = 2,
  FLAG_FACTOR_FAILURE = 4,
  FLAG_PAIR_OVERFLOW = 8,
  FLAG_MU_GRAD_MISSING = 16,  // adjoint: a friction row's mu derivative has no closed form
  FLAG_LEARNED_SEED = 32,     // the step started from the learned seed (lower initial residual)
};

constexpr int kMaxJoints = 64;      // joints per world supported by the fused path
constexpr int kMaxPairMembers = 64; // per_body_pair_preconditioner.py:59

struct DeviceModel {
  int W, B, J, D, S, G, SC;
  float gravity[3];
  float *body_mass, *body_inv_mass, *body_inertia, *body_inv_inertia, *body_com;
  int32_t *joint_type, *joint_parent, *joint_child, *joint_qd_start, *joint_enabled;
  float *joint_X_p, *joint_X_c, *joint_compliance;
  float *joint_axis;
  int32_t* joint_dof_mode;
  float *joint_target_ke, *joint_target_kd;
  int32_t* shape_body;
  float *shape_mu, *shape_friction_axis, *shape_mu_perp, *shape_mu_stiction_scale, *shape_v_stribeck,
      *shape_stribeck_lateral_only;
  int32_t* shape_world;
  int flat_shapes;  // G + W*S
};

// Row layout derived from world 0 (core/model.py:83-168)
struct RowLayout {
  int Nj, Nctrl, K, Nc, off_ctrl, off_n, off_f;
  int32_t joint_off[kMaxJoints];  // first joint row of each joint
  int32_t ctrl_off[kMaxJoints];   // control row of each joint (-1 = none)
};

struct PlainBuffers;  // fb_plain.hip

}  // namespace ofb

struct ofb_solver {
  ofb_config cfg;
  int32_t path;
  ofb::DeviceModel m;
  ofb::RowLayout L;
  // packed contacts
  int32_t* ccount;   // [W] raw per-world count (may exceed K, core/contacts.py:73-76)
  float* cf;         // [W*CF*K]
  int32_t* ci;       // [W*2*K]
  // contact packing scratch (stable partition by world)
  int32_t rigid_contact_cap;
  int32_t* pk_keys;     // [rigid_contact_cap] world of each flat contact (valid prefix)
  int32_t* pk_members;  // [W*K] flat index of each world's contacts, in reservation order
  // outputs
  int32_t* topo;     // joint of each joint row [Nj], row within joint [Nj], joint of each control row [Nctrl]
  float* lambda;     // [W*Nc]
  float* lambda_prev;// [W*Nc] lambda_prev_iter as the last step left it (fused path)
  float* lambda_seed;// [W*Nc]
  int32_t *st_nr, *st_pcr, *st_best, *st_flags;
  float* st_merit;
  ofb::PlainBuffers* plain;
  // largest number of waves per workgroup of a fused class launch (ofb_set_fused_launch;
  // the Python autotuner picks it)
  int fused_worlds_per_block;
  int num_simds;  // SIMDs of the device (4 per CU)
  // caller buffers of the per-Newton-iteration history (ofb_set_iteration_history)
  float* hist_merit;
  int32_t* hist_pcr;
  int hist_cap;
  // per-step bucketing of worlds by active-row class (fused path)
  int32_t* cls_counts;
  int32_t* cls_lists;
  int cls_wpb[16];   // worlds per workgroup of each class launch (LDS-limited)
  // the one-world-per-wave classes run on side streams beside the paired class (the
  // dense class, the R = 2 class and the larger classes one after another)
  hipStream_t side_stream[8];
  hipEvent_t ev_fork, ev_join[8];
  int side_ready;  // side streams and join events created
  // cross-step contact warm start: snapshot of the previous step (fb_warmstart.hip)
  int32_t *ws_count, *ws_b0, *ws_b1;
  float *ws_p, *ws_n, *ws_ln, *ws_lt;
  float *ws_feat, *ws_fscale, *ws_y;  // learned warm start: features [W*K*12], scale [W*K], output [W*K*3]
  int32_t* ws_off;                     // [W+1] offsets of the worlds' contacts in the packed MLP rows
  float* lambda_seed2;                // [W*Nc] learned seed
  const struct ofb_mlp* ws_net;       // attached network (not owned)
  // merit line search candidates (core/engine_data.py:15-39), device and host copies
  float* ls_steps;
  int32_t ls_count;
  float ls_host[512];
  // events around the solve of the last step, and around its paired-class kernel
  hipEvent_t ev_solve0, ev_solve1, ev_kernel0, ev_kernel1;
  int events_ready;
  int kernel_timed;   // the last step recorded the kernel events
  // optional per-class timing of the fused path (ofb_set_class_timing): events around
  // each class launch on the stream it runs on; bit c of cls_timed: class c was timed
  int cls_timing;
  // the fused path writes lambda_prev_iter (the adjoint's input) only when set
  int keep_lambda_prev;
  int cls_timed;
  hipEvent_t ev_cls0[16], ev_cls1[16];
  int cls_events_ready;
  int solve_timed;    // the last step recorded the solve events
};

namespace ofb {
int fused_step(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl, float dt,
               hipStream_t stream);
int fused_classes_alloc(ofb_solver* s);
int fused_set_launch(ofb_solver* s, int waves_per_block, int pair_small_worlds);
int fused_set_layout(ofb_solver* s, int large_world_slots);
int fused_class_times(const ofb_solver* s, int capacity, int32_t* rows_cap, int32_t* worlds, float* ms,
                      int32_t* count);
int fused_class_shapes(const ofb_solver* s, int capacity, int32_t* slots, int32_t* lanes, int32_t* max_rows,
                       int32_t* count);
void fused_classes_free(ofb_solver* s);
int fused_adjoint(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl,
                  const float* lambda, const float* lambda_prev, const ofb_adjoint_in* gin,
                  const ofb_adjoint_out* gout, float dt, hipStream_t stream);
int plain_create(ofb_solver* s);
void plain_destroy(ofb_solver* s);
int plain_step(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl, float dt,
               hipStream_t stream);
int reduce_contacts(ofb_solver* s, const float* body_q, hipStream_t stream);
int warm_start_apply(ofb_solver* s, const ofb_state_in* in, float dt, hipStream_t stream);
int warm_start_snapshot(ofb_solver* s, const ofb_state_in* in, float dt, hipStream_t stream);
int warm_start_learned_seed(ofb_solver* s, const float* y, hipStream_t stream);
// network forward over the W * K contact slots, valid ones only (c < count[w]), through
// a packed row list (off: [W + 1] scratch)
int mlp_forward_packed(const ofb_mlp* net, const float* x, const int32_t* count, int W, int K, int32_t* off,
                       float* y, hipStream_t st);
void set_error(const char* fmt, ...);
}  // namespace ofb
