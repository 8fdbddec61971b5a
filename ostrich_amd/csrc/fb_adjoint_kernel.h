// fb_adjoint_kernel.h -- adjoint of one solver step on gfx950 (OstrichEngine.step_backward,
// src/ostrich/core/base_engine.py:528-743), one wavefront per world like the forward
// kernel and built from the same pieces: row assembly (fb_rows.h), the LDS matrix-free
// operator, the Jacobi / block-Jacobi preconditioners and the register PCR.
//
// At the state the forward step left (q+, u+, lambda, lambda_prev_iter) the kernel
//   1. rebuilds J, C and the residual R = [h_d | h_c] (compute_linear_system, :533);
//   2. forms the driving force f = -(g_u + dt G(q+)^T g_q) - [0; dt (R(q+)(-com)) x g_p]
//      and w_u0 = M^-1 f per body (compute_body_adjoint_init_kernel,
//      adjoint/adjoint_utils.py:12-79);
//   3. solves (J M^-1 J^T + C + reg + gamma) w_lambda = J w_u0 with the forward's PCR
//      (compute_adjoint_rhs_kernel :82-117, base_engine.py:633-644);
//   4. sets w_u = w_u0 - M^-1 J^T w_lambda (subtract_constraint_feedback_kernel :120-169);
//   5. returns w^T dR/dxi for the step's inputs, which the reference obtains by replaying
//      its residual kernels on a Warp tape seeded with (w_u, w_lambda)
//      (base_engine.py:664-712): in closed form for u- (-M(q+) w_u), the wrench (-dt w_u),
//      the joint targets and gains (control_target_grad_kernel, residual.py:201-257) and mu
//      of the isotropic friction cone; by central differences of the on-chip residual in
//      each of the 7 pose coordinates of every body for q+ and q-, so every row type is
//      covered; then the explicit carry dq+/dq- (body_pose_prev_grad_kernel,
//      residual.py:260-339) of g_q + w^T dR/dq+.
//
// The central differences are most of the kernel's work (2 x 14 B residual evaluations
// per world).  With P.split > 1 a workgroup holds one world in `split` wavefronts: wave 0
// runs steps 1-4 and publishes lambda and w_lambda, every wave copies wave 0's body
// block and lambda_prev into its own LDS region and evaluates every split-th pose
// coordinate.  Each coordinate is computed by the same code on the same data, so the
// result does not depend on the split.

#pragma once

#include "fb_fused_kernel.h"

namespace ofb {

struct AdjParams {
  // names read by the row functions (fb_rows.h) match FusedParams
  DeviceModel m;
  RowLayout L;
  ofb_config cfg;
  float dt;
  const float* target_pos;
  const float* target_vel;
  const int32_t* ccount;
  const float* cf;
  const int32_t* ci;
  const int32_t* topo;
  // state of the step
  const float* q_prev;   // [W*B*7] step start
  const float* qd_prev;  // [W*B*6]
  const float* body_f;   // [W*B*6]
  const float* q;        // [W*B*7] after the step
  const float* qd;       // [W*B*6]
  const float* lam;      // [W*Nc] reference layout
  const float* lam_prev; // [W*Nc]
  // gradients
  const float* g_q;      // [W*B*7] dL/dq+
  const float* g_qd;     // [W*B*6] dL/du+
  float* g_q_prev;       // outputs, each may be null
  float* g_qd_prev;
  float* g_f;
  float* g_tpos;
  float* g_tvel;
  float* g_ke;
  float* g_kd;
  float* g_mu;
  float* w_lam;
  float* w_u;
  int32_t* st_pcr;
  int32_t* st_flags;
  int lds_per_wave;
  const int32_t* wlist;
  const int32_t* wcount;
  int rows_cap;
  // waves per world (1..8): wave 0 of a world runs the solve; with more, every wave
  // takes a share of the pose central differences on its own LDS copy of the world
  int split;
  int xfer_off;  // LDS offset (floats, wave 0's region) of lambda / w_lambda for the others
  int fd_off;    // LDS offset of the row-local difference table (< 0: full-residual differences)
  int rs_off;    // LDS offset (floats, each wave's region) of the lane-private body sums (< 0: none)
};

// Words of the row-local difference table: per row slot and side (body 0 / body 1) the 14
// pose-coordinate differences of the row's term, the two body indices, and per body the
// 7 q+ differences of its dynamics term.
__host__ __device__ inline int adj_fd_words(int rows_cap, int B) { return 30 * rows_cap + 7 * B; }

// LDS floats of the adjoint's own scratch per wave: pose pull-back [14 B], control
// gradients [4 D] (target_pos, target_vel, ke, kd), mu gradients [S + G]
__host__ __device__ inline int adj_words(int B, int D, int SC) { return 14 * B + 4 * D + SC; }

// The adjoint keeps w_u of each body in the slot of the forward's best-candidate pose
// (BBQ, unused here): BVV is the Schur operator's scratch.
constexpr int BWU = BBQ;

// The weighted residual splits into per-row terms, each a function of the row's two bodies
// only, plus per-body dynamics terms:
//   s = sum_rows [w_lambda h_c + (-lambda dt)(w_u(b0) . J0 + w_u(b1) . J1)]
//       + sum_bodies w_u . (M(q+)(u - u-) - dt (f + m g + gyro)),
// the J^T lambda part of h_d being the rows' scatter.  term of one row at poses read by pr:
template <class PR>
__device__ __forceinline__ float adjoint_row_term(const AdjParams& P, float* sm, const float* lamprev, int w, int i,
                                                  int n_c, float lam, float wl, const PR& pr, int* b0, int* b1) {
  RowOut o = assemble_row(P, sm, lamprev, w, i, n_c, lam, pr, false);
  *b0 = o.b0;
  *b1 = o.b1;
  float t = wl * o.hc;
  if (o.active) {
    float d = 0.f;
    if (o.b0 >= 0) d += sdot(o.J0, ldsv(sm + o.b0 * BS + BWU));
    if (o.b1 >= 0) d += sdot(o.J1, ldsv(sm + o.b1 * BS + BWU));
    t += d * (-lam * P.dt);
  }
  return t;
}

// weighted residual s = sum_rows w_lambda h_c + sum_bodies w_u . h_d at the pose in LDS
template <int R>
__device__ __forceinline__ float adjoint_weighted_residual(const AdjParams& P, float* sm, const float* lamprev,
                                                           int lane, int w, int n_r, int n_c, const float (&lam)[R],
                                                           const float (&wl)[R], const BodySum<WAVE, false>& S) {
  const int B = P.m.B;
  dynamics_bodies(P, sm, lane, false);
  wsync();
  S.begin(BHD, false);
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    if (i < n_r) {
      RowOut o = assemble_row(P, sm, lamprev, w, i, n_c, lam[k], LdsPoses(), false);
      s += wl[k] * o.hc;
      S.add(BHD, i, o.active, o.b0, o.b1, o.J0 * (-lam[k] * P.dt), o.J1 * (-lam[k] * P.dt));
    }
  }
  S.end(BHD, false, n_r);
  wsync();
  for (int e = lane; e < 6 * B; e += WAVE) {
    const float* bb = sm + (e / 6) * BS;
    s += bb[BWU + e % 6] * bb[BHD + e % 6];
  }
  return wave_sum(s);
}

// adjoint.friction_linearization = "frozen" (freeze_contact_mode_kernel and its soft
// variant, adjoint/adjoint_friction.py:20-158, launched in base_engine.py:547-590): the
// friction rows of contact c take the compliance of the contact's mode at the converged
// impulses -- sticking (|lambda_t| <= 0.9 mu lambda_n): compliance.joint; sliding: 100 --
// or, with soft blending, a sigmoid of (ratio - 0.9) / temperature between the two; a
// contact without normal impulse (< 1e-6) drops its friction rows from the system.
template <class PT>
__device__ __forceinline__ RowOut frozen_friction(const PT& P, int w, int c, RowOut o) {
  const RowLayout& L = P.L;
  const float* lw = P.lam + (size_t)w * L.Nc;
  const float ln = lw[L.off_n + c], t1 = lw[L.off_f + 2 * c], t2 = lw[L.off_f + 2 * c + 1];
  if (ln < 1e-6f) {
    o.active = false;
    return o;
  }
  const ContactGeo g = load_contact(P, w, c);
  const int SC = P.m.SC;
  const float mu = 0.5f * (P.m.shape_mu[w * SC + g.s0] + P.m.shape_mu[w * SC + g.s1]);
  const float ratio = sqrtf(t1 * t1 + t2 * t2 + 1e-10f) / (mu * ln + 1e-10f);
  const float stick = P.cfg.compliance_joint, slide = 100.f;
  if (P.cfg.adjoint_soft_blending) {
    const float e = fminf(fmaxf(-(ratio - 0.9f) / fmaxf(P.cfg.adjoint_soft_temperature, 1e-8f), -20.f), 20.f);
    const float sg = 1.f / (1.f + expf(e));
    o.C = (1.f - sg) * stick + sg * slide;
  } else {
    o.C = ratio > 0.9f ? slide : stick;
  }
  return o;
}

template <int R, bool BJ>
__global__ void __launch_bounds__(256) adjoint_step_kernel(AdjParams P) {
  extern __shared__ __align__(16) float lds[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = threadIdx.x / WAVE;
  const int split = P.split;  // > 1 only with one world per workgroup
  const int sub = wave % split;
  const int idx = blockIdx.x * (blockDim.x / WAVE / split) + wave / split;
  if (idx >= *P.wcount) return;  // uniform over a world's waves
  const int w = P.wlist[idx];
  float* sm = lds + wave * P.lds_per_wave;
  float* sm0 = lds + (wave - sub) * P.lds_per_wave;  // wave 0 of this world
  const int B = P.m.B, D = P.m.D, SC = P.m.SC;
  const RowLayout& L = P.L;
  const float dt = P.dt;
  float* lamprev = sm + B * BS;
  float* gpose = sm0 + B * BS + L.Nc + 1;  // [B][14]: w^T dR/dq+ (7), w^T dR/dq- (7)
  float* gctl = gpose + 14 * B;            // [4][D]
  float* gmu = gctl + 4 * D;               // [SC]
  const BJView V = bj_view(gmu + SC, P.rows_cap);
  BJRegs<R> G;
  int flags = 0;
  const int n_c = min(P.ccount[w], L.K);
  const int n_r = L.Nj + L.Nctrl + 3 * n_c;
  // row -> body sums: lane-private copies when the launch gave each wave a region for
  // them, else LDS atomics
  const BodySum<WAVE, false> S{sm, sm + (P.rs_off >= 0 ? P.rs_off : 0), lane, B, P.rs_off >= 0 && rs_enabled(B),
                                 nullptr};
  float lam[R], x[R];
  int pcr_it = 0;
  if (sub == 0) {

  // ---- state: q+ / u+ in the iterate slots, q- / u- in the step-start slots ----
  for (int b = lane; b < B; b += WAVE) {
    float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
#pragma unroll
    for (int i = 0; i < 7; ++i) { bb[BQ + i] = P.q[k * 7 + i]; bb[BQP + i] = P.q_prev[k * 7 + i]; }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      bb[BU + i] = P.qd[k * 6 + i];
      bb[BUP + i] = P.qd_prev[k * 6 + i];
      bb[BF + i] = P.body_f[k * 6 + i];
    }
    bb[BM] = P.m.body_mass[k];
    bb[BMINV] = P.m.body_inv_mass[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) bb[BCOM + i] = P.m.body_com[k * 3 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) { bb[BIB + i] = P.m.body_inertia[k * 9 + i]; bb[BIIB + i] = P.m.body_inv_inertia[k * 9 + i]; }
  }
  for (int e = lane; e < 14 * B + 4 * D + SC; e += WAVE) gpose[e] = 0.f;
  float Cr[R], Pi[R], rhs[R];
  sv J0[R], J1[R];
  int b0[R], b1[R];
  bool act[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    lam[k] = 0.f;
    act[k] = false; b0[k] = -1; b1[k] = -1;
    if (i < n_r) {
      size_t o = (size_t)w * L.Nc + original_row(L, i, n_c);
      lam[k] = P.lam[o];
      lamprev[i] = P.lam_prev[o];
    }
  }
  wsync();

  // ---- 1. J, C at the converged state ----
  dynamics_bodies(P, sm, lane, true);
  wsync();
  S.begin(BHD, false);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    Cr[k] = 0.f;
    J0[k] = svzero(); J1[k] = svzero();
    if (i < n_r) {
      RowOut o = assemble_row(P, sm, lamprev, w, i, n_c, lam[k], LdsPoses(), false);
      S.add(BHD, i, o.active, o.b0, o.b1, o.J0 * (-lam[k] * dt), o.J1 * (-lam[k] * dt));
      J0[k] = o.J0; J1[k] = o.J1; b0[k] = o.b0; b1[k] = o.b1;
      if (!o.active && i >= L.Nj + L.Nctrl) lam[k] = 0.f;
      if (P.cfg.adjoint_friction_frozen && i >= L.Nj + L.Nctrl + n_c && o.active)
        o = frozen_friction(P, w, (i - L.Nj - L.Nctrl - n_c) >> 1, o);
      act[k] = o.active;
      Cr[k] = o.C + P.cfg.regularization + (o.active ? P.cfg.adjoint_regularization : 0.f);
    }
  }
  wsync();

  // ---- 2. w_u0 = M^-1 f_drive (Gt_matvec, kinematic_mapping.py:208-279) ----
  for (int b = lane; b < B; b += WAVE) {
    float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
    float gq[7], gu[6];
#pragma unroll
    for (int i = 0; i < 7; ++i) gq[i] = P.g_q[k * 7 + i];
#pragma unroll
    for (int i = 0; i < 6; ++i) gu[i] = P.g_qd[k * 6 + i];
    const float t2 = bb[BQ + 3], t3 = bb[BQ + 4], t4 = bb[BQ + 5], t1 = bb[BQ + 6];
    const float d2 = gq[3], d3 = gq[4], d4 = gq[5], d1 = gq[6];
    const float hs = dt * 0.5f;
    v3 fv = mk(-(gu[0] + gq[0] * dt), -(gu[1] + gq[1] * dt), -(gu[2] + gq[2] * dt));
    v3 fw = mk(-(gu[3] + (-t2 * d1 + t1 * d2 - t4 * d3 + t3 * d4) * hs),
               -(gu[4] + (-t3 * d1 + t4 * d2 + t1 * d3 - t2 * d4) * hs),
               -(gu[5] + (-t4 * d1 - t3 * d2 + t2 * d3 + t1 * d4) * hs));
    qt rq = qt{t2, t3, t4, t1};
    v3 lever = qrot(rq, ld3(bb + BCOM) * -1.f);
    fw = fw - cross(lever, mk(gq[0], gq[1], gq[2])) * dt;
    stsv(bb + BWU, apply_minv(bb, sv{fv, fw}));
  }
  wsync();

  // ---- 3. A w_lambda = J w_u0 ----
  float bsq = 0.f;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    rhs[k] = 0.f; Pi[k] = 0.f;
    if (act[k]) {
      float s = 0.f, d = 0.f;
      if (b0[k] >= 0) {
        const float* bb = sm + b0[k] * BS;
        s += sdot(J0[k], ldsv(bb + BWU));
        d += sdot(J0[k], apply_minv(bb, J0[k]));
      }
      if (b1[k] >= 0) {
        const float* bb = sm + b1[k] * BS;
        s += sdot(J1[k], ldsv(bb + BWU));
        d += sdot(J1[k], apply_minv(bb, J1[k]));
      }
      rhs[k] = s;
      Pi[k] = 1.f / (d + Cr[k] + 1e-6f);
    }
    bsq += rhs[k] * rhs[k];
  }
  bsq = wave_sum(bsq);
  if (BJ) {
    bj_structure<R>(V, lane, n_r, B, act, b0, b1, G, flags);
    bj_factor<R>(V, sm, lane, b0, b1, J0, J1, Cr, G, flags);
  }
  const float tol_sq = P.cfg.lin_tol >= 0.f ? P.cfg.lin_tol * P.cfg.lin_tol : -1.f;
  const float latol_sq = P.cfg.lin_atol >= 0.f ? P.cfg.lin_atol * P.cfg.lin_atol : -1.f;
  WorldSum<WAVE> WS{nullptr, 0};  // one world per wave: reductions stay inside the wave
  pcr_it = pcr_solve<R, BJ, WAVE, false>(sm, lane, B, n_r, S, V, G, act, b0, b1, J0, J1, Cr, Pi, rhs, bsq, tol_sq,
                                         latol_sq, P.cfg.lin_max_iters, x, WS);

  // ---- 4. w_u = w_u0 - M^-1 J^T w_lambda ----
  S.begin(BACC, true);
#pragma unroll
  for (int k = 0; k < R; ++k) S.add(BACC, lane + WAVE * k, act[k], b0[k], b1[k], J0[k] * x[k], J1[k] * x[k]);
  S.end(BACC, true, n_r);
  wsync();
  for (int b = lane; b < B; b += WAVE) {
    float* bb = sm + b * BS;
    sv a = apply_minv(bb, ldsv(bb + BACC));
    sv u0 = ldsv(bb + BWU);
    stsv(bb + BWU, sv{u0.a - a.a, u0.b - a.b});
  }
  wsync();

  // ---- 5a. closed-form parts of w^T dR/dxi ----
  for (int b = lane; b < B; b += WAVE) {
    const float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
    sv wu = ldsv(bb + BWU);
    // h_d = M(q+)(u - u-) - dt (f + m g + gyro): dh_d/du- = -M(q+), dh_d/df = -dt
    m3 Iw = rot_inertia(qt{bb[BQ + 3], bb[BQ + 4], bb[BQ + 5], bb[BQ + 6]}, ldm3(bb + BIB));
    v3 gw = mv(Iw, wu.b);
    float gup[6] = {-bb[BM] * wu.a.x, -bb[BM] * wu.a.y, -bb[BM] * wu.a.z, -gw.x, -gw.y, -gw.z};
    float wuv[6] = {wu.a.x, wu.a.y, wu.a.z, wu.b.x, wu.b.y, wu.b.z};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      if (P.g_qd_prev) P.g_qd_prev[k * 6 + i] = gup[i];
      if (P.g_f) P.g_f[k * 6 + i] = -dt * wuv[i];
      if (P.w_u) P.w_u[k * 6 + i] = wuv[i];
    }
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    if (i >= n_r || !act[k]) continue;
    if (i >= L.Nj && i < L.Nj + L.Nctrl) {  // control_target_grad_kernel (residual.py:201-257)
      const int J = P.m.J;
      int j = P.topo[2 * L.Nj + (i - L.Nj)];
      int qs = P.m.joint_qd_start[w * J + j];
      int mode = P.m.joint_dof_mode[w * D + qs];
      float ke = P.m.joint_target_ke[w * D + qs], kd = P.m.joint_target_kd[w * D + qs];
      if (mode == 1) {
        atomicAdd(&gctl[qs], x[k] * (-1.f / dt));
        float den = dt * dt * ke + dt * kd;
        if (den > 1e-6f) {
          float inv2 = 1.f / (den * den);
          atomicAdd(&gctl[2 * D + qs], x[k] * (-dt * dt * inv2) * lam[k] * dt);
          atomicAdd(&gctl[3 * D + qs], x[k] * (-dt * inv2) * lam[k] * dt);
        }
      } else if (mode == 2) {
        atomicAdd(&gctl[D + qs], -x[k]);
        float den = dt * ke;
        if (den > 1e-6f) atomicAdd(&gctl[2 * D + qs], x[k] * (-dt / (den * den)) * lam[k] * dt);
      }
    } else if (i >= L.Nj + L.Nctrl + n_c && !P.cfg.adjoint_friction_frozen) {  // friction: h = v_t + w(mu) lambda
      int f = i - L.Nj - L.Nctrl - n_c;
      float dw = 0.f;
      (void)friction_row(P, sm, lamprev, w, f >> 1, f & 1, n_c, lam[k], &dw);
      float contrib = x[k] * lam[k];
      if (dw != dw) {
        if (contrib != 0.f) flags |= FLAG_MU_GRAD_MISSING;
      } else if (dw != 0.f) {
        ContactGeo g = load_contact(P, w, f >> 1);
        if (g.s0 >= 0) atomicAdd(&gmu[g.s0], 0.5f * dw * contrib);
        if (g.s1 >= 0) atomicAdd(&gmu[g.s1], 0.5f * dw * contrib);
      }
    }
  }
  wsync();

  if (split > 1) {  // publish lambda and w_lambda for the other waves of this world
    float* xl = sm + P.xfer_off;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + WAVE * k;
      if (i < n_r) { xl[i] = lam[k]; xl[P.rows_cap + i] = x[k]; }
    }
  }
  }  // sub == 0
  if (split > 1) {
    __syncthreads();
    if (sub > 0) {
      for (int e = lane; e < B * BS + L.Nc + 1; e += WAVE) sm[e] = sm0[e];
      const float* xl = sm0 + P.xfer_off;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        int i = lane + WAVE * k;
        lam[k] = i < n_r ? xl[i] : 0.f;
        x[k] = i < n_r ? xl[P.rows_cap + i] : 0.f;
      }
      wsync();
    }
  }

  // ---- 5b. w^T dR/dq+ and w^T dR/dq- by central differences ----
  // with frozen friction the friction rows are not seeded (base_engine.py:704-711): their
  // own residual term drops out, their share of h_d stays
  float ws_[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int i = lane + WAVE * k;
    ws_[k] = (P.cfg.adjoint_friction_frozen && i >= L.Nj + L.Nctrl + n_c) ? 0.f : x[k];
  }
  if (P.cfg.adjoint_pose_vjp && P.fd_off >= 0) {
    // row-local: each lane differences its own rows' terms in the 14 coordinates of each
    // of the row's bodies (the shifted pose is read privately, the LDS blocks stay
    // unperturbed); wave 0 then sums them per body in row order
    const float h = P.cfg.adjoint_pose_eps;
    float* T = sm0 + P.fd_off;
    int* tb = (int*)(T + 28 * P.rows_cap);
    float* Td = T + 30 * P.rows_cap;
    if constexpr (R == 1) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int i = lane + WAVE * k;
        if (i >= n_r) continue;
        // item 0 evaluates the row unshifted to learn its bodies; item 1 + 2 (14 side + c)
        // + sign evaluates it with coordinate c of body `side` shifted by +h / -h.  One
        // call site, so the row code is inlined once.
        int rb0 = -1, rb1 = -1;
        float tp = 0.f;
#pragma unroll 1
        for (int it = 0; it < 1 + 56; ++it) {
          const int side = it > 0 ? (it - 1) / 28 : 0, c = it > 0 ? ((it - 1) >> 1) % 14 : 0;
          const int sgn = it > 0 ? (it - 1) & 1 : 0;
          if (it > 0 && c % split != sub) continue;
          const int b = it == 0 ? -1 : (side ? rb1 : rb0);
          if (it > 0 && b < 0) {
            if (sgn) T[(2 * i + side) * 14 + c] = 0.f;
            continue;
          }
          const float t = adjoint_row_term(P, sm, lamprev, w, i, n_c, lam[k], ws_[k],
                                           ShiftedPose{b, c, sgn ? -h : h}, &rb0, &rb1);
          if (it == 0) {
            if (sub == 0) { tb[2 * i] = rb0; tb[2 * i + 1] = rb1; }
          } else if (sgn == 0) {
            tp = t;
          } else {
            T[(2 * i + side) * 14 + c] = (tp - t) / (2.f * h);
          }
        }
      }
    }
    for (int e = lane; e < 7 * B; e += WAVE) {  // dynamics terms: q+ only
      const int b = e / 7, c = e % 7;
      if (c % split != sub) continue;
      const float* bb = sm + b * BS;
      const tf q0 = ldtf(bb + BQ);
      const ShiftedPose sp{b, c, h}, sn{b, c, -h};
      const sv wu = ldsv(bb + BWU);
      const float dp = sdot(wu, dyn_residual(P, bb, sp.shift(q0, b, c).q));
      const float dn = sdot(wu, dyn_residual(P, bb, sn.shift(q0, b, c).q));
      Td[e] = (dp - dn) / (2.f * h);
    }
  } else if (P.cfg.adjoint_pose_vjp) {
    const float h = P.cfg.adjoint_pose_eps;
    for (int d = sub; d < 14 * B; d += split) {
      const int b = d / 14, kk = d % 14;
      float* ptr = sm + b * BS + (kk < 7 ? BQ + kk : BQP + kk - 7);
      const float orig = *ptr;
      wsync();
      if (lane == 0) *ptr = orig + h;
      wsync();
      float sp = adjoint_weighted_residual<R>(P, sm, lamprev, lane, w, n_r, n_c, lam, ws_, S);
      if (lane == 0) *ptr = orig - h;
      wsync();
      float sn = adjoint_weighted_residual<R>(P, sm, lamprev, lane, w, n_r, n_c, lam, ws_, S);
      if (lane == 0) { *ptr = orig; gpose[d] = (sp - sn) / (2.f * h); }
      wsync();
    }
  }

  if (split > 1) {
    __syncthreads();
    if (sub > 0) return;
  }
  if (P.cfg.adjoint_pose_vjp && P.fd_off >= 0) {  // sum the row-local differences per body
    wsync();
    const float* T = sm0 + P.fd_off;
    const int* tb = (const int*)(T + 28 * P.rows_cap);
    const float* Td = T + 30 * P.rows_cap;
    for (int e = lane; e < 14 * B; e += WAVE) {
      const int b = e / 14, c = e % 14;
      float acc = 0.f;
      for (int i = 0; i < n_r; ++i) {
        if (tb[2 * i] == b) acc += T[(2 * i) * 14 + c];
        if (tb[2 * i + 1] == b) acc += T[(2 * i + 1) * 14 + c];
      }
      if (c < 7) acc += Td[b * 7 + c];
      gpose[e] = acc;
    }
    wsync();
  }

  // ---- 5c. carry dq+/dq- of g_q + w^T dR/dq+ (residual.py:260-339), outputs ----
  for (int b = lane; b < B; b += WAVE) {
    const float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
    float g[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) g[i] = P.g_q[k * 7 + i] + gpose[b * 14 + i];
    const float w1 = bb[BU + 3], w2 = bb[BU + 4], w3 = bb[BU + 5], s = dt * 0.5f;
    float o[7];
    o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    o[3] = g[3] + s * (w3 * g[4] - w2 * g[5] - w1 * g[6]);
    o[4] = g[4] + s * (-w3 * g[3] + w1 * g[5] - w2 * g[6]);
    o[5] = g[5] + s * (w2 * g[3] - w1 * g[4] - w3 * g[6]);
    o[6] = g[6] + s * (w1 * g[3] + w2 * g[4] + w3 * g[5]);
    if (P.g_q_prev) {
#pragma unroll
      for (int i = 0; i < 7; ++i) P.g_q_prev[k * 7 + i] = o[i] + gpose[b * 14 + 7 + i];
    }
  }
  for (int d = lane; d < D; d += WAVE) {
    size_t o = (size_t)w * D + d;
    if (P.g_tpos) P.g_tpos[o] = gctl[d];
    if (P.g_tvel) P.g_tvel[o] = gctl[D + d];
    if (P.g_ke) P.g_ke[o] = gctl[2 * D + d];
    if (P.g_kd) P.g_kd[o] = gctl[3 * D + d];
  }
  if (P.g_mu)
    for (int s_ = lane; s_ < SC; s_ += WAVE) P.g_mu[(size_t)w * SC + s_] = gmu[s_];
  if (P.w_lam) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + WAVE * k;
      if (i < n_r) P.w_lam[(size_t)w * L.Nc + original_row(L, i, n_c)] = x[k];
    }
    for (int c = n_c + lane; c < L.K; c += WAVE) {
      size_t o = (size_t)w * L.Nc;
      P.w_lam[o + L.off_n + c] = 0.f;
      P.w_lam[o + L.off_f + 2 * c] = 0.f;
      P.w_lam[o + L.off_f + 2 * c + 1] = 0.f;
    }
  }
  int any = 0;
  if (__any(flags & FLAG_FACTOR_FAILURE)) any |= FLAG_FACTOR_FAILURE;
  if (__any(flags & FLAG_PAIR_OVERFLOW)) any |= FLAG_PAIR_OVERFLOW;
  if (__any(flags & FLAG_MU_GRAD_MISSING)) any |= FLAG_MU_GRAD_MISSING;
  if (lane == 0) {
    P.st_pcr[w] = pcr_it;
    P.st_flags[w] = (P.st_flags[w] & FLAG_CONTACT_OVERFLOW) | any;
  }
}

template <int R, bool BJ>
int adjoint_class_op(int op, const AdjParams& P, int blocks, int threads, size_t lds, hipStream_t stream) {
  if (op == 1)
    return hipFuncSetAttribute((const void*)adjoint_step_kernel<R, BJ>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds) == hipSuccess ? 0 : -4;
  hipLaunchKernelGGL((adjoint_step_kernel<R, BJ>), dim3(blocks), dim3(threads), lds, stream, P);
  return 0;
}

// defined in fb_adjoint_classes.hip, one per (preconditioner, class set) translation unit
int adjoint_class_op_bj0_set0(int R, int op, const AdjParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int adjoint_class_op_bj0_set1(int R, int op, const AdjParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int adjoint_class_op_bj1_set0(int R, int op, const AdjParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int adjoint_class_op_bj1_set1(int R, int op, const AdjParams& P, int blocks, int threads, size_t lds, hipStream_t st);

}  // namespace ofb
