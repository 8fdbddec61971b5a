// fb_rows.h -- per-row and per-body arithmetic of the Newton system, shared by the fused
// and the plain execution paths.
//
// Body data is addressed through a "body block" pointer: B consecutive blocks of BS
// floats (layout below) that live in LDS for the fused kernel and in HBM for the plain
// kernels.  The functions are templates on the launch-parameter struct, which must
// provide: m (DeviceModel), L (RowLayout), cfg (ofb_config), dt, topo, cf, ci,
// target_pos, target_vel.
#pragma once

#include <hip/hip_runtime.h>

#include "fb_math.h"
#include "fb_solver.h"

namespace ofb {

namespace rows {

constexpr int WAVE = 64;
constexpr int BS = 101;  // LDS floats per body (odd stride spreads bodies over banks)

// body block offsets (floats)
enum {
  BQ = 0,       // pose (7)
  BU = 7,       // velocity (6)
  BQP = 13,     // pose at step start (7)
  BUP = 20,     // velocity at step start (6)
  BF = 26,      // external wrench (6)
  BHD = 32,     // dynamics residual h_d (6)
  BIW = 38,     // world inverse inertia (9)
  BACC = 47,    // J^T x accumulator (6)
  BVV = 53,     // M^-1 acc (6)
  BBQ = 59,     // best pose (7)
  BBU = 66,     // best velocity (6)
  BM = 72,      // mass
  BMINV = 73,   // inverse mass
  BCOM = 74,    // com (3)
  BIB = 77,     // body inertia (9)
  BIIB = 86,    // body inverse inertia (9)
  BU0 = 95,     // velocity before the line-searched step (6)
};


__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Synchronises the lanes of one world: the wave barrier for a world of one wave (or half
// of one), the workgroup barrier for a world that spans several waves (such a world has
// its workgroup to itself).
template <int GW>
__device__ __forceinline__ void gsync() {
  if constexpr (GW <= WAVE) wsync();
  else __syncthreads();
}

#define OFB_DPP(v, ctrl, rmask)                                                                  \
  __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), (ctrl), (rmask), 0xF, false))

// Lane groups: a fused kernel runs one world per group of G lanes, G = 64 (the wave) or
// G = 32 (each half of the wave holds its own world).  The group reductions below stay
// inside a group, so the two worlds of a wave never mix; every lane of an active group
// must be active (the other group may have retired).
// 32-lane group sums that every lane of the group holds bit for bit (the stop rules and
// the PCR scalars of a world are evaluated by each of its lanes, so they must agree):
// pair, quad, half-row and row exchanges, each adding a lane's partial sum to its
// partner's, give every lane of a 16-lane row the same row sum (a + b and b + a round
// alike); v_permlane16_swap_b32 with both operands v then returns (rows 0,0,2,2) and
// (rows 1,1,3,3), whose sum is the half's total in every lane of the half -- no
// read-back through scalar registers.  The adds are kept out of FMA contraction: a
// product fused into the first exchange would round differently in the two lanes of a
// pair.
__device__ __forceinline__ float row_sum_sym(float v) {
#pragma clang fp contract(off)
  v = v + OFB_DPP(v, 0xB1, 0xF);   // quad_perm [1,0,3,2]
  v = v + OFB_DPP(v, 0x4E, 0xF);   // quad_perm [2,3,0,1]
  v = v + OFB_DPP(v, 0x141, 0xF);  // row_half_mirror: quads 0<->1, 2<->3
  v = v + OFB_DPP(v, 0x140, 0xF);  // row_mirror: half-rows 0<->1
  return v;
}
__device__ __forceinline__ float half_sum_from_rows(float v) {
#pragma clang fp contract(off)
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// 32-lane halves to the wave total: v_permlane32_swap_b32 with both operands v returns
// (half 0, half 0) and (half 1, half 1), whose sum every lane holds
__device__ __forceinline__ float wave_sum_from_halves(float v) {
#pragma clang fp contract(off)
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// sum over the 64 lanes, the same bits in every lane; all lanes must be active
__device__ __forceinline__ float wave_sum(float v) { return wave_sum_from_halves(half_sum_from_rows(row_sum_sym(v))); }

template <int G>
__device__ __forceinline__ float group_sum(float v) {
  static_assert(G == 32 || G == 64, "lane groups are 32 or 64 lanes");
  if (G == 64) return wave_sum(v);
  return half_sum_from_rows(row_sum_sym(v));
}

// two group sums whose DPP chains interleave (one chain's wait states are the other's
// issue slots)
template <int G>
__device__ __forceinline__ void group_sum2(float& a, float& b) {
#pragma clang fp contract(off)
  static_assert(G == 32 || G == 64, "lane groups are 32 or 64 lanes");
  a = a + OFB_DPP(a, 0xB1, 0xF);
  b = b + OFB_DPP(b, 0xB1, 0xF);
  a = a + OFB_DPP(a, 0x4E, 0xF);
  b = b + OFB_DPP(b, 0x4E, 0xF);
  a = a + OFB_DPP(a, 0x141, 0xF);
  b = b + OFB_DPP(b, 0x141, 0xF);
  a = a + OFB_DPP(a, 0x140, 0xF);
  b = b + OFB_DPP(b, 0x140, 0xF);
  a = half_sum_from_rows(a);
  b = half_sum_from_rows(b);
  if (G == 64) {
    a = wave_sum_from_halves(a);
    b = wave_sum_from_halves(b);
  }
}

template <int G>
__device__ __forceinline__ int group_max_int(int v) {
  for (int off = G / 2; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, G));
  return v;
}

template <int G>
__device__ __forceinline__ bool group_any(bool pred) {
  const unsigned long long b = __ballot(pred);
  if (G == 64) return b != 0ull;
  return ((threadIdx.x & 32) ? (b >> 32) : (b & 0xffffffffull)) != 0ull;
}

__device__ __forceinline__ v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ tf ldtf(const float* p) { return tf{mk(p[0], p[1], p[2]), qt{p[3], p[4], p[5], p[6]}}; }
__device__ __forceinline__ sv ldsv(const float* p) { return sv{mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5])}; }
__device__ __forceinline__ m3 ldm3(const float* p) {
  m3 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.m[i] = p[i];
  return r;
}
__device__ __forceinline__ void sttf(float* p, const tf& x) {
  p[0] = x.p.x; p[1] = x.p.y; p[2] = x.p.z; p[3] = x.q.x; p[4] = x.q.y; p[5] = x.q.z; p[6] = x.q.w;
}
__device__ __forceinline__ void stsv(float* p, const sv& x) {
  p[0] = x.a.x; p[1] = x.a.y; p[2] = x.a.z; p[3] = x.b.x; p[4] = x.b.y; p[5] = x.b.z;
}
__device__ __forceinline__ void atomic_add_sv(float* p, const sv& x) {
  atomicAdd(p + 0, x.a.x); atomicAdd(p + 1, x.a.y); atomicAdd(p + 2, x.a.z);
  atomicAdd(p + 3, x.b.x); atomicAdd(p + 4, x.b.y); atomicAdd(p + 5, x.b.z);
}
// M^-1 v with M^-1 = diag(m^-1 I3, I_w^-1)
__device__ __forceinline__ sv apply_minv(const float* body, const sv& v) {
  m3 Iw = ldm3(body + BIW);
  return sv{v.a * body[BMINV], mv(Iw, v.b)};
}

// ---------------------------------------------------------------------------------
// Per-row computations.  Each returns the linearisation (Jhat pair, C, active) and the
// residual h_c, and scatters -Jhat^T lambda dt into the dynamics residual with LDS atomics
// (with scatter = false the caller sums the contributions itself).
// ---------------------------------------------------------------------------------
struct RowOut {
  sv J0, J1;
  int b0, b1;
  float hc, C;
  bool active;
};

// ---------------------------------------------------------------------------------
// Step-constant row data.  Everything a row reads from the model, the control input and
// the packed contacts is gathered into a static record, together with everything of a
// contact row that depends only on the step-start poses: the reference evaluates the
// contact and friction Jacobians, the effective masses and the friction frame at
// body_q_prev (contact_constraint.py:118-166, friction_constraint.py:79-156 and
// :379-469, constraints/utils.py:6-33), so they are constant over the Newton loop.  The
// per-iteration row functions (*_row_s) work from a record and the current body state,
// with the same arithmetic in the same order as evaluating everything in place.  The
// fused kernel builds the records once per step into LDS; the plain kernels and the
// adjoint build them on the fly (the *_row wrappers).
// ---------------------------------------------------------------------------------

// joints (joint_constraint.py:62-162, control_constraint.py:117-338)
struct JointStatic {
  int parent, child, enabled, type;
  float comp;       // compliance (the joint's own, or the config's when negative)
  int dof_mode;     // mode of the first dof (revolute / prismatic)
  float ke, kd;     // its gains
  float target;     // its control target: position (mode 1) or velocity (mode 2)
  tf Xp, Xc;        // joint frames in the parent and child bodies
  v3 ax;            // axis of the first dof (revolute / prismatic), else 0
};

template <class PT>
__device__ __forceinline__ JointStatic joint_static(const PT& P, int w, int j) {
  const int J = P.m.J, D = P.m.D;
  const int k = w * J + j;
  JointStatic s;
  s.parent = P.m.joint_parent[k];
  s.child = P.m.joint_child[k];
  s.enabled = P.m.joint_enabled[k];
  s.type = P.m.joint_type[k];
  float c = P.m.joint_compliance[k];
  s.comp = c < 0.f ? P.cfg.compliance_joint : c;
  s.Xp = ldtf(P.m.joint_X_p + (size_t)k * 7);
  s.Xc = ldtf(P.m.joint_X_c + (size_t)k * 7);
  s.ax = mk(0.f, 0.f, 0.f);
  s.dof_mode = 0;
  s.ke = 0.f; s.kd = 0.f; s.target = 0.f;
  if (s.type == 0 || s.type == 1) {
    const int qs = P.m.joint_qd_start[k];
    s.ax = ld3(P.m.joint_axis + (size_t)(w * D + qs) * 3);
    s.dof_mode = P.m.joint_dof_mode[w * D + qs];
    s.ke = P.m.joint_target_ke[w * D + qs];
    s.kd = P.m.joint_target_kd[w * D + qs];
    if (s.dof_mode == 1) s.target = P.target_pos[w * D + qs];
    else if (s.dof_mode == 2) s.target = P.target_vel[w * D + qs];
  }
  return s;
}

// LDS record of a joint: 28 floats (7 quads)
constexpr int JS_WORDS = 28;
__device__ __forceinline__ void store_joint_static(float* d, const JointStatic& s) {
  float4* q = reinterpret_cast<float4*>(d);
  q[0] = make_float4(__int_as_float(s.parent), __int_as_float(s.child), __int_as_float(s.enabled),
                     __int_as_float(s.type));
  q[1] = make_float4(s.comp, __int_as_float(s.dof_mode), s.ke, s.kd);
  q[2] = make_float4(s.Xp.p.x, s.Xp.p.y, s.Xp.p.z, s.Xp.q.x);
  q[3] = make_float4(s.Xp.q.y, s.Xp.q.z, s.Xp.q.w, s.Xc.p.x);
  q[4] = make_float4(s.Xc.p.y, s.Xc.p.z, s.Xc.q.x, s.Xc.q.y);
  q[5] = make_float4(s.Xc.q.z, s.Xc.q.w, s.ax.x, s.ax.y);
  q[6] = make_float4(s.ax.z, s.target, 0.f, 0.f);
}
__device__ __forceinline__ JointStatic load_joint_static(const float* d) {
  const float4* q = reinterpret_cast<const float4*>(d);
  const float4 a = q[0], b = q[1], c = q[2], e = q[3], f = q[4], g = q[5], h = q[6];
  JointStatic s;
  s.parent = __float_as_int(a.x); s.child = __float_as_int(a.y);
  s.enabled = __float_as_int(a.z); s.type = __float_as_int(a.w);
  s.comp = b.x; s.dof_mode = __float_as_int(b.y); s.ke = b.z; s.kd = b.w;
  s.Xp = tf{mk(c.x, c.y, c.z), qt{c.w, e.x, e.y, e.z}};
  s.Xc = tf{mk(e.w, f.x, f.y), qt{f.z, f.w, g.x, g.y}};
  s.ax = mk(g.z, g.w, h.x);
  s.target = h.y;
  return s;
}

// How the row functions read body poses: straight from the LDS body blocks (every
// solver pass), or with one coordinate of one body shifted (the adjoint's central
// differences, which evaluate a row at a perturbed pose without touching the shared
// block).  coord 0..6: q+ (p, quaternion xyzw), 7..13: the step-start pose q-.
struct LdsPoses {
  __device__ __forceinline__ tf q(const float* sm, int b) const { return ldtf(sm + b * BS + BQ); }
  __device__ __forceinline__ tf qp(const float* sm, int b) const { return ldtf(sm + b * BS + BQP); }
};
struct ShiftedPose {
  int body, coord;
  float delta;
  __device__ __forceinline__ tf shift(tf x, int b, int k) const {
    if (b != body) return x;
    x.p.x += k == 0 ? delta : 0.f;
    x.p.y += k == 1 ? delta : 0.f;
    x.p.z += k == 2 ? delta : 0.f;
    x.q.x += k == 3 ? delta : 0.f;
    x.q.y += k == 4 ? delta : 0.f;
    x.q.z += k == 5 ? delta : 0.f;
    x.q.w += k == 6 ? delta : 0.f;
    return x;
  }
  __device__ __forceinline__ tf q(const float* sm, int b) const { return shift(ldtf(sm + b * BS + BQ), b, coord); }
  __device__ __forceinline__ tf qp(const float* sm, int b) const {
    return shift(ldtf(sm + b * BS + BQP), b, coord - 7);
  }
};

// world frames of a joint at the current poses (joint_kinematics.py:5-40)
template <class PR = LdsPoses>
__device__ __forceinline__ void joint_frames(const JointStatic& s, const float* sm, tf& Xwp, tf& Xwc, v3& r_p,
                                             v3& r_c, v3& pos_p, v3& pos_c, v3& com_p, const PR& pr = PR()) {
  const float* bc = sm + s.child * BS;
  tf bq = pr.q(sm, s.child);
  Xwc = tmul(bq, s.Xc);
  pos_c = Xwc.p;
  r_c = pos_c - tpoint(bq, ld3(bc + BCOM));
  tf bp = tident();
  com_p = mk(0.f, 0.f, 0.f);
  if (s.parent >= 0) { bp = pr.q(sm, s.parent); com_p = ld3(sm + s.parent * BS + BCOM); }
  Xwp = tmul(bp, s.Xp);
  pos_p = Xwp.p;
  r_p = pos_p - tpoint(bp, com_p);
}

// row k of a joint (constraints/joint_constraint.py:62-162, joint_kinematics.py:5-261)
template <class PT, class PR = LdsPoses>
__device__ RowOut joint_row_s(const PT& P, float* sm, const JointStatic& s, int k, float lam, bool scatter = true,
                              const PR& pr = PR()) {
  RowOut o;
  o.J0 = svzero(); o.J1 = svzero(); o.hc = 0.f; o.C = 0.f; o.active = false;
  o.b0 = s.parent;
  o.b1 = s.child;
  if (s.enabled == 0 || o.b1 < 0) return o;
  const int type = s.type;
  if (type < 0 || type > 3) return o;
  const float comp_ = s.comp;
  tf Xwp, Xwc; v3 r_p, r_c, pos_p, pos_c, com_p;
  joint_frames(s, sm, Xwp, Xwc, r_p, r_c, pos_p, pos_c, com_p, pr);
  const v3 ax = s.ax;
  sv Jp = svzero(), Jc = svzero();
  float err = 0.f;
  bool act = false;
  // compute_joint_row (joint_kinematics.py:191-261)
  int mode = -1, idx = 0;  // 0 linear, 1 angular, 2 revolute-ortho, 3 prismatic-ortho
  if (type == 0) { if (k < 2) { mode = 3; idx = k; } else if (k < 5) { mode = 1; idx = k - 2; } }
  else if (type == 1) { if (k < 3) { mode = 0; idx = k; } else if (k < 5) { mode = 2; idx = k - 3; } }
  else if (type == 2) { if (k < 3) { mode = 0; idx = k; } }
  else if (type == 3) { if (k < 3) { mode = 0; idx = k; } else if (k < 6) { mode = 1; idx = k - 3; } }
  if (mode == 0) {  // get_linear_component :43-74
    v3 a = unit(idx);
    Jc = sv{a, cross(r_c, a)};
    Jp = sv{-a, -cross(r_p, a)};
    err = comp(pos_c - pos_p, idx);
    act = true;
  } else if (mode == 1) {  // get_angular_component :77-113
    v3 aw = qrot(Xwp.q, unit(idx));
    Jc = sv{mk(0.f, 0.f, 0.f), aw};
    Jp = sv{mk(0.f, 0.f, 0.f), -aw};
    qt qr = qmul(qconj(Xwp.q), Xwc.q);
    float e = 2.f * (idx == 0 ? qr.x : (idx == 1 ? qr.y : qr.z));
    err = qr.w < 0.f ? -e : e;
    act = true;
  } else if (mode == 2) {  // get_revolute_angular_component :116-147
    v3 b1, b2;
    orthogonal_basis(ax, b1, b2);
    v3 axc = qrot(Xwc.q, ax);
    v3 tgt = qrot(Xwp.q, idx == 0 ? b1 : b2);
    err = dot(axc, tgt);
    v3 ra = cross(axc, tgt);
    Jc = sv{mk(0.f, 0.f, 0.f), ra};
    Jp = sv{mk(0.f, 0.f, 0.f), -ra};
    act = true;
  } else if (mode == 3) {  // get_prismatic_linear_component :150-183
    v3 b1, b2;
    orthogonal_basis(ax, b1, b2);
    v3 nv = qrot(Xwp.q, idx == 0 ? b1 : b2);
    err = dot(pos_c - pos_p, nv);
    Jc = sv{nv, cross(r_c, nv)};
    Jp = sv{-nv, cross(nv, pos_c - com_p)};  // com_p is the body-frame com, as the reference passes it
    act = true;
  }
  if (!act) return o;
  o.active = true;
  float dt = P.dt;
  if (scatter) {
    if (o.b0 >= 0) atomic_add_sv(sm + o.b0 * BS + BHD, Jp * (-lam * dt));
    atomic_add_sv(sm + o.b1 * BS + BHD, Jc * (-lam * dt));
  }
  o.hc = (err + comp_ * lam) / dt;
  o.J0 = Jp; o.J1 = Jc;
  o.C = comp_ / (dt * dt);
  return o;
}

template <class PT, class PR = LdsPoses>
__device__ RowOut joint_row(const PT& P, float* sm, int w, int row, float lam, bool /*full*/, const PR& pr = PR(),
                            bool scatter = true) {
  return joint_row_s(P, sm, joint_static(P, w, P.topo[row]), P.topo[P.L.Nj + row], lam, scatter, pr);
}

// control row of a joint (constraints/control_constraint.py:117-338)
template <class PT, class PR = LdsPoses>
__device__ RowOut control_row_s(const PT& P, float* sm, const JointStatic& s, float lam, bool scatter = true,
                                const PR& pr = PR()) {
  RowOut o;
  o.J0 = svzero(); o.J1 = svzero(); o.hc = 0.f; o.C = 0.f; o.active = false;
  o.b0 = s.parent;
  o.b1 = s.child;
  if (s.enabled == 0) return o;
  const int type = s.type;
  if (type != 0 && type != 1) return o;
  const int mode = s.dof_mode;
  if (mode == 0 || o.b1 < 0) return o;
  tf Xwp, Xwc; v3 r_p, r_c, pos_p, pos_c, com_p;
  joint_frames(s, sm, Xwp, Xwc, r_p, r_c, pos_p, pos_c, com_p, pr);
  const v3 ax = s.ax;
  sv up = svzero();
  if (o.b0 >= 0) up = ldsv(sm + o.b0 * BS + BU);
  sv uc = ldsv(sm + o.b1 * BS + BU);
  const float target = s.target;
  const float ke = s.ke, kd = s.kd;
  float dt = P.dt;
  float qv, qdv;
  sv Jp, Jc;
  bool angular;
  if (type == 1) {
    qt qr = qmul(qconj(Xwp.q), Xwc.q);
    float sh = qr.x * ax.x + qr.y * ax.y + qr.z * ax.z;
    qv = 2.f * atan2f(sh, qr.w);
    v3 aw = qrot(Xwp.q, ax);
    qdv = dot(uc.b - up.b, aw);
    Jc = sv{mk(0.f, 0.f, 0.f), aw};
    Jp = sv{mk(0.f, 0.f, 0.f), -aw};
    angular = true;
  } else {
    v3 aw = qrot(Xwp.q, ax);
    v3 delta = pos_c - pos_p;
    qv = dot(delta, aw);
    v3 vc = uc.a + cross(uc.b, r_c);
    v3 vp = up.a + cross(up.b, r_p);
    qdv = dot(vc - vp, aw) + dot(delta, cross(up.b, aw));
    Jc = sv{aw, cross(r_c, aw)};
    Jp = sv{-aw, -cross(pos_c - com_p, aw)};
    angular = false;
  }
  float err = 0.f, alpha = 0.f;
  if (mode == 1) {
    float raw = qv - target;
    if (angular) {
      const float PI = 3.141592653589793f, TWO_PI = 6.283185307179586f;
      float s_ = raw + PI;
      err = (s_ - TWO_PI * floorf(s_ / TWO_PI)) - PI;
    } else {
      err = raw;
    }
    err = err / dt;
    float den = dt * dt * ke + dt * kd;
    alpha = den > 1e-6f ? 1.f / den : 1.0e8f;
  } else if (mode == 2) {
    err = qdv - target;
    float den = dt * ke;
    alpha = den > 1e-6f ? 1.f / den : 1.0e8f;
  }
  o.active = true;
  if (scatter) {
    if (o.b0 >= 0) atomic_add_sv(sm + o.b0 * BS + BHD, Jp * (-lam * dt));
    atomic_add_sv(sm + o.b1 * BS + BHD, Jc * (-lam * dt));
  }
  o.hc = err + alpha * lam * dt;
  o.J0 = Jp; o.J1 = Jc;
  o.C = alpha;
  return o;
}

template <class PT, class PR = LdsPoses>
__device__ RowOut control_row(const PT& P, float* sm, int w, int crow, float lam, const PR& pr = PR(),
                              bool scatter = true) {
  return control_row_s(P, sm, joint_static(P, w, P.topo[2 * P.L.Nj + crow]), lam, scatter, pr);
}

struct ContactGeo {
  v3 p0, p1, n;
  float th0, th1;
  int s0, s1, b0, b1;
};

template <class PT>
__device__ __forceinline__ ContactGeo load_contact(const PT& P, int w, int c) {
  const int K = P.L.K;
  const float* f = P.cf + (size_t)w * CF * K + c;
  ContactGeo g;
  g.p0 = mk(f[CF_P0X * K], f[CF_P0Y * K], f[CF_P0Z * K]);
  g.p1 = mk(f[CF_P1X * K], f[CF_P1Y * K], f[CF_P1Z * K]);
  g.n = mk(f[CF_NX * K], f[CF_NY * K], f[CF_NZ * K]);
  g.th0 = f[CF_TH0 * K];
  g.th1 = f[CF_TH1 * K];
  const int32_t* ii = P.ci + (size_t)w * 2 * K + c;
  g.s0 = ii[0];
  g.s1 = ii[K];
  const int SC = P.m.SC;
  g.b0 = g.s0 >= 0 ? P.m.shape_body[w * SC + g.s0] : -1;
  g.b1 = g.s1 >= 0 ? P.m.shape_body[w * SC + g.s1] : -1;
  return g;
}

// effective mass J M^-1 J^T with the inverse inertia rotated by the step-start pose
// (constraints/utils.py:6-33)
template <class PR = LdsPoses>
__device__ __forceinline__ float eff_mass_prev(const float* sm, int b0, int b1, const sv& J0, const sv& J1,
                                               const PR& pr = PR()) {
  float val = 0.f;
  if (b0 >= 0) {
    const float* bb = sm + b0 * BS;
    m3 Iw = rot_inertia(pr.qp(sm, b0).q, ldm3(bb + BIIB));
    val += dot(J0.a, J0.a * bb[BMINV]) + dot(J0.b, mv(Iw, J0.b));
  }
  if (b1 >= 0) {
    const float* bb = sm + b1 * BS;
    m3 Iw = rot_inertia(pr.qp(sm, b1).q, ldm3(bb + BIIB));
    val += dot(J1.a, J1.a * bb[BMINV]) + dot(J1.b, mv(Iw, J1.b));
  }
  return val;
}

// contact normal row: the step-constant part (contact_constraint.py:118-166): geometry,
// bodies, the Jacobian at the step-start poses and the FB preconditioner dt^2 J M^-1 J^T
struct NormalStatic {
  int b0, b1, skip;
  v3 p0, p1, n;
  float th0, th1, precond;
  sv Ja, Jb;
  v3 r0, r1;  // lever arms (contact point less the centre of mass) at the step-start poses
};

template <class PT, class PR = LdsPoses>
__device__ __forceinline__ NormalStatic normal_static(const PT& P, const float* sm, int w, int c,
                                                      const PR& pr = PR()) {
  ContactGeo g = load_contact(P, w, c);
  NormalStatic s;
  s.skip = g.s0 == g.s1;
  s.b0 = g.b0; s.b1 = g.b1;
  s.p0 = g.p0; s.p1 = g.p1; s.n = g.n; s.th0 = g.th0; s.th1 = g.th1;
  tf q0p = tident(), q1p = tident();
  v3 com0 = mk(0.f, 0.f, 0.f), com1 = com0;
  if (g.b0 >= 0) { q0p = pr.qp(sm, g.b0); com0 = ld3(sm + g.b0 * BS + BCOM); }
  if (g.b1 >= 0) { q1p = pr.qp(sm, g.b1); com1 = ld3(sm + g.b1 * BS + BCOM); }
  v3 p0_adj_prev = tpoint(q0p, g.p0) - g.n * g.th0;
  v3 p1_adj_prev = tpoint(q1p, g.p1) + g.n * g.th1;
  s.Ja = svzero(); s.Jb = svzero();
  s.r0 = mk(0.f, 0.f, 0.f); s.r1 = s.r0;
  if (g.b0 >= 0) { s.r0 = p0_adj_prev - tpoint(q0p, com0); s.Ja = sv{g.n, cross(s.r0, g.n)}; }
  if (g.b1 >= 0) { s.r1 = p1_adj_prev - tpoint(q1p, com1); s.Jb = sv{-g.n, cross(s.r1, -g.n)}; }
  const float dt = P.dt;
  s.precond = dt * dt * eff_mass_prev(sm, g.b0, g.b1, s.Ja, s.Jb, pr);
  return s;
}

// contact normal row at the current poses (constraints/contact_constraint.py:48-374)
template <class PT, class PR = LdsPoses>
__device__ RowOut normal_row_s(const PT& P, float* sm, const NormalStatic& s, float lam, bool scatter = true,
                               const PR& pr = PR()) {
  RowOut o;
  o.J0 = svzero(); o.J1 = svzero(); o.hc = 0.f; o.C = 0.f; o.active = false; o.b0 = -1; o.b1 = -1;
  if (s.skip) return o;
  o.b0 = s.b0; o.b1 = s.b1;
  tf q0 = tident(), q1 = tident();
  if (s.b0 >= 0) q0 = pr.q(sm, s.b0);
  if (s.b1 >= 0) q1 = pr.q(sm, s.b1);
  v3 p0_adj = tpoint(q0, s.p0) - s.n * s.th0;
  v3 p1_adj = tpoint(q1, s.p1) + s.n * s.th1;
  float gap = dot(s.n, p0_adj - p1_adj);
  float dt = P.dt;
  const float precond = s.precond;
  float phi, dphi_dg, dphi_dl;
  if (P.cfg.contact_mode == OFB_CONTACT_EQUALITY) {
    phi = gap; dphi_dg = 1.f; dphi_dl = 0.f;
  } else {  // scaled_fisher_burmeister_diff (mechanics/complementarity.py:22-47)
    float sa = P.cfg.contact_fb_alpha * gap, sb = precond * lam;
    float nrm = sqrtf(sa * sa + sb * sb + P.cfg.contact_fb_eps_sq);
    phi = sa + sb - nrm;
    dphi_dg = P.cfg.contact_fb_alpha * (1.f - sa / nrm);
    dphi_dl = precond * (1.f - sb / nrm);
  }
  o.J0 = s.Ja * dphi_dg;
  o.J1 = s.Jb * dphi_dg;
  o.active = true;
  if (scatter) {
    if (s.b0 >= 0) atomic_add_sv(sm + s.b0 * BS + BHD, o.J0 * (-lam * dt));
    if (s.b1 >= 0) atomic_add_sv(sm + s.b1 * BS + BHD, o.J1 * (-lam * dt));
  }
  o.hc = phi / dt;
  o.C = (dphi_dl + P.cfg.compliance_contact) / (dt * dt);
  return o;
}

template <class PT, class PR = LdsPoses>
__device__ RowOut normal_row(const PT& P, float* sm, int w, int c, float lam, const PR& pr = PR(),
                             bool scatter = true) {
  return normal_row_s(P, sm, normal_static(P, sm, w, c, pr), lam, scatter, pr);
}

// friction rows of a contact: the step-constant part - skip rule, shape parameters, the
// friction frame and both tangent Jacobians at the step-start poses, the effective-mass
// scale r (friction_constraint.py:39-194, :379-469)
struct FrictionStatic {
  int b0, b1, skip;
  int closed;  // isotropic frame without anisotropy axes (d w / d mu has a closed form)
  sv Jt1a, Jt2a, Jt1b, Jt2b;
  float rr, mux, muy, gate_mu, mss, vst, lat;
  v3 t1, t2;  // the friction frame
};

template <class PT, class PR = LdsPoses>
__device__ __forceinline__ FrictionStatic friction_static(const PT& P, const float* sm, int w, int c,
                                                          const PR& pr = PR()) {
  ContactGeo g = load_contact(P, w, c);
  FrictionStatic s;
  s.b0 = g.b0; s.b1 = g.b1;
  s.closed = 1;
  s.Jt1a = svzero(); s.Jt2a = svzero(); s.Jt1b = svzero(); s.Jt2b = svzero();
  s.rr = 0.f; s.mux = 0.f; s.muy = 0.f; s.gate_mu = 0.f; s.mss = -1.f; s.vst = -1.f; s.lat = -1.f;
  s.t1 = mk(0.f, 0.f, 0.f); s.t2 = s.t1;
  bool skip = g.s0 == g.s1;
  if (P.cfg.friction_mode == OFB_FRICTION_BILATERAL_PATCH && !skip) {  // bilateral_row_is_skipped :39-76
    for (int jj = 0; jj < c; ++jj) {
      const int32_t* ii = P.ci + (size_t)w * 2 * P.L.K + jj;
      if (ii[0] == g.s0 && ii[P.L.K] == g.s1) { skip = true; break; }
    }
  }
  s.skip = skip;
  if (skip) return s;
  const int SC = P.m.SC;
  float mu0 = P.m.shape_mu[w * SC + g.s0], mu1 = P.m.shape_mu[w * SC + g.s1];
  s.gate_mu = fmaxf(mu0, mu1);
  tf q0p = tident(), q1p = tident();
  v3 com0 = mk(0.f, 0.f, 0.f), com1 = com0;
  if (g.b0 >= 0) { q0p = pr.qp(sm, g.b0); com0 = ld3(sm + g.b0 * BS + BCOM); }
  if (g.b1 >= 0) { q1p = pr.qp(sm, g.b1); com1 = ld3(sm + g.b1 * BS + BCOM); }
  // resolve_friction_frame :79-156
  v3 ax0 = ld3(P.m.shape_friction_axis + (size_t)(w * SC + g.s0) * 3);
  v3 ax1 = ld3(P.m.shape_friction_axis + (size_t)(w * SC + g.s1) * 3);
  float mp0 = P.m.shape_mu_perp[w * SC + g.s0], mp1 = P.m.shape_mu_perp[w * SC + g.s1];
  v3 t1, t2;
  float mux, muy;
  bool an0 = dot(ax0, ax0) > 1e-6f && g.b0 >= 0;
  bool an1 = dot(ax1, ax1) > 1e-6f && g.b1 >= 0;
  if (!an0 && !an1) {
    orthogonal_basis(g.n, t1, t2);
    mux = muy = (mu0 + mu1) * 0.5f;
  } else {
    v3 aw;
    float mal, mpa, mot;
    if (an0) { aw = qrot(q0p.q, ax0); mal = mu0; mpa = mp0 >= 0.f ? mp0 : mu0; mot = mu1; }
    else { aw = qrot(q1p.q, ax1); mal = mu1; mpa = mp1 >= 0.f ? mp1 : mu1; mot = mu0; }
    v3 at = aw - g.n * dot(aw, g.n);
    float tl2 = dot(at, at);
    if (tl2 < 1e-8f) {
      orthogonal_basis(g.n, t1, t2);
      mux = muy = (mal + mot) * 0.5f;
    } else {
      t1 = at * (1.f / sqrtf(tl2));
      t2 = cross(g.n, t1);
      mux = (mal + mot) * 0.5f;
      muy = (mpa + mot) * 0.5f;
    }
  }
  s.closed = (!an0 && !an1) ? 1 : 0;
  s.mux = mux; s.muy = muy;
  s.t1 = t1; s.t2 = t2;
  // resolve_stribeck_params :159-176 / resolve_stribeck_lateral_only :179-194
  float ss0 = P.m.shape_mu_stiction_scale[w * SC + g.s0], vs0 = P.m.shape_v_stribeck[w * SC + g.s0];
  float ss1 = P.m.shape_mu_stiction_scale[w * SC + g.s1], vs1 = P.m.shape_v_stribeck[w * SC + g.s1];
  float mss = -1.f, vst = -1.f;
  if (ss0 > 0.f && vs0 > 0.f) { mss = ss0; vst = vs0; }
  else if (ss1 > 0.f && vs1 > 0.f) { mss = ss1; vst = vs1; }
  float lo0 = P.m.shape_stribeck_lateral_only[w * SC + g.s0], lo1 = P.m.shape_stribeck_lateral_only[w * SC + g.s1];
  s.lat = lo0 > 0.f ? lo0 : (lo1 > 0.f ? lo1 : -1.f);
  s.mss = mss; s.vst = vst;
  // compute_friction_core :379-469
  if (g.b0 >= 0) {
    v3 r_ = (tpoint(q0p, g.p0) - g.n * g.th0) - tpoint(q0p, com0);
    s.Jt1a = sv{t1, cross(r_, t1)};
    s.Jt2a = sv{t2, cross(r_, t2)};
  }
  if (g.b1 >= 0) {
    v3 r_ = (tpoint(q1p, g.p1) + g.n * g.th1) - tpoint(q1p, com1);
    s.Jt1b = sv{-t1, -cross(r_, t1)};
    s.Jt2b = sv{-t2, -cross(r_, t2)};
  }
  float wt1 = eff_mass_prev(sm, g.b0, g.b1, s.Jt1a, s.Jt1b, pr);
  float wt2 = eff_mass_prev(sm, g.b0, g.b1, s.Jt2a, s.Jt2b, pr);
  s.rr = (wt1 + wt2) * 0.5f;
  return s;
}

// LDS record of a contact, 9 quads: bodies and skip flags; the geometry (body-frame
// points, normal, thicknesses) and the normal row's preconditioner; the lever arms r0, r1
// at the step-start poses; the friction frame t1, t2 and the friction row parameters.
// The four Jacobians are rebuilt from them on load (normal (n, r x n), tangents (t, r x t)
// and their negations for the second body, as the builders form them), which holds a
// contact in 36 floats instead of 68.  Lanes read the records of different contacts with
// 128-bit loads; an odd number of quads per record puts consecutive lanes' quads on
// distinct banks.
constexpr int CS_WORDS = 36;
__device__ __forceinline__ void store_contact_static(float* d, const NormalStatic& n, const FrictionStatic& f) {
  float4* q = reinterpret_cast<float4*>(d);
  q[0] = make_float4(__int_as_float(n.b0), __int_as_float(n.b1), __int_as_float(n.skip), __int_as_float(f.skip));
  q[1] = make_float4(n.p0.x, n.p0.y, n.p0.z, n.p1.x);
  q[2] = make_float4(n.p1.y, n.p1.z, n.n.x, n.n.y);
  q[3] = make_float4(n.n.z, n.th0, n.th1, n.precond);
  q[4] = make_float4(n.r0.x, n.r0.y, n.r0.z, __int_as_float(f.closed));
  q[5] = make_float4(n.r1.x, n.r1.y, n.r1.z, f.rr);
  q[6] = make_float4(f.t1.x, f.t1.y, f.t1.z, f.mux);
  q[7] = make_float4(f.t2.x, f.t2.y, f.t2.z, f.muy);
  q[8] = make_float4(f.gate_mu, f.mss, f.vst, f.lat);
}
__device__ __forceinline__ NormalStatic load_normal_static(const float* d) {
  const float4* q = reinterpret_cast<const float4*>(d);
  const float4 a = q[0], b = q[1], c = q[2], e = q[3], r0 = q[4], r1 = q[5];
  NormalStatic s;
  s.b0 = __float_as_int(a.x); s.b1 = __float_as_int(a.y); s.skip = __float_as_int(a.z);
  s.p0 = mk(b.x, b.y, b.z);
  s.p1 = mk(b.w, c.x, c.y);
  s.n = mk(c.z, c.w, e.x);
  s.th0 = e.y; s.th1 = e.z; s.precond = e.w;
  s.r0 = mk(r0.x, r0.y, r0.z);
  s.r1 = mk(r1.x, r1.y, r1.z);
  s.Ja = svzero(); s.Jb = svzero();
  if (s.b0 >= 0) s.Ja = sv{s.n, cross(s.r0, s.n)};
  if (s.b1 >= 0) s.Jb = sv{-s.n, cross(s.r1, -s.n)};
  return s;
}
__device__ __forceinline__ FrictionStatic load_friction_static(const float* d) {
  const float4* q = reinterpret_cast<const float4*>(d);
  const float4 a = q[0], r0 = q[4], r1 = q[5], t1 = q[6], t2 = q[7], e = q[8];
  FrictionStatic s;
  s.b0 = __float_as_int(a.x); s.b1 = __float_as_int(a.y); s.skip = __float_as_int(a.w);
  s.closed = __float_as_int(r0.w);
  s.rr = r1.w; s.mux = t1.w; s.muy = t2.w;
  s.gate_mu = e.x; s.mss = e.y; s.vst = e.z; s.lat = e.w;
  s.t1 = mk(t1.x, t1.y, t1.z);
  s.t2 = mk(t2.x, t2.y, t2.z);
  s.Jt1a = svzero(); s.Jt2a = svzero(); s.Jt1b = svzero(); s.Jt2b = svzero();
  if (!s.skip) {
    const v3 ra = mk(r0.x, r0.y, r0.z), rb = mk(r1.x, r1.y, r1.z);
    if (s.b0 >= 0) {
      s.Jt1a = sv{s.t1, cross(ra, s.t1)};
      s.Jt2a = sv{s.t2, cross(ra, s.t2)};
    }
    if (s.b1 >= 0) {
      s.Jt1b = sv{-s.t1, -cross(rb, s.t1)};
      s.Jt2b = sv{-s.t2, -cross(rb, s.t2)};
    }
  }
  return s;
}

// friction row t in {0,1} of contact c at the current velocities and lambda_prev
// (constraints/friction_constraint.py:79-908).
// dwdmu (adjoint only, may be null): d(weight)/d(mu_x) of the isotropic cone, written as
// NaN where the weight's dependence on mu has no closed form here (anisotropic axes,
// Stribeck scaling, the elliptic cone) and 0 where the row does not depend on mu.
template <class PT>
__device__ RowOut friction_row_s(const PT& P, float* sm, const float* lamprev, const FrictionStatic& s, int c,
                                 int t, int n_c, float lam, float* dwdmu = nullptr, bool scatter = true) {
  if (dwdmu) *dwdmu = 0.f;
  RowOut o;
  o.J0 = svzero(); o.J1 = svzero(); o.hc = 0.f; o.C = 0.f; o.active = false; o.b0 = -1; o.b1 = -1;
  if (s.skip) return o;
  // lambda_prev of this contact: normal row then the two friction rows (compacted index)
  const int base_n = P.L.Nj + P.L.Nctrl;
  float fn_prev = lamprev[base_n + c];
  if (s.gate_mu * fn_prev <= P.cfg.friction_load_gate) return o;
  o.b0 = s.b0; o.b1 = s.b1;
  float l1p = lamprev[base_n + n_c + 2 * c], l2p = lamprev[base_n + n_c + 2 * c + 1];
  sv vel0 = svzero(), vel1 = svzero();
  if (s.b0 >= 0) vel0 = ldsv(sm + s.b0 * BS + BU);
  if (s.b1 >= 0) vel1 = ldsv(sm + s.b1 * BS + BU);
  float mux = s.mux, muy = s.muy;
  const float mss = s.mss, vst = s.vst, lat = s.lat, rr = s.rr;
  // compute_friction_model :197-371
  float vt0 = sdot(s.Jt1a, vel0) + sdot(s.Jt1b, vel1);
  float vt1 = sdot(s.Jt2a, vel0) + sdot(s.Jt2b, vel1);
  if (mss > 0.f && vst > 0.f) {
    float sp = sqrtf(vt0 * vt0 + vt1 * vt1);
    float factor = 1.f + (mss - 1.f) * expf(-sp / vst);
    if (lat > 0.f && mux != muy) { mux = mux * factor; }
    else { mux = mux * factor; muy = muy * factor; }
  }
  float dt = P.dt;
  float wx = 0.f, wy = 0.f;
  if (P.cfg.friction_mode == OFB_FRICTION_CONE) {
    const float eps = 1e-8f;
    float denom_eps = 1e-6f * dt;
    float d0 = vt0 * dt, d1 = vt1 * dt;
    float i0 = l1p * dt, i1 = l2p * dt;
    if (mux == muy) {
      float dn = sqrtf(d0 * d0 + d1 * d1 + eps);
      float raw = sqrtf(i0 * i0 + i1 * i1);
      float lim = mux * (fn_prev * dt);
      float gap = lim - fminf(raw, lim);
      float sa = dn, sb = rr * gap;
      float phi = sa + sb - sqrtf(sa * sa + sb * sb);
      float den = rr * raw + phi + denom_eps;
      float wv = rr * ((dn - phi) / den);
      if (dwdmu) {  // lim = mu fn_prev dt enters through gap
        float nrm = sqrtf(sa * sa + sb * sb);
        float dphi = rr * (1.f - (nrm > 0.f ? sb / nrm : 0.f)) * (raw < lim ? 1.f : 0.f) * fn_prev * dt;
        bool closed = s.closed && !(mss > 0.f && vst > 0.f);
        *dwdmu = !closed ? __int_as_float(0x7fc00000)
                         : ((wv < 0.f || wv > 1e5f) ? 0.f : -rr * dphi * (den + (dn - phi)) / (den * den));
      }
      wv = fminf(fmaxf(wv, 0.f), 1e5f);
      wx = wv; wy = wv;
    } else {
      float mxs = fmaxf(mux, 1e-6f), mys = fmaxf(muy, 1e-6f);
      float dxt = mxs * d0, dyt = mys * d1;
      float dn = sqrtf(dxt * dxt + dyt * dyt + eps);
      float imx = 1.f / mxs, imy = 1.f / mys;
      float fxt = i0 * imx, fyt = i1 * imy;
      float raw = sqrtf(fxt * fxt + fyt * fyt);
      float lim = fn_prev * dt;
      float gap = lim - fminf(raw, lim);
      float sa = dn, sb = rr * gap;
      float phi = sa + sb - sqrtf(sa * sa + sb * sb);
      float wt = rr * ((dn - phi) / (rr * raw + phi + denom_eps));
      wt = fminf(fmaxf(wt, 0.f), 1e5f);
      wx = wt * imx * imx;
      wy = wt * imy * imy;
      if (dwdmu) *dwdmu = __int_as_float(0x7fc00000);
    }
  }
  o.active = true;
  sv Ja = t == 0 ? s.Jt1a : s.Jt2a;
  sv Jb = t == 0 ? s.Jt1b : s.Jt2b;
  float wgt = t == 0 ? wx : wy;
  float vt = t == 0 ? vt0 : vt1;
  if (scatter) {
    if (s.b0 >= 0) atomic_add_sv(sm + s.b0 * BS + BHD, Ja * (-dt * lam));
    if (s.b1 >= 0) atomic_add_sv(sm + s.b1 * BS + BHD, Jb * (-dt * lam));
  }
  o.hc = vt + wgt * lam;
  o.J0 = Ja; o.J1 = Jb;
  o.C = wgt / dt + P.cfg.compliance_friction;
  return o;
}

template <class PT, class PR = LdsPoses>
__device__ RowOut friction_row(const PT& P, float* sm, const float* lamprev, int w, int c, int t, int n_c,
                               float lam, float* dwdmu = nullptr, const PR& pr = PR(), bool scatter = true) {
  return friction_row_s(P, sm, lamprev, friction_static(P, sm, w, c, pr), c, t, n_c, lam, dwdmu, scatter);
}

// dispatch on the compacted row index
template <class PT, class PR = LdsPoses>
__device__ __forceinline__ RowOut assemble_row(const PT& P, float* sm, const float* lamprev, int w, int i,
                                               int n_c, float lam, const PR& pr = PR(), bool scatter = true) {
  const int Nj = P.L.Nj, Nct = P.L.Nctrl;
  if (i < Nj) return joint_row(P, sm, w, i, lam, true, pr, scatter);
  if (i < Nj + Nct) return control_row(P, sm, w, i - Nj, lam, pr, scatter);
  int c = i - Nj - Nct;
  if (c < n_c) return normal_row(P, sm, w, c, lam, pr, scatter);
  int f = c - n_c;
  return friction_row(P, sm, lamprev, w, f >> 1, f & 1, n_c, lam, nullptr, pr, scatter);
}

// dispatch on the compacted row index with the step's static records in LDS (js: one
// record per joint, cs: one per contact); rj, rk: the row's joint and its row within the
// joint (joint rows), the row's joint (control rows)
template <class PT>
__device__ __forceinline__ RowOut assemble_row_cached(const PT& P, float* sm, const float* lamprev, const float* js,
                                                      const float* cs, int i, int n_c, int rj, int rk, float lam,
                                                      bool scatter = true) {
  const int Nj = P.L.Nj, Nct = P.L.Nctrl;
  if (i < Nj) return joint_row_s(P, sm, load_joint_static(js + JS_WORDS * rj), rk, lam, scatter);
  if (i < Nj + Nct) return control_row_s(P, sm, load_joint_static(js + JS_WORDS * rj), lam, scatter);
  int c = i - Nj - Nct;
  if (c < n_c) return normal_row_s(P, sm, load_normal_static(cs + CS_WORDS * c), lam, scatter);
  int f = c - n_c;
  return friction_row_s(P, sm, lamprev, load_friction_static(cs + CS_WORDS * (f >> 1)), f >> 1, f & 1, n_c, lam,
                        nullptr, scatter);
}

__device__ __forceinline__ int original_row(const RowLayout& L, int i, int n_c) {
  if (i < L.Nj + L.Nctrl) return i;
  int c = i - L.Nj - L.Nctrl;
  if (c < n_c) return L.off_n + c;
  int f = c - n_c;
  return L.off_f + f;
}

// dynamics residual of one body at rotation q (constraints/dynamics_constraint.py:6-50):
// M(q)(u - u-) - dt (f + m g + gyro)
template <class PT>
__device__ __forceinline__ sv dyn_residual(const PT& P, const float* bb, qt q) {
  const v3 g = mk(P.m.gravity[0], P.m.gravity[1], P.m.gravity[2]);
  m3 Iw = rot_inertia(q, ldm3(bb + BIB));
  sv u = ldsv(bb + BU), up = ldsv(bb + BUP), f = ldsv(bb + BF);
  float m = bb[BM];
  v3 wv = u.b;
  v3 gyro = cross(wv, mv(Iw, wv)) * P.cfg.gyro_scale;
  v3 lin = (u.a - up.a) * m - (f.a + g * m) * P.dt;
  v3 ang = mv(Iw, u.b - up.b) - (f.b + gyro) * P.dt;
  return sv{lin, ang};
}

// dynamics residual of every body; lanes over bodies
template <class PT>
__device__ __forceinline__ void dynamics_bodies(const PT& P, float* sm, int lane, bool with_inv, int stride = WAVE) {
  const int B = P.m.B;
  for (int b = lane; b < B; b += stride) {
    float* bb = sm + b * BS;
    qt q = qt{bb[BQ + 3], bb[BQ + 4], bb[BQ + 5], bb[BQ + 6]};
    if (with_inv) {
      m3 Iwi = rot_inertia(q, ldm3(bb + BIIB));
#pragma unroll
      for (int k = 0; k < 9; ++k) bb[BIW + k] = Iwi.m[k];
    }
    stsv(bb + BHD, dyn_residual(P, bb, q));
  }
}

// integrate_body_pose (mechanics/pose_integration.py:6-38)
__device__ __forceinline__ tf integrate(const sv& vel, const tf& pp, v3 com, float dt) {
  v3 pcom = pp.p + qrot(pp.q, com);
  v3 op = pcom + vel.a * dt;
  float s = dt * 0.5f;
  float t2 = pp.q.x, t3 = pp.q.y, t4 = pp.q.z, t1 = pp.q.w;
  v3 wv = vel.b;
  qt rq;
  rq.x = pp.q.x + (t1 * wv.x + t4 * wv.y - t3 * wv.z) * s;
  rq.y = pp.q.y + (-t4 * wv.x + t1 * wv.y + t2 * wv.z) * s;
  rq.z = pp.q.z + (t3 * wv.x - t2 * wv.y + t1 * wv.z) * s;
  rq.w = pp.q.w + (-t2 * wv.x - t3 * wv.y - t4 * wv.z) * s;
  float l = sqrtf(rq.x * rq.x + rq.y * rq.y + rq.z * rq.z + rq.w * rq.w);
  qt rn = l > 0.f ? qt{rq.x / l, rq.y / l, rq.z / l, rq.w / l} : qt{0.f, 0.f, 0.f, 0.f};
  return tf{op - qrot(rn, com), rn};
}

}  // namespace rows

}  // namespace ofb
