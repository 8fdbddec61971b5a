// fb_oracle.cpp -- CPU restatement of Ostrich's Fisher-Burmeister Newton contact solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
// oracle/_build/libfb_oracle.so and may be loaded only by tests/, by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, always as the checker
// or as the timed CPU baseline, never as part of the product path.
//
// It restates, serially and world by world, the kernel sequence the reference runs for
// `OstrichEngine.step` (src/ostrich/core/engine.py:25-71).  Every function cites the
// reference file:line it follows.  It is templated on the arithmetic type so the same
// code yields the FP32 mirror and the FP64 reference solve.
//
// Loop order mirrors Warp's CPU device (a launch is a serial loop over the thread ids,
// so scatter-adds land in row order); the batch-global stop flags of the Newton and PCR
// loops are reproduced in OFB_TERM_GLOBAL mode.

#include "../include/ostrich_fb.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

// ----------------------------------------------------------------------------------------
// Small linear algebra (Warp semantics for transform / quaternion helpers)
// ----------------------------------------------------------------------------------------
template <class T> struct V3 { T x, y, z; };
template <class T> struct Qt { T x, y, z, w; };
template <class T> struct Tf { V3<T> p; Qt<T> q; };
template <class T> struct Sp { T v[6]; };
template <class T> struct M3 { T m[9]; };

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> mul(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> neg(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T comp(V3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class T> inline V3<T> normalize(V3<T> a) {
  T l = std::sqrt(dot(a, a));
  return l > T(0) ? mul(a, T(1) / l) : v3<T>(0, 0, 0);
}

// quat_rotate for a unit quaternion: v (2w^2-1) + 2 (q.v) q + 2 w (q x v)
template <class T> inline V3<T> qrot(Qt<T> q, V3<T> v) {
  V3<T> u{q.x, q.y, q.z};
  T c = T(2) * q.w * q.w - T(1);
  T d = T(2) * dot(u, v);
  V3<T> uxv = cross(u, v);
  return add(add(mul(v, c), mul(u, d)), mul(uxv, T(2) * q.w));
}
template <class T> inline Qt<T> qmul(Qt<T> a, Qt<T> b) {
  return {a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y,
          a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
          a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> inline Qt<T> qconj(Qt<T> a) { return {-a.x, -a.y, -a.z, a.w}; }
template <class T> inline Qt<T> qnormalize(Qt<T> a) {
  T l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
  if (l > T(0)) { T s = T(1) / l; return {a.x * s, a.y * s, a.z * s, a.w * s}; }
  return {T(0), T(0), T(0), T(0)};
}
template <class T> inline M3<T> qmat(Qt<T> q) {
  T x = q.x, y = q.y, z = q.z, w = q.w;
  M3<T> R;
  R.m[0] = T(1) - T(2) * (y * y + z * z); R.m[1] = T(2) * (x * y - z * w); R.m[2] = T(2) * (x * z + y * w);
  R.m[3] = T(2) * (x * y + z * w); R.m[4] = T(1) - T(2) * (x * x + z * z); R.m[5] = T(2) * (y * z - x * w);
  R.m[6] = T(2) * (x * z - y * w); R.m[7] = T(2) * (y * z + x * w); R.m[8] = T(1) - T(2) * (x * x + y * y);
  return R;
}
template <class T> inline V3<T> mv(const M3<T>& A, V3<T> v) {
  return {A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
          A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}
// R I R^T  (mechanics/mass.py:21-33)
template <class T> inline M3<T> world_inertia(Qt<T> q, const M3<T>& I) {
  M3<T> R = qmat(q), RI, out;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += R.m[i * 3 + k] * I.m[k * 3 + j];
      RI.m[i * 3 + j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += RI.m[i * 3 + k] * R.m[j * 3 + k];
      out.m[i * 3 + j] = s;
    }
  return out;
}
template <class T> inline V3<T> tpoint(const Tf<T>& X, V3<T> p) { return add(X.p, qrot(X.q, p)); }
template <class T> inline Tf<T> tmul(const Tf<T>& A, const Tf<T>& B) {
  return Tf<T>{add(A.p, qrot(A.q, B.p)), qmul(A.q, B.q)};
}
template <class T> inline Tf<T> tident() { return Tf<T>{v3<T>(0, 0, 0), Qt<T>{0, 0, 0, 1}}; }

template <class T> inline Sp<T> sp(V3<T> a, V3<T> b) { return Sp<T>{{a.x, a.y, a.z, b.x, b.y, b.z}}; }
template <class T> inline Sp<T> spzero() { return Sp<T>{{0, 0, 0, 0, 0, 0}}; }
template <class T> inline V3<T> top(const Sp<T>& s) { return {s.v[0], s.v[1], s.v[2]}; }
template <class T> inline V3<T> bot(const Sp<T>& s) { return {s.v[3], s.v[4], s.v[5]}; }
template <class T> inline T sdot(const Sp<T>& a, const Sp<T>& b) {
  T s = 0;
  for (int i = 0; i < 6; ++i) s += a.v[i] * b.v[i];
  return s;
}
template <class T> inline Sp<T> sscale(const Sp<T>& a, T s) {
  Sp<T> r;
  for (int i = 0; i < 6; ++i) r.v[i] = a.v[i] * s;
  return r;
}
// [m * top, I * bot]  (mechanics/mass.py:10-18)
template <class T> inline Sp<T> spatial_momentum(T m, const M3<T>& I, const Sp<T>& v) {
  return sp(mul(top(v), m), mv(I, bot(v)));
}

// orthogonal_basis (mechanics/geometry.py:4-20)
template <class T> inline void orthogonal_basis(V3<T> a, V3<T>& b1, V3<T>& b2) {
  V3<T> v;
  if (std::fabs(a.x) <= std::fabs(a.y) && std::fabs(a.x) <= std::fabs(a.z)) v = v3<T>(1, 0, 0);
  else if (std::fabs(a.y) <= std::fabs(a.z)) v = v3<T>(0, 1, 0);
  else v = v3<T>(0, 0, 1);
  b1 = normalize(cross(a, v));
  b2 = cross(a, b1);
}

// scaled_fisher_burmeister (mechanics/complementarity.py:4-19)
template <class T> inline T fb(T a, T b, T alpha, T beta) {
  T sa = alpha * a, sb = beta * b;
  return sa + sb - std::sqrt(sa * sa + sb * sb);
}
// scaled_fisher_burmeister_diff (mechanics/complementarity.py:22-47)
template <class T> inline void fb_diff(T a, T b, T alpha, T beta, T eps_sq, T& val, T& da, T& db) {
  T sa = alpha * a, sb = beta * b;
  T nrm = std::sqrt(sa * sa + sb * sb + eps_sq);
  val = sa + sb - nrm;
  da = alpha * (T(1) - sa / nrm);
  db = beta * (T(1) - sb / nrm);
}

// ----------------------------------------------------------------------------------------
// Batch state
// ----------------------------------------------------------------------------------------
inline int joint_row_count(int type) {  // core/model.py:88-99
  static const int m[8] = {5, 5, 3, 6, 0, 1, 6, 0};
  return (type >= 0 && type < 8) ? m[type] : 0;
}
inline bool joint_rows_supported(int type) { return type >= 0 && type <= 3; }

template <class T> struct Batch {
  const ofb_model& M;
  const ofb_config& cfg;
  int W, B, J, D, S, G, SC, Cmax, Nj, Nctrl, Nc, Nu;
  std::vector<int> joint_off, ctrl_off;  // per joint (world 0 topology), -1 if none
  T dt;

  // packed contacts [W*Cmax]
  std::vector<int> ccount, cs0, cs1;
  std::vector<T> cp0, cp1, cn, ct0, ct1;

  // body state [W*B]
  std::vector<Tf<T>> q, q_prev;
  std::vector<Sp<T>> u, u_prev, f_ext, h_d, du, tmp_body;
  std::vector<M3<T>> Iw_inv;
  std::vector<T> tpos, tvel;  // [W*D]

  // rows [W*Nc]
  std::vector<T> lam, lam_prev, dlam, h_c, C, active, rhs, Pinv;
  std::vector<Sp<T>> J0, J1;
  std::vector<T> dw_dmu;  // [W*Cmax] d(friction weight)/d(mu_x) of the isotropic cone, NaN if not closed-form
  std::vector<int> b0, b1;

  // PCR
  std::vector<T> r, z, Az, p, Ap, y;
  std::vector<T> b_sq, r_sq, zAz, zAz_old, yAp;

  // block-Jacobi
  struct PairBlocks {
    std::vector<int> count;   // [npairs]
    std::vector<int> members; // [npairs*64]
    std::vector<T> L;         // [npairs*64*64]
    std::vector<int> fail;    // [npairs]
  };
  std::vector<PairBlocks> pairs;  // per world
  std::vector<int> flags;         // per world

  // candidates [iters][W*...]
  std::vector<std::vector<Tf<T>>> cand_q;
  std::vector<std::vector<Sp<T>>> cand_u;
  std::vector<std::vector<T>> cand_lam;
  std::vector<std::vector<T>> cand_merit;
  std::vector<T> merit;
  std::vector<int> nr_iters, pcr_iters, best;

  Batch(const ofb_model& m, const ofb_config& c) : M(m), cfg(c) {
    W = m.num_worlds; B = m.bodies_per_world; J = m.joints_per_world; D = m.dofs_per_world;
    S = m.shapes_per_world; G = m.num_global_shapes; SC = S + G; Cmax = c.max_contacts_per_world;
    // offsets from world 0 (core/model.py:83-168)
    joint_off.assign(J, -1); ctrl_off.assign(J, -1);
    Nj = 0; Nctrl = 0;
    for (int j = 0; j < J; ++j) {
      int t = m.joint_type[j];
      joint_off[j] = Nj;
      Nj += joint_row_count(t);
    }
    for (int j = 0; j < J; ++j) {
      int t = m.joint_type[j];
      if (t == 0 || t == 1) {
        int qs = m.joint_qd_start[j];
        if (m.joint_dof_mode[qs] != 0) { ctrl_off[j] = Nctrl; ++Nctrl; }
      }
    }
    Nc = Nj + Nctrl + 3 * Cmax;
    Nu = 6 * B;
    size_t WB = size_t(W) * B, WC = size_t(W) * Nc, WK = size_t(W) * Cmax;
    ccount.assign(W, 0); cs0.assign(WK, 0); cs1.assign(WK, 0);
    cp0.assign(WK * 3, 0); cp1.assign(WK * 3, 0); cn.assign(WK * 3, 0); ct0.assign(WK, 0); ct1.assign(WK, 0);
    q.resize(WB); q_prev.resize(WB);
    u.resize(WB); u_prev.resize(WB); f_ext.resize(WB); h_d.resize(WB); du.resize(WB); tmp_body.resize(WB);
    Iw_inv.resize(WB);
    tpos.assign(size_t(W) * D, 0); tvel.assign(size_t(W) * D, 0);
    lam.assign(WC, 0); lam_prev.assign(WC, 0); dlam.assign(WC, 0); h_c.assign(WC, 0); C.assign(WC, 0);
    active.assign(WC, 0); rhs.assign(WC, 0); Pinv.assign(WC, 0);
    dw_dmu.assign(WK, T(0));
    J0.assign(WC, spzero<T>()); J1.assign(WC, spzero<T>());
    b0.assign(WC, -1); b1.assign(WC, -1);
    r.assign(WC, 0); z.assign(WC, 0); Az.assign(WC, 0); p.assign(WC, 0); Ap.assign(WC, 0); y.assign(WC, 0);
    b_sq.assign(W, 0); r_sq.assign(W, 0); zAz.assign(W, 0); zAz_old.assign(W, 0); yAp.assign(W, 0);
    flags.assign(W, 0);
    int it = c.nr_max_iters;
    cand_q.assign(it, std::vector<Tf<T>>(WB)); cand_u.assign(it, std::vector<Sp<T>>(WB));
    cand_lam.assign(it, std::vector<T>(WC, 0)); cand_merit.assign(it, std::vector<T>(W, 0));
    merit.assign(W, 0); nr_iters.assign(W, 0); pcr_iters.assign(W, 0); best.assign(W, 0);
    // constraint body index for joint and control rows is fixed (fill kernels,
    // constraints/utils.py:460-527)
    for (int w = 0; w < W; ++w) {
      for (int j = 0; j < J; ++j) {
        int t = m.joint_type[w * J + j];
        int p_ = m.joint_parent[w * J + j], c_ = m.joint_child[w * J + j];
        for (int k = 0; k < joint_row_count(t); ++k) {
          size_t rr = size_t(w) * Nc + joint_off[j] + k;
          b0[rr] = p_; b1[rr] = c_;
          bool en = m.joint_enabled[w * J + j] != 0 && c_ >= 0;
          active[rr] = en ? T(1) : T(0);  // fill_joint_constraint_active_mask_kernel
        }
        if (ctrl_off[j] >= 0) {
          size_t rr = size_t(w) * Nc + Nj + ctrl_off[j];
          b0[rr] = p_; b1[rr] = c_;
        }
      }
    }
    if (c.preconditioner == OFB_PRECOND_PER_BODY_PAIR) {
      int np = (B + 1) * (B + 1);
      pairs.resize(W);
      for (auto& pb : pairs) {
        pb.count.assign(np, 0); pb.members.assign(size_t(np) * 64, 0);
        pb.L.assign(size_t(np) * 64 * 64, 0); pb.fail.assign(np, 0);
      }
    }
  }

  // --- model accessors ---
  T mass(int w, int b) const { return T(M.body_mass[w * B + b]); }
  T inv_mass(int w, int b) const { return T(M.body_inv_mass[w * B + b]); }
  M3<T> inertia(int w, int b) const { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = T(M.body_inertia[(w * B + b) * 9 + i]); return r; }
  M3<T> inv_inertia(int w, int b) const { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = T(M.body_inv_inertia[(w * B + b) * 9 + i]); return r; }
  V3<T> com(int w, int b) const { const float* c = M.body_com + (w * B + b) * 3; return v3<T>(c[0], c[1], c[2]); }
  Tf<T> jX(const float* base, int w, int j) const {
    const float* x = base + (w * J + j) * 7;
    return Tf<T>{v3<T>(x[0], x[1], x[2]), Qt<T>{T(x[3]), T(x[4]), T(x[5]), T(x[6])}};
  }
  V3<T> axis(int w, int d) const { const float* a = M.joint_axis + (w * D + d) * 3; return v3<T>(a[0], a[1], a[2]); }
  int shape_body(int w, int s) const { return M.shape_body[w * SC + s]; }
  T smu(int w, int s) const { return T(M.shape_mu[w * SC + s]); }
  V3<T> saxis(int w, int s) const { const float* a = M.shape_friction_axis + (w * SC + s) * 3; return v3<T>(a[0], a[1], a[2]); }

  size_t ri(int w, int row) const { return size_t(w) * Nc + row; }
  size_t bi(int w, int b) const { return size_t(w) * B + b; }
  V3<T> cvec(const std::vector<T>& a, int w, int c) const {
    size_t k = (size_t(w) * Cmax + c) * 3; return v3<T>(a[k], a[k + 1], a[k + 2]);
  }
};

// ----------------------------------------------------------------------------------------
// Contact packing: core/contacts.py:25-99 (serial thread order == input order)
// ----------------------------------------------------------------------------------------
template <class T> void pack_contacts(Batch<T>& Bt, const ofb_contacts& c) {
  std::fill(Bt.ccount.begin(), Bt.ccount.end(), 0);
  int n = *c.count;
  int G = Bt.G, S = Bt.S;
  for (int i = 0; i < c.max_contacts; ++i) {
    int s0 = c.shape0[i], s1 = c.shape1[i];
    if (i >= n || s0 == s1) continue;
    int w0 = s0 >= 0 ? Bt.M.shape_world[s0] : -1;
    int w1 = s1 >= 0 ? Bt.M.shape_world[s1] : -1;
    int w = w0 >= 0 ? w0 : (w1 >= 0 ? w1 : -1);
    if (w < 0) continue;
    int slot = Bt.ccount[w]++;
    if (slot >= Bt.Cmax) { Bt.flags[w] |= 2; continue; }
    size_t k = size_t(w) * Bt.Cmax + slot;
    for (int d = 0; d < 3; ++d) {
      Bt.cp0[k * 3 + d] = T(c.point0[i * 3 + d]);
      Bt.cp1[k * 3 + d] = T(c.point1[i * 3 + d]);
      Bt.cn[k * 3 + d] = -T(c.normal[i * 3 + d]);  // contacts.py:80-83
    }
    Bt.ct0[k] = T(c.margin0[i]);
    Bt.ct1[k] = T(c.margin1[i]);
    auto local = [&](int s) { return s < G ? S + s : (s - G) % S; };  // contacts.py:10-22
    Bt.cs0[k] = s0 >= 0 ? local(s0) : s0;
    Bt.cs1[k] = s1 >= 0 ? local(s1) : s1;
  }
}

// ----------------------------------------------------------------------------------------
// Contact reduction: collision/top_k.py, fps.py, cluster.py.  Runs on the packed contacts
// with the step-start poses (the reference calls it from load_data, where body_pose_prev
// holds the step's input pose, base_engine.py:317-320).  Each reducer is restated as the
// serial loop Warp's CPU device runs for its (world, contact) launch: leaders in contact
// order, then a scatter that reserves slots in contact order, so survivors keep their
// packed order.  A world holds min(count, Cmax) contacts (the reference loops to the raw
// count, past the arrays, when a world overflows).  The reducer's geometry is evaluated in
// double for both precisions of the oracle (and on the GPU): its decisions are
// comparisons, and the near-ties of flat contact patches would otherwise fall by rounding.
// ----------------------------------------------------------------------------------------
template <class T> struct RedContact {
  int b0, b1;
  T depth;
  V3<T> mid, n;
};

template <class T> RedContact<double> red_contact(const Batch<T>& Bt, const double* q, int w, int c) {
  const size_t k = size_t(w) * Bt.Cmax + c;
  RedContact<double> r;
  auto body = [&](int s) { return s < 0 ? -1 : Bt.M.shape_body[w * Bt.SC + s]; };  // _resolve_body
  r.b0 = body(Bt.cs0[k]);
  r.b1 = body(Bt.cs1[k]);
  auto pose = [&](int b) {
    if (b < 0) return tident<double>();
    const double* x = q + (size_t(w) * Bt.B + b) * 7;
    return Tf<double>{v3<double>(x[0], x[1], x[2]), Qt<double>{x[3], x[4], x[5], x[6]}};
  };
  r.n = v3<double>(Bt.cn[k * 3], Bt.cn[k * 3 + 1], Bt.cn[k * 3 + 2]);
  V3<double> p0 = v3<double>(Bt.cp0[k * 3], Bt.cp0[k * 3 + 1], Bt.cp0[k * 3 + 2]);
  V3<double> p1 = v3<double>(Bt.cp1[k * 3], Bt.cp1[k * 3 + 1], Bt.cp1[k * 3 + 2]);
  // _contact_depth / _depth_and_midpoint (top_k.py:55-87, fps.py:50-86)
  V3<double> pa = sub(tpoint(pose(r.b0), p0), mul(r.n, double(Bt.ct0[k])));
  V3<double> pb = add(tpoint(pose(r.b1), p1), mul(r.n, double(Bt.ct1[k])));
  r.depth = dot(r.n, sub(pb, pa));
  r.mid = mul(add(pa, pb), 0.5);
  return r;
}

template <class T> inline bool deeper(const RedContact<T>& a, int ia, const RedContact<T>& b, int ib) {
  return a.depth > b.depth || (a.depth == b.depth && ia < ib);
}

template <class T> void reduce_contacts(Batch<T>& Bt, const double* q) {
  const int policy = Bt.cfg.contact_reduction;
  if (policy == OFB_REDUCE_NONE) return;
  const int K = Bt.cfg.contact_max_per_pair;
  std::vector<RedContact<double>> C;
  std::vector<int> keep;
  for (int w = 0; w < Bt.W; ++w) {
    const int n = std::min(Bt.ccount[w], Bt.Cmax);
    C.resize(n);
    for (int c = 0; c < n; ++c) C[c] = red_contact(Bt, q, w, c);
    keep.assign(n, 0);
    auto same = [&](int i, int j) { return C[i].b0 == C[j].b0 && C[i].b1 == C[j].b1; };
    if (policy == OFB_REDUCE_TOP_K) {  // rank_within_pair_kernel (top_k.py:90-143)
      for (int c = 0; c < n; ++c) {
        int rank = 0;
        for (int j = 0; j < n; ++j)
          if (j != c && same(j, c) && deeper(C[j], j, C[c], c)) ++rank;
        keep[c] = rank < K;
      }
    } else {
      for (int c = 0; c < n; ++c) {
        bool leader = true;  // smallest contact index of its pair
        for (int j = 0; j < c; ++j)
          if (same(j, c)) leader = false;
        if (!leader) continue;
        if (policy == OFB_REDUCE_FPS) {  // fps_per_pair_kernel (fps.py:89-225)
          int seed = -1;
          double best = -1.0e30;
          for (int j = 0; j < n; ++j)
            if (same(j, c) && C[j].depth > best) { best = C[j].depth; seed = j; }
          if (seed < 0) continue;
          keep[seed] = 1;
          for (int r = 1; r < K; ++r) {
            int pick = -1;
            double pick_d = -1.0;
            for (int j = 0; j < n; ++j) {
              if (!same(j, c) || keep[j]) continue;
              double md = 1.0e30;
              for (int k = 0; k < n; ++k) {
                if (!keep[k] || !same(k, c)) continue;
                V3<double> d = sub(C[j].mid, C[k].mid);
                double dd = dot(d, d);
                if (dd < md) md = dd;
              }
              if (md > pick_d) { pick_d = md; pick = j; }
            }
            if (pick < 0) break;
            keep[pick] = 1;
          }
        } else {  // cluster_per_pair_kernel (cluster.py:94-238)
          const double pos_sq = double(Bt.cfg.cluster_pos) * double(Bt.cfg.cluster_pos);
          const double ndot = double(Bt.cfg.cluster_normal_dot);
          for (int j = 0; j < n; ++j) {  // pass 1: representatives of this pair
            if (!same(j, c)) continue;
            bool rep = true;
            for (int k = 0; k < n; ++k) {
              if (k == j || !same(k, c)) continue;
              V3<double> d = sub(C[j].mid, C[k].mid);
              if (dot(C[j].n, C[k].n) > ndot && dot(d, d) < pos_sq && deeper(C[k], k, C[j], j)) rep = false;
            }
            keep[j] = rep ? 1 : 0;
          }
          // pass 2, as the reference writes it: every contact of the world still marked
          // (this pair's and earlier pairs' representatives) is ranked against this
          // pair's marked representatives, in contact order, and dropped at rank >= K
          for (int j = 0; j < n; ++j) {
            if (!keep[j]) continue;
            int rank = 0;
            for (int k = 0; k < n; ++k)
              if (k != j && keep[k] && same(k, c) && deeper(C[k], k, C[j], j)) ++rank;
            if (rank >= K) keep[j] = 0;
          }
        }
      }
    }
    // scatter in contact order (scatter_top_k_kernel / scatter_by_keep_kernel)
    int m = 0;
    for (int c = 0; c < n; ++c) {
      if (!keep[c]) continue;
      const size_t src = size_t(w) * Bt.Cmax + c, dst = size_t(w) * Bt.Cmax + m;
      if (dst != src) {
        for (int d = 0; d < 3; ++d) {
          Bt.cp0[dst * 3 + d] = Bt.cp0[src * 3 + d];
          Bt.cp1[dst * 3 + d] = Bt.cp1[src * 3 + d];
          Bt.cn[dst * 3 + d] = Bt.cn[src * 3 + d];
        }
        Bt.ct0[dst] = Bt.ct0[src];
        Bt.ct1[dst] = Bt.ct1[src];
        Bt.cs0[dst] = Bt.cs0[src];
        Bt.cs1[dst] = Bt.cs1[src];
      }
      ++m;
    }
    Bt.ccount[w] = m;
  }
}

// ----------------------------------------------------------------------------------------
// Joint kinematics: constraints/joint_kinematics.py:5-261
// ----------------------------------------------------------------------------------------
template <class T> struct JointFrames {
  Tf<T> Xwp, Xwc;
  V3<T> r_p, r_c, pos_p, pos_c, com_p;  // com_p is the BODY-frame com (joint_constraint.py:113)
};

template <class T> inline void joint_transforms(const Tf<T>& body_q, V3<T> body_com, const Tf<T>& Xl,
                                                Tf<T>& Xw, V3<T>& r, V3<T>& pos) {
  Xw = tmul(body_q, Xl);
  V3<T> com_w = tpoint(body_q, body_com);
  pos = Xw.p;
  r = sub(pos, com_w);
}

template <class T> JointFrames<T> joint_frames(const Batch<T>& Bt, const std::vector<Tf<T>>& q, int w, int j) {
  JointFrames<T> F;
  int p_ = Bt.M.joint_parent[w * Bt.J + j], c_ = Bt.M.joint_child[w * Bt.J + j];
  joint_transforms(q[Bt.bi(w, c_)], Bt.com(w, c_), Bt.jX(Bt.M.joint_X_c, w, j), F.Xwc, F.r_c, F.pos_c);
  Tf<T> Xbp = tident<T>();
  V3<T> com_p = v3<T>(0, 0, 0);
  if (p_ >= 0) { Xbp = q[Bt.bi(w, p_)]; com_p = Bt.com(w, p_); }
  joint_transforms(Xbp, com_p, Bt.jX(Bt.M.joint_X_p, w, j), F.Xwp, F.r_p, F.pos_p);
  F.com_p = com_p;
  return F;
}

template <class T> inline V3<T> unit_axis(int i) {
  return i == 0 ? v3<T>(1, 0, 0) : (i == 1 ? v3<T>(0, 1, 0) : v3<T>(0, 0, 1));
}

// compute_joint_row (joint_kinematics.py:191-261)
template <class T> bool joint_row(int type, int row, const JointFrames<T>& F, V3<T> axis_local,
                                  Sp<T>& Jp, Sp<T>& Jc, T& err) {
  Jp = spzero<T>(); Jc = spzero<T>(); err = T(0);
  auto linear = [&](int ai) {  // get_linear_component :43-74
    V3<T> a = unit_axis<T>(ai);
    Jc = sp(a, cross(F.r_c, a));
    Jp = sp(neg(a), neg(cross(F.r_p, a)));
    err = comp(sub(F.pos_c, F.pos_p), ai);
  };
  auto angular = [&](int ai) {  // get_angular_component :77-113
    V3<T> aw = qrot(F.Xwp.q, unit_axis<T>(ai));
    Jc = sp(v3<T>(0, 0, 0), aw);
    Jp = sp(v3<T>(0, 0, 0), neg(aw));
    Qt<T> qr = qmul(qconj(F.Xwp.q), F.Xwc.q);
    T e = T(2) * (ai == 0 ? qr.x : (ai == 1 ? qr.y : qr.z));
    if (qr.w < T(0)) e = -e;
    err = e;
  };
  auto rev_angular = [&](int oi) {  // get_revolute_angular_component :116-147
    V3<T> b1, b2;
    orthogonal_basis(axis_local, b1, b2);
    V3<T> axc = qrot(F.Xwc.q, axis_local);
    V3<T> tgt = qrot(F.Xwp.q, oi == 0 ? b1 : b2);
    err = dot(axc, tgt);
    V3<T> ra = cross(axc, tgt);
    Jc = sp(v3<T>(0, 0, 0), ra);
    Jp = sp(v3<T>(0, 0, 0), neg(ra));
  };
  auto pris_linear = [&](int oi) {  // get_prismatic_linear_component :150-183
    V3<T> b1, b2;
    orthogonal_basis(axis_local, b1, b2);
    V3<T> nv = qrot(F.Xwp.q, oi == 0 ? b1 : b2);
    err = dot(sub(F.pos_c, F.pos_p), nv);
    Jc = sp(nv, cross(F.r_c, nv));
    V3<T> rpd = sub(F.pos_c, F.com_p);
    Jp = sp(neg(nv), cross(nv, rpd));
  };
  if (type == 0) {
    if (row < 2) { pris_linear(row); return true; }
    if (row < 5) { angular(row - 2); return true; }
  } else if (type == 1) {
    if (row < 3) { linear(row); return true; }
    if (row < 5) { rev_angular(row - 3); return true; }
  } else if (type == 2) {
    if (row < 3) { linear(row); return true; }
  } else if (type == 3) {
    if (row < 3) { linear(row); return true; }
    if (row < 6) { angular(row - 3); return true; }
  }
  return false;
}

// ----------------------------------------------------------------------------------------
// Row assembly (compute_linear_system, core/linear_system.py:140-341) and residual
// (compute_residual, core/residual.py:18-173).  `full` = write J, C, mask too.
// ----------------------------------------------------------------------------------------
template <class T> void dynamics_rows(Batch<T>& Bt, int w) {  // dynamics_constraint.py:6-50
  V3<T> g = v3<T>(Bt.M.gravity[0], Bt.M.gravity[1], Bt.M.gravity[2]);
  for (int b = 0; b < Bt.B; ++b) {
    size_t k = Bt.bi(w, b);
    T m = Bt.mass(w, b);
    M3<T> Iw = world_inertia(Bt.q[k].q, Bt.inertia(w, b));
    Sp<T> du_;
    for (int i = 0; i < 6; ++i) du_.v[i] = Bt.u[k].v[i] - Bt.u_prev[k].v[i];
    V3<T> wv = bot(Bt.u[k]);
    Sp<T> fg = sp(mul(g, m), v3<T>(0, 0, 0));
    Sp<T> fgy = sp(v3<T>(0, 0, 0), mul(cross(wv, mv(Iw, wv)), T(Bt.cfg.gyro_scale)));
    Sp<T> md = spatial_momentum(m, Iw, du_);
    for (int i = 0; i < 6; ++i)
      Bt.h_d[k].v[i] = md.v[i] - (Bt.f_ext[k].v[i] + fg.v[i] + fgy.v[i]) * Bt.dt;
  }
}

template <class T> inline void add_body(Batch<T>& Bt, int w, int b, const Sp<T>& d) {
  if (b < 0) return;
  Sp<T>& h = Bt.h_d[Bt.bi(w, b)];
  for (int i = 0; i < 6; ++i) h.v[i] += d.v[i];
}

template <class T> void joint_rows(Batch<T>& Bt, int w, bool full) {  // joint_constraint.py:62-249
  T dt = Bt.dt;
  for (int j = 0; j < Bt.J; ++j) {
    if (Bt.M.joint_enabled[w * Bt.J + j] == 0) continue;
    int type = Bt.M.joint_type[w * Bt.J + j];
    int p_ = Bt.M.joint_parent[w * Bt.J + j], c_ = Bt.M.joint_child[w * Bt.J + j];
    int off = Bt.joint_off[j];
    T comp_ = T(Bt.M.joint_compliance[w * Bt.J + j]);
    if (comp_ < T(0)) comp_ = T(Bt.cfg.compliance_joint);
    JointFrames<T> F = joint_frames(Bt, Bt.q, w, j);
    V3<T> ax = v3<T>(0, 0, 0);
    if (type == 0 || type == 1) ax = Bt.axis(w, Bt.M.joint_qd_start[w * Bt.J + j]);
    int rc = joint_rows_supported(type) ? joint_row_count(type) : 0;
    for (int i = 0; i < rc; ++i) {
      Sp<T> Jp, Jc; T err;
      bool act = joint_row(type, i, F, ax, Jp, Jc, err);
      size_t rr = Bt.ri(w, off + i);
      if (full) Bt.active[rr] = act ? T(1) : T(0);
      if (!act) continue;
      T l = Bt.lam[rr];
      if (p_ >= 0) add_body(Bt, w, p_, sscale(Jp, -l * dt));
      add_body(Bt, w, c_, sscale(Jc, -l * dt));
      Bt.h_c[rr] = (err + comp_ * l) / dt;
      if (full) { Bt.J0[rr] = Jp; Bt.J1[rr] = Jc; Bt.C[rr] = comp_ / (dt * dt); }
    }
  }
}

template <class T> void control_rows(Batch<T>& Bt, int w, bool full) {  // control_constraint.py:230-438
  T dt = Bt.dt;
  for (int j = 0; j < Bt.J; ++j) {
    if (Bt.M.joint_enabled[w * Bt.J + j] == 0) continue;
    int type = Bt.M.joint_type[w * Bt.J + j];
    if (type != 0 && type != 1) continue;
    int qs = Bt.M.joint_qd_start[w * Bt.J + j];
    int mode = Bt.M.joint_dof_mode[w * Bt.D + qs];
    if (mode == 0) continue;
    int p_ = Bt.M.joint_parent[w * Bt.J + j], c_ = Bt.M.joint_child[w * Bt.J + j];
    if (c_ < 0) continue;
    size_t rr = Bt.ri(w, Bt.Nj + Bt.ctrl_off[j]);
    if (full) Bt.active[rr] = T(1);
    JointFrames<T> F = joint_frames(Bt, Bt.q, w, j);
    V3<T> ax = Bt.axis(w, qs);
    Sp<T> up = spzero<T>();
    if (p_ >= 0) up = Bt.u[Bt.bi(w, p_)];
    Sp<T> uc = Bt.u[Bt.bi(w, c_)];
    T target = T(0);
    if (mode == 1) target = Bt.tpos[size_t(w) * Bt.D + qs];
    else if (mode == 2) target = Bt.tvel[size_t(w) * Bt.D + qs];
    T ke = T(Bt.M.joint_target_ke[w * Bt.D + qs]), kd = T(Bt.M.joint_target_kd[w * Bt.D + qs]);
    T l = Bt.lam[rr];
    Sp<T> Jp, Jc;
    T qv = 0, qdv = 0;
    bool angular = false;
    if (type == 1) {  // compute_revolute_q_qd :9-38, compute_revolute_jacobians :79-90
      Qt<T> qr = qmul(qconj(F.Xwp.q), F.Xwc.q);
      T sh = qr.x * ax.x + qr.y * ax.y + qr.z * ax.z;
      qv = T(2) * std::atan2(sh, qr.w);
      V3<T> aw = qrot(F.Xwp.q, ax);
      qdv = dot(sub(bot(uc), bot(up)), aw);
      Jc = sp(v3<T>(0, 0, 0), aw);
      Jp = sp(v3<T>(0, 0, 0), neg(aw));
      angular = true;
    } else {  // compute_prismatic_q_qd :41-73, compute_prismatic_jacobians :93-111
      V3<T> aw = qrot(F.Xwp.q, ax);
      V3<T> delta = sub(F.pos_c, F.pos_p);
      qv = dot(delta, aw);
      V3<T> vc = add(top(uc), cross(bot(uc), F.r_c));
      V3<T> vp = add(top(up), cross(bot(up), F.r_p));
      qdv = dot(sub(vc, vp), aw) + dot(delta, cross(bot(up), aw));
      Jc = sp(aw, cross(F.r_c, aw));
      V3<T> rpd = sub(F.pos_c, F.com_p);
      Jp = sp(neg(aw), neg(cross(rpd, aw)));
    }
    // compute_control_properties :117-166
    T err = 0, alpha = 0;
    if (mode == 1) {
      T raw = qv - target;
      if (angular) {
        const T PI = T(3.141592653589793), TWO_PI = T(6.283185307179586);
        T sh = raw + PI;
        T md = sh - TWO_PI * std::floor(sh / TWO_PI);
        err = md - PI;
      } else {
        err = raw;
      }
      err = err / dt;
      T den = dt * dt * ke + dt * kd;
      alpha = den > T(1e-6) ? T(1) / den : T(1.0e8);
    } else if (mode == 2) {
      err = qdv - target;
      T den = dt * ke;
      alpha = den > T(1e-6) ? T(1) / den : T(1.0e8);
    }
    if (p_ >= 0) add_body(Bt, w, p_, sscale(Jp, -l * dt));
    add_body(Bt, w, c_, sscale(Jc, -l * dt));
    Bt.h_c[rr] = err + alpha * l * dt;
    if (full) { Bt.J0[rr] = Jp; Bt.J1[rr] = Jc; Bt.C[rr] = alpha; }
  }
}

template <class T> struct BodyView {
  Tf<T> pose, pose_prev; Sp<T> vel; T m_inv; M3<T> I_inv; V3<T> com;
};

template <class T> BodyView<T> body_view(const Batch<T>& Bt, int w, int b) {
  BodyView<T> v;
  v.pose = tident<T>(); v.pose_prev = tident<T>(); v.vel = spzero<T>(); v.m_inv = T(0);
  for (int i = 0; i < 9; ++i) v.I_inv.m[i] = T(0);
  v.com = v3<T>(0, 0, 0);
  if (b >= 0) {
    size_t k = Bt.bi(w, b);
    v.pose = Bt.q[k]; v.pose_prev = Bt.q_prev[k]; v.vel = Bt.u[k];
    v.m_inv = Bt.inv_mass(w, b); v.I_inv = Bt.inv_inertia(w, b); v.com = Bt.com(w, b);
  }
  return v;
}

// compute_effective_mass (constraints/utils.py:6-33), inertia rotated by pose_prev
template <class T> T effective_mass(const BodyView<T>& A, const BodyView<T>& Bv, const Sp<T>& J0,
                                    const Sp<T>& J1, int b0, int b1) {
  T val = 0;
  if (b0 >= 0) { M3<T> Iw = world_inertia(A.pose_prev.q, A.I_inv); val += sdot(J0, spatial_momentum(A.m_inv, Iw, J0)); }
  if (b1 >= 0) { M3<T> Iw = world_inertia(Bv.pose_prev.q, Bv.I_inv); val += sdot(J1, spatial_momentum(Bv.m_inv, Iw, J1)); }
  return val;
}

template <class T> void contact_rows(Batch<T>& Bt, int w, bool full) {  // contact_constraint.py:48-374
  T dt = Bt.dt;
  int on = Bt.Nj + Bt.Nctrl;
  for (int c = 0; c < Bt.Cmax; ++c) {
    size_t rr = Bt.ri(w, on + c);
    size_t k = size_t(w) * Bt.Cmax + c;
    int s0 = Bt.cs0[k], s1 = Bt.cs1[k];
    if (c >= Bt.ccount[w] || s0 == s1) {
      if (full) {
        Bt.active[rr] = 0; Bt.lam[rr] = 0; Bt.J0[rr] = spzero<T>(); Bt.J1[rr] = spzero<T>(); Bt.C[rr] = 0;
      }
      Bt.h_c[rr] = 0;
      continue;
    }
    int bb0 = s0 >= 0 ? Bt.shape_body(w, s0) : -1;
    int bb1 = s1 >= 0 ? Bt.shape_body(w, s1) : -1;
    V3<T> n = Bt.cvec(Bt.cn, w, c);
    T fn = Bt.lam[rr];
    BodyView<T> A = body_view(Bt, w, bb0), Bv = body_view(Bt, w, bb1);
    V3<T> p0l = Bt.cvec(Bt.cp0, w, c), p1l = Bt.cvec(Bt.cp1, w, c);
    T th0 = Bt.ct0[k], th1 = Bt.ct1[k];
    // compute_contact_core :48-138
    V3<T> p0_adj_prev = sub(tpoint(A.pose_prev, p0l), mul(n, th0));
    V3<T> p1_adj_prev = add(tpoint(Bv.pose_prev, p1l), mul(n, th1));
    V3<T> p0_adj = sub(tpoint(A.pose, p0l), mul(n, th0));
    V3<T> p1_adj = add(tpoint(Bv.pose, p1l), mul(n, th1));
    Sp<T> Ja = spzero<T>(), Jb = spzero<T>();
    if (bb0 >= 0) { V3<T> lev = sub(p0_adj_prev, tpoint(A.pose_prev, A.com)); Ja = sp(n, cross(lev, n)); }
    if (bb1 >= 0) { V3<T> lev = sub(p1_adj_prev, tpoint(Bv.pose_prev, Bv.com)); Jb = sp(neg(n), cross(lev, neg(n))); }
    T gap = dot(n, sub(p0_adj, p1_adj));
    T em = effective_mass(A, Bv, Ja, Jb, bb0, bb1);
    T precond = dt * dt * em;
    T phi, dphi_dg, dphi_dl;
    if (Bt.cfg.contact_mode == OFB_CONTACT_EQUALITY) { phi = gap; dphi_dg = 1; dphi_dl = 0; }
    else fb_diff(gap, fn, T(Bt.cfg.contact_fb_alpha), precond, T(Bt.cfg.contact_fb_eps_sq), phi, dphi_dg, dphi_dl);
    Sp<T> Jh0 = sscale(Ja, dphi_dg), Jh1 = sscale(Jb, dphi_dg);
    if (full) { Bt.active[rr] = 1; Bt.b0[rr] = bb0; Bt.b1[rr] = bb1; }
    if (bb0 >= 0) add_body(Bt, w, bb0, sscale(Jh0, -fn * dt));
    if (bb1 >= 0) add_body(Bt, w, bb1, sscale(Jh1, -fn * dt));
    Bt.h_c[rr] = phi / dt;
    if (full) { Bt.C[rr] = (dphi_dl + T(Bt.cfg.compliance_contact)) / (dt * dt); Bt.J0[rr] = Jh0; Bt.J1[rr] = Jh1; }
  }
}

// resolve_friction_frame (friction_constraint.py:79-156)
template <class T> void friction_frame(V3<T> n, V3<T> ax0, const Tf<T>& pp0, int b0, V3<T> ax1, const Tf<T>& pp1,
                                       int b1, T mu0, T mp0, T mu1, T mp1, V3<T>& t1, V3<T>& t2, T& mux, T& muy) {
  const T eps_len = T(1e-6);
  bool an0 = dot(ax0, ax0) > eps_len && b0 >= 0;
  bool an1 = dot(ax1, ax1) > eps_len && b1 >= 0;
  if (!an0 && !an1) {
    orthogonal_basis(n, t1, t2);
    T mu = (mu0 + mu1) * T(0.5);
    mux = mu; muy = mu;
    return;
  }
  V3<T> aw; T mal, mpa, mot;
  if (an0) { aw = qrot(pp0.q, ax0); mal = mu0; mpa = mp0 >= T(0) ? mp0 : mu0; mot = mu1; }
  else { aw = qrot(pp1.q, ax1); mal = mu1; mpa = mp1 >= T(0) ? mp1 : mu1; mot = mu0; }
  V3<T> at = sub(aw, mul(n, dot(aw, n)));
  T tl2 = dot(at, at);
  if (tl2 < T(1e-8)) {
    orthogonal_basis(n, t1, t2);
    T mi = (mal + mot) * T(0.5);
    mux = mi; muy = mi;
    return;
  }
  t1 = mul(at, T(1) / std::sqrt(tl2));
  t2 = cross(n, t1);
  mux = (mal + mot) * T(0.5);
  muy = (mpa + mot) * T(0.5);
}

template <class T> void friction_rows(Batch<T>& Bt, int w, bool full) {  // friction_constraint.py:379-908
  T dt = Bt.dt;
  int on = Bt.Nj + Bt.Nctrl, of = on + Bt.Cmax;
  const int SCn = Bt.SC;
  for (int c = 0; c < Bt.Cmax; ++c) {
    size_t r0 = Bt.ri(w, of + 2 * c), r1 = r0 + 1;
    size_t k = size_t(w) * Bt.Cmax + c;
    int s0 = Bt.cs0[k], s1 = Bt.cs1[k];
    auto deactivate = [&]() {
      Bt.dw_dmu[k] = 0;
      if (full) {
        Bt.active[r0] = 0; Bt.active[r1] = 0; Bt.lam[r0] = 0; Bt.lam[r1] = 0;
        Bt.J0[r0] = spzero<T>(); Bt.J0[r1] = spzero<T>(); Bt.J1[r0] = spzero<T>(); Bt.J1[r1] = spzero<T>();
        Bt.C[r0] = 0; Bt.C[r1] = 0;
      }
      Bt.h_c[r0] = 0; Bt.h_c[r1] = 0;
    };
    if (c >= Bt.ccount[w]) { deactivate(); continue; }
    bool skip = s0 == s1;
    if (Bt.cfg.friction_mode == OFB_FRICTION_BILATERAL_PATCH) {  // bilateral_row_is_skipped :39-76
      for (int jj = 0; jj < c; ++jj) {
        size_t kk = size_t(w) * Bt.Cmax + jj;
        if (Bt.cs0[kk] == s0 && Bt.cs1[kk] == s1) { skip = true; break; }
      }
    }
    if (skip) { deactivate(); continue; }
    T mu0 = Bt.smu(w, s0), mu1 = Bt.smu(w, s1);
    T fn_prev = Bt.lam_prev[Bt.ri(w, on + c)];
    T mu_max = std::max(mu0, mu1);
    if (mu_max * fn_prev <= T(Bt.cfg.friction_load_gate)) { deactivate(); continue; }
    int bb0 = s0 >= 0 ? Bt.shape_body(w, s0) : -1;
    int bb1 = s1 >= 0 ? Bt.shape_body(w, s1) : -1;
    if (full) {
      Bt.active[r0] = 1; Bt.active[r1] = 1;
      Bt.b0[r0] = bb0; Bt.b1[r0] = bb1; Bt.b0[r1] = bb0; Bt.b1[r1] = bb1;
    }
    V3<T> n = Bt.cvec(Bt.cn, w, c);
    BodyView<T> A = body_view(Bt, w, bb0), Bv = body_view(Bt, w, bb1);
    V3<T> t1, t2; T mux, muy;
    friction_frame(n, Bt.saxis(w, s0), A.pose_prev, bb0, Bt.saxis(w, s1), Bv.pose_prev, bb1, mu0,
                   T(Bt.M.shape_mu_perp[w * SCn + s0]), mu1, T(Bt.M.shape_mu_perp[w * SCn + s1]), t1, t2, mux, muy);
    // resolve_stribeck_params :159-176, resolve_stribeck_lateral_only :179-194
    T ss0 = T(Bt.M.shape_mu_stiction_scale[w * SCn + s0]), vs0 = T(Bt.M.shape_v_stribeck[w * SCn + s0]);
    T ss1 = T(Bt.M.shape_mu_stiction_scale[w * SCn + s1]), vs1 = T(Bt.M.shape_v_stribeck[w * SCn + s1]);
    T mss = -1, vst = -1;
    if (ss0 > T(0) && vs0 > T(0)) { mss = ss0; vst = vs0; }
    else if (ss1 > T(0) && vs1 > T(0)) { mss = ss1; vst = vs1; }
    T lo0 = T(Bt.M.shape_stribeck_lateral_only[w * SCn + s0]), lo1 = T(Bt.M.shape_stribeck_lateral_only[w * SCn + s1]);
    T lat = lo0 > T(0) ? lo0 : (lo1 > T(0) ? lo1 : T(-1));
    T l1 = Bt.lam[r0], l2 = Bt.lam[r1];
    T l1p = Bt.lam_prev[r0], l2p = Bt.lam_prev[r1];
    V3<T> p0l = Bt.cvec(Bt.cp0, w, c), p1l = Bt.cvec(Bt.cp1, w, c);
    T th0 = Bt.ct0[k], th1 = Bt.ct1[k];
    // compute_friction_core :379-469
    Sp<T> Jt1a = spzero<T>(), Jt2a = spzero<T>(), Jt1b = spzero<T>(), Jt2b = spzero<T>();
    if (bb0 >= 0) {
      V3<T> r_ = sub(sub(tpoint(A.pose_prev, p0l), mul(n, th0)), tpoint(A.pose_prev, A.com));
      Jt1a = sp(t1, cross(r_, t1)); Jt2a = sp(t2, cross(r_, t2));
    }
    if (bb1 >= 0) {
      V3<T> r_ = sub(add(tpoint(Bv.pose_prev, p1l), mul(n, th1)), tpoint(Bv.pose_prev, Bv.com));
      Jt1b = sp(neg(t1), neg(cross(r_, t1))); Jt2b = sp(neg(t2), neg(cross(r_, t2)));
    }
    T wt1 = effective_mass(A, Bv, Jt1a, Jt1b, bb0, bb1);
    T wt2 = effective_mass(A, Bv, Jt2a, Jt2b, bb0, bb1);
    T rr_ = (wt1 + wt2) * T(0.5);
    // compute_friction_model :197-371
    T vt0 = sdot(Jt1a, A.vel) + sdot(Jt1b, Bv.vel);
    T vt1 = sdot(Jt2a, A.vel) + sdot(Jt2b, Bv.vel);
    if (mss > T(0) && vst > T(0)) {
      T s = std::sqrt(vt0 * vt0 + vt1 * vt1);
      T factor = T(1) + (mss - T(1)) * std::exp(-s / vst);
      if (lat > T(0) && mux != muy) { mux = mux * factor; }
      else { mux = mux * factor; muy = muy * factor; }
    }
    T wx, wy;
    if (Bt.cfg.friction_mode != OFB_FRICTION_CONE) {
      wx = 0; wy = 0;
      Bt.dw_dmu[k] = 0;
    } else {
      const T eps = T(1e-8);
      T denom_eps = T(1e-6) * dt;
      if (mux == muy) {
        T mu = mux;
        T d0 = vt0 * dt, d1 = vt1 * dt;
        T dn = std::sqrt(d0 * d0 + d1 * d1 + eps);
        T i0 = l1p * dt, i1 = l2p * dt;
        T raw = std::sqrt(i0 * i0 + i1 * i1);
        T lim = mu * (fn_prev * dt);
        T clamped = std::min(raw, lim);
        T gap = lim - clamped;
        T phi = fb(dn, gap, T(1), rr_);
        T den = rr_ * raw + phi + denom_eps;
        T num = dn - phi;
        T wv = rr_ * (num / den);
        // d wv / d mu for the adjoint: lim = mu fn_prev dt enters through gap
        {
          T sb = rr_ * gap, nrm = std::sqrt(dn * dn + sb * sb);
          T dgap = raw < lim ? T(1) : T(0);
          T dphi = rr_ * (T(1) - (nrm > T(0) ? sb / nrm : T(0))) * dgap * fn_prev * dt;
          T dw = (wv < T(0) || wv > T(1e5)) ? T(0) : -rr_ * dphi * (den + num) / (den * den);
          V3<T> a0 = Bt.saxis(w, s0), a1 = Bt.saxis(w, s1);
          bool aniso = (dot(a0, a0) > T(1e-6) && bb0 >= 0) || (dot(a1, a1) > T(1e-6) && bb1 >= 0);
          bool strib = mss > T(0) && vst > T(0);
          Bt.dw_dmu[k] = (aniso || strib) ? std::numeric_limits<T>::quiet_NaN() : dw;
        }
        wv = std::max(wv, T(0));
        wv = std::min(wv, T(1e5));
        wx = wv; wy = wv;
      } else {
        const T mu_floor = T(1e-6);
        T mxs = std::max(mux, mu_floor), mys = std::max(muy, mu_floor);
        T d0 = vt0 * dt, d1 = vt1 * dt;
        T dxt = mxs * d0, dyt = mys * d1;
        T dn = std::sqrt(dxt * dxt + dyt * dyt + eps);
        T i0 = l1p * dt, i1 = l2p * dt;
        T imx = T(1) / mxs, imy = T(1) / mys;
        T fxt = i0 * imx, fyt = i1 * imy;
        T raw = std::sqrt(fxt * fxt + fyt * fyt);
        T lim = fn_prev * dt;
        T clamped = std::min(raw, lim);
        T gap = lim - clamped;
        T phi = fb(dn, gap, T(1), rr_);
        T den = rr_ * raw + phi + denom_eps;
        T num = dn - phi;
        T wt = rr_ * (num / den);
        wt = std::max(wt, T(0));
        wt = std::min(wt, T(1e5));
        wx = wt * imx * imx; wy = wt * imy * imy;
        Bt.dw_dmu[k] = std::numeric_limits<T>::quiet_NaN();
      }
    }
    Sp<T> d0 = spzero<T>(), d1 = spzero<T>();
    for (int i = 0; i < 6; ++i) {
      d0.v[i] = -dt * (Jt1a.v[i] * l1 + Jt2a.v[i] * l2);
      d1.v[i] = -dt * (Jt1b.v[i] * l1 + Jt2b.v[i] * l2);
    }
    if (bb0 >= 0) add_body(Bt, w, bb0, d0);
    if (bb1 >= 0) add_body(Bt, w, bb1, d1);
    Bt.h_c[r0] = vt0 + wx * l1;
    Bt.h_c[r1] = vt1 + wy * l2;
    if (full) {
      Bt.J0[r0] = Jt1a; Bt.J0[r1] = Jt2a; Bt.J1[r0] = Jt1b; Bt.J1[r1] = Jt2b;
      Bt.C[r0] = wx / dt + T(Bt.cfg.compliance_friction);
      Bt.C[r1] = wy / dt + T(Bt.cfg.compliance_friction);
    }
  }
}

template <class T> void world_inv_inertia(Batch<T>& Bt, int w) {  // linear_system.py:120-137
  for (int b = 0; b < Bt.B; ++b) Bt.Iw_inv[Bt.bi(w, b)] = world_inertia(Bt.q[Bt.bi(w, b)].q, Bt.inv_inertia(w, b));
}

template <class T> Sp<T> minv(const Batch<T>& Bt, int w, int b, const Sp<T>& v) {
  return spatial_momentum(Bt.inv_mass(w, b), Bt.Iw_inv[Bt.bi(w, b)], v);
}

template <class T> void linear_system(Batch<T>& Bt, int w) {  // linear_system.py:140-341
  world_inv_inertia(Bt, w);
  dynamics_rows(Bt, w);
  joint_rows(Bt, w, true);
  control_rows(Bt, w, true);
  contact_rows(Bt, w, true);
  friction_rows(Bt, w, true);
  for (int i = 0; i < Bt.Nc; ++i) {  // compute_schur_complement_rhs_kernel :17-65
    size_t rr = Bt.ri(w, i);
    if (Bt.active[rr] == T(0)) { Bt.rhs[rr] = 0; continue; }
    T s = 0;
    if (Bt.b0[rr] >= 0) s += sdot(Bt.J0[rr], minv(Bt, w, Bt.b0[rr], Bt.h_d[Bt.bi(w, Bt.b0[rr])]));
    if (Bt.b1[rr] >= 0) s += sdot(Bt.J1[rr], minv(Bt, w, Bt.b1[rr], Bt.h_d[Bt.bi(w, Bt.b1[rr])]));
    Bt.rhs[rr] = (s - Bt.h_c[rr]) / Bt.dt;
  }
}

template <class T> void residual(Batch<T>& Bt, int w) {  // core/residual.py:18-173
  dynamics_rows(Bt, w);
  joint_rows(Bt, w, false);
  control_rows(Bt, w, false);
  contact_rows(Bt, w, false);
  friction_rows(Bt, w, false);
}

template <class T> T merit_of(const Batch<T>& Bt, int w) {  // TiledSqNorm over res.full
  T s = 0;
  for (int b = 0; b < Bt.B; ++b)
    for (int i = 0; i < 6; ++i) { T v = Bt.h_d[Bt.bi(w, b)].v[i]; s += v * v; }
  for (int i = 0; i < Bt.Nc; ++i) { T v = Bt.h_c[Bt.ri(w, i)]; s += v * v; }
  return s;
}

// ----------------------------------------------------------------------------------------
// Operator, preconditioners, PCR (optim/*.py)
// ----------------------------------------------------------------------------------------
// z = beta*y + alpha*(J M^-1 J^T x + (C + reg) x)  (system_operator.py:36-184)
template <class T> void matvec(Batch<T>& Bt, int w, const std::vector<T>& x, const std::vector<T>& yv,
                               std::vector<T>& zout, T alpha, T beta) {
  for (int b = 0; b < Bt.B; ++b) Bt.tmp_body[Bt.bi(w, b)] = spzero<T>();
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    if (Bt.active[rr] == T(0)) continue;
    T xi = x[rr];
    if (Bt.b0[rr] >= 0) { Sp<T> dv = minv(Bt, w, Bt.b0[rr], sscale(Bt.J0[rr], xi)); Sp<T>& t = Bt.tmp_body[Bt.bi(w, Bt.b0[rr])]; for (int k = 0; k < 6; ++k) t.v[k] += dv.v[k]; }
    if (Bt.b1[rr] >= 0) { Sp<T> dv = minv(Bt, w, Bt.b1[rr], sscale(Bt.J1[rr], xi)); Sp<T>& t = Bt.tmp_body[Bt.bi(w, Bt.b1[rr])]; for (int k = 0; k < 6; ++k) t.v[k] += dv.v[k]; }
  }
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    T Ax = 0;
    if (Bt.active[rr] > T(0)) {
      T jv = 0;
      if (Bt.b0[rr] >= 0) jv += sdot(Bt.J0[rr], Bt.tmp_body[Bt.bi(w, Bt.b0[rr])]);
      if (Bt.b1[rr] >= 0) jv += sdot(Bt.J1[rr], Bt.tmp_body[Bt.bi(w, Bt.b1[rr])]);
      Ax = jv + (Bt.C[rr] + T(Bt.cfg.regularization)) * x[rr];
    }
    T res = alpha * Ax;
    if (beta != T(0)) res += beta * yv[rr];
    zout[rr] = res;
  }
}

template <class T> void jacobi_update(Batch<T>& Bt, int w) {  // preconditioner.py:9-50
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    if (Bt.active[rr] == T(0)) continue;
    T s = 0;
    if (Bt.b0[rr] >= 0) s += sdot(Bt.J0[rr], minv(Bt, w, Bt.b0[rr], Bt.J0[rr]));
    if (Bt.b1[rr] >= 0) s += sdot(Bt.J1[rr], minv(Bt, w, Bt.b1[rr], Bt.J1[rr]));
    T dA = s + Bt.C[rr] + T(Bt.cfg.regularization);
    Bt.Pinv[rr] = T(1) / (dA + T(1e-6));
  }
}

// per_body_pair_preconditioner.py:62-321 (pair ids, member lists in row order,
// dense blocks, column Cholesky with failure flag)
template <class T> void pair_update(Batch<T>& Bt, int w) {
  auto& pb = Bt.pairs[w];
  const int NB = Bt.B, np = (NB + 1) * (NB + 1);
  std::fill(pb.count.begin(), pb.count.end(), 0);
  std::fill(pb.fail.begin(), pb.fail.end(), 0);
  std::vector<int> total(np, 0);
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    if (Bt.active[rr] == T(0)) continue;
    int e1 = Bt.b0[rr] < 0 ? NB : Bt.b0[rr], e2 = Bt.b1[rr] < 0 ? NB : Bt.b1[rr];
    int lo = std::min(e1, e2), hi = std::max(e1, e2);
    int pid = lo * (NB + 1) + hi;
    int slot = total[pid]++;
    if (slot < 64) { pb.members[size_t(pid) * 64 + slot] = i; pb.count[pid] = slot + 1; }
    else Bt.flags[w] |= 8;  // overflow members fall back to Jacobi
  }
  for (int pid = 0; pid < np; ++pid) {
    int n = pb.count[pid];
    if (n == 0) continue;
    int stride = NB + 1, bhi = pid % stride, blo = pid / stride;
    std::vector<T> A(size_t(n) * n, 0);
    auto jfor = [&](int tb, int row) -> Sp<T> {
      size_t rr = Bt.ri(w, row);
      if (tb >= NB) return spzero<T>();
      if (Bt.b0[rr] == tb) return Bt.J0[rr];
      if (Bt.b1[rr] == tb) return Bt.J1[rr];
      return spzero<T>();
    };
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        int ci = pb.members[size_t(pid) * 64 + i], cj = pb.members[size_t(pid) * 64 + j];
        T val = 0;
        if (blo < NB) val += sdot(jfor(blo, ci), minv(Bt, w, blo, jfor(blo, cj)));
        if (bhi < NB) val += sdot(jfor(bhi, ci), minv(Bt, w, bhi, jfor(bhi, cj)));
        if (i == j) val += Bt.C[Bt.ri(w, ci)] + T(Bt.cfg.regularization);
        A[size_t(i) * n + j] = val;
      }
    T* L = &pb.L[size_t(pid) * 64 * 64];
    std::fill(L, L + 64 * 64, T(0));
    for (int kk = 0; kk < n; ++kk) {
      T ds = 0;
      for (int qq = 0; qq < kk; ++qq) ds += L[kk * 64 + qq] * L[kk * 64 + qq];
      T dv = A[size_t(kk) * n + kk] - ds;
      if (dv <= T(0)) { pb.fail[pid] = 1; Bt.flags[w] |= 4; break; }
      T Lkk = std::sqrt(dv);
      L[kk * 64 + kk] = Lkk;
      T inv = T(1) / Lkk;
      for (int i = kk + 1; i < n; ++i) {
        T rs = 0;
        for (int qq = 0; qq < kk; ++qq) rs += L[i * 64 + qq] * L[kk * 64 + qq];
        L[i * 64 + kk] = (A[size_t(i) * n + kk] - rs) * inv;
      }
    }
  }
  jacobi_update(Bt, w);  // _update_jacobi_fallback :606-626
}

// z = beta*y + alpha*P x
template <class T> void precond_apply(Batch<T>& Bt, int w, const std::vector<T>& x, const std::vector<T>& yv,
                                      std::vector<T>& zout, T alpha, T beta) {
  if (Bt.cfg.preconditioner != OFB_PRECOND_PER_BODY_PAIR) {  // preconditioner.py:53-80
    for (int i = 0; i < Bt.Nc; ++i) {
      size_t rr = Bt.ri(w, i);
      T px = 0;
      if (Bt.active[rr] > T(0)) px = Bt.Pinv[rr] * x[rr];
      zout[rr] = beta == T(0) ? alpha * px : beta * yv[rr] + alpha * px;
    }
    return;
  }
  // per_body_pair_preconditioner.py:324-410 (+ overflow rows use the Jacobi fallback)
  auto& pb = Bt.pairs[w];
  const int NB = Bt.B, np = (NB + 1) * (NB + 1);
  std::vector<T> out(Bt.Nc);
  for (int i = 0; i < Bt.Nc; ++i) out[i] = beta * yv[Bt.ri(w, i)];
  std::vector<char> covered(Bt.Nc, 0);
  T ws[64];
  for (int pid = 0; pid < np; ++pid) {
    int n = pb.count[pid];
    if (n == 0) continue;
    const int* mem = &pb.members[size_t(pid) * 64];
    for (int kk = 0; kk < n; ++kk) covered[mem[kk]] = 1;
    if (pb.fail[pid]) {
      for (int kk = 0; kk < n; ++kk) { int ci = mem[kk]; out[ci] += alpha * Bt.Pinv[Bt.ri(w, ci)] * x[Bt.ri(w, ci)]; }
      continue;
    }
    const T* L = &pb.L[size_t(pid) * 64 * 64];
    for (int kk = 0; kk < n; ++kk) ws[kk] = x[Bt.ri(w, mem[kk])];
    for (int i = 0; i < n; ++i) {
      T s = ws[i];
      for (int j = 0; j < i; ++j) s -= L[i * 64 + j] * ws[j];
      ws[i] = s / L[i * 64 + i];
    }
    for (int ii = 0; ii < n; ++ii) {
      int i = n - 1 - ii;
      T s = ws[i];
      for (int j = i + 1; j < n; ++j) s -= L[j * 64 + i] * ws[j];
      ws[i] = s / L[i * 64 + i];
    }
    for (int kk = 0; kk < n; ++kk) out[mem[kk]] += alpha * ws[kk];
  }
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    if (!covered[i] && Bt.active[rr] > T(0)) out[i] += alpha * Bt.Pinv[rr] * x[rr];
    zout[rr] = out[i];
  }
}

template <class T> T wdot(const Batch<T>& Bt, int w, const std::vector<T>& a, const std::vector<T>& b) {
  T s = 0;
  for (int i = 0; i < Bt.Nc; ++i) s += a[Bt.ri(w, i)] * b[Bt.ri(w, i)];
  return s;
}

// PCRSolver.solve (optim/pcr_solver.py:170-295) over the worlds in `ws`.
template <class T> void pcr_solve(Batch<T>& Bt, const std::vector<int>& ws, bool global) {
  T tol_sq = Bt.cfg.lin_tol >= 0 ? T(Bt.cfg.lin_tol) * T(Bt.cfg.lin_tol) : T(-1);
  T atol_sq = Bt.cfg.lin_atol >= 0 ? T(Bt.cfg.lin_atol) * T(Bt.cfg.lin_atol) : T(-1);
  for (int w : ws) {
    Bt.b_sq[w] = wdot(Bt, w, Bt.rhs, Bt.rhs);
    matvec(Bt, w, Bt.dlam, Bt.rhs, Bt.r, T(-1), T(1));
    precond_apply(Bt, w, Bt.r, Bt.z, Bt.z, T(1), T(0));
    matvec(Bt, w, Bt.z, Bt.Az, Bt.Az, T(1), T(0));
    for (int i = 0; i < Bt.Nc; ++i) { size_t rr = Bt.ri(w, i); Bt.p[rr] = Bt.z[rr]; Bt.Ap[rr] = Bt.Az[rr]; }
    Bt.zAz[w] = wdot(Bt, w, Bt.z, Bt.Az);
    Bt.r_sq[w] = wdot(Bt, w, Bt.r, Bt.r);
  }
  std::vector<char> run(Bt.W, 0);
  for (int w : ws) run[w] = 1;
  int iter = 0;
  while (true) {
    bool any = false;
    for (int w : ws) {
      if (!run[w]) continue;
      Bt.pcr_iters[w] += 1;
      Bt.zAz_old[w] = Bt.zAz[w];
      precond_apply(Bt, w, Bt.Ap, Bt.y, Bt.y, T(1), T(0));
      Bt.yAp[w] = wdot(Bt, w, Bt.y, Bt.Ap);
      T a = Bt.yAp[w] > T(0) ? Bt.zAz_old[w] / Bt.yAp[w] : T(0);
      for (int i = 0; i < Bt.Nc; ++i) {
        size_t rr = Bt.ri(w, i);
        Bt.dlam[rr] += a * Bt.p[rr];
        Bt.r[rr] -= a * Bt.Ap[rr];
        Bt.z[rr] -= a * Bt.y[rr];
      }
      matvec(Bt, w, Bt.z, Bt.Az, Bt.Az, T(1), T(0));
      Bt.zAz[w] = wdot(Bt, w, Bt.z, Bt.Az);
      T bt = Bt.zAz_old[w] > T(0) ? Bt.zAz[w] / Bt.zAz_old[w] : T(0);
      for (int i = 0; i < Bt.Nc; ++i) {
        size_t rr = Bt.ri(w, i);
        Bt.p[rr] = Bt.z[rr] + bt * Bt.p[rr];
        Bt.Ap[rr] = Bt.Az[rr] + bt * Bt.Ap[rr];
      }
      Bt.r_sq[w] = wdot(Bt, w, Bt.r, Bt.r);
    }
    ++iter;
    for (int w : ws) {
      T target = atol_sq;
      T rel = Bt.b_sq[w] * tol_sq;
      if (rel > target) target = rel;
      bool unconv = Bt.r_sq[w] > target;
      if (global) { if (unconv) any = true; }
      else run[w] = (run[w] && unconv && iter < Bt.cfg.lin_max_iters) ? 1 : 0;
    }
    if (global) {
      if (!any || iter >= Bt.cfg.lin_max_iters) break;
    } else {
      bool anyrun = false;
      for (int w : ws) anyrun |= run[w] != 0;
      if (!anyrun) break;
    }
  }
}

// ----------------------------------------------------------------------------------------
// Newton step pieces
// ----------------------------------------------------------------------------------------
// integrate_body_pose (mechanics/pose_integration.py:6-38, kinematic_mapping.py:70-141)
template <class T> Tf<T> integrate_pose(const Sp<T>& vel, const Tf<T>& pp, V3<T> com, T dt) {
  V3<T> pcom = add(pp.p, qrot(pp.q, com));
  V3<T> v = top(vel), wv = bot(vel);
  V3<T> out_p = add(pcom, mul(v, dt));
  T s = dt * T(0.5);
  T t2 = pp.q.x, t3 = pp.q.y, t4 = pp.q.z, t1 = pp.q.w;
  Qt<T> rq;
  rq.x = pp.q.x + (t1 * wv.x + t4 * wv.y - t3 * wv.z) * s;
  rq.y = pp.q.y + (-t4 * wv.x + t1 * wv.y + t2 * wv.z) * s;
  rq.z = pp.q.z + (t3 * wv.x - t2 * wv.y + t1 * wv.z) * s;
  rq.w = pp.q.w + (-t2 * wv.x - t3 * wv.y - t4 * wv.z) * s;
  Qt<T> rn = qnormalize(rq);
  return Tf<T>{sub(out_p, qrot(rn, com)), rn};
}

template <class T> void newton_update(Batch<T>& Bt, int w) {
  // compute_dbody_qd_from_dbody_lambda (linear_system.py:344-377)
  for (int b = 0; b < Bt.B; ++b) Bt.tmp_body[Bt.bi(w, b)] = spzero<T>();
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    T dl = Bt.dlam[rr];
    if (Bt.b0[rr] >= 0) { Sp<T>& t = Bt.tmp_body[Bt.bi(w, Bt.b0[rr])]; for (int k = 0; k < 6; ++k) t.v[k] += dl * Bt.J0[rr].v[k]; }
    if (Bt.b1[rr] >= 0) { Sp<T>& t = Bt.tmp_body[Bt.bi(w, Bt.b1[rr])]; for (int k = 0; k < 6; ++k) t.v[k] += dl * Bt.J1[rr].v[k]; }
  }
  for (int b = 0; b < Bt.B; ++b) {
    size_t k = Bt.bi(w, b);
    Sp<T> a;
    for (int i = 0; i < 6; ++i) a.v[i] = Bt.tmp_body[k].v[i] * Bt.dt - Bt.h_d[k].v[i];
    Bt.du[k] = minv(Bt, w, b, a);
  }
  // apply_standard_newton_step (core/newton_step.py:22-54)
  for (int b = 0; b < Bt.B; ++b) {
    size_t k = Bt.bi(w, b);
    for (int i = 0; i < 6; ++i) Bt.u[k].v[i] = Bt.u[k].v[i] + Bt.du[k].v[i];
  }
  for (int i = 0; i < Bt.Nc; ++i) { size_t rr = Bt.ri(w, i); Bt.lam[rr] = Bt.lam[rr] + Bt.dlam[rr]; }
  for (int b = 0; b < Bt.B; ++b) {
    size_t k = Bt.bi(w, b);
    Bt.q[k] = integrate_pose(Bt.u[k], Bt.q_prev[k], Bt.com(w, b), Bt.dt);
  }
}

// _compute_linesearch_step_size_array (core/engine_data.py:15-39), in double, as float
std::vector<float> linesearch_step_list(const ofb_config& c) {
  const int nc = c.linesearch_conservative_count, no = c.linesearch_optimistic_count;
  std::vector<double> v;
  const double a = std::log10((double)c.linesearch_min_step), b = std::log10((double)c.linesearch_conservative_upper);
  for (int i = 0; i < nc; ++i) v.push_back(std::pow(10.0, nc == 1 ? a : (i == nc - 1 ? b : a + i * ((b - a) / (nc - 1)))));
  const double h = c.linesearch_optimistic_window / 2.0, lo = 1.0 - h, hi = 1.0 + h;
  for (int i = 0; i < no; ++i) v.push_back(no == 1 ? lo : (i == no - 1 ? hi : lo + i * ((hi - lo) / (no - 1))));
  std::sort(v.begin(), v.end());
  size_t best = 0;
  for (size_t i = 1; i < v.size(); ++i)
    if (std::fabs(v[i] - 1.0) < std::fabs(v[best] - 1.0)) best = i;
  v[best] = 1.0;
  return std::vector<float>(v.begin(), v.end());
}

// perform_linesearch (core/linesearch.py:330-339): candidates u0 + a du, integrate from
// the step-start pose, lambda0 + a dlambda, lambda_prev unchanged; first argmin of the
// merit; the chosen iterate's residual stays in h_d / h_c.
template <class T> void newton_update_ls(Batch<T>& Bt, int w, const std::vector<float>& steps) {
  for (int b = 0; b < Bt.B; ++b) Bt.tmp_body[Bt.bi(w, b)] = spzero<T>();
  for (int i = 0; i < Bt.Nc; ++i) {
    size_t rr = Bt.ri(w, i);
    T dl = Bt.dlam[rr];
    if (Bt.b0[rr] >= 0) { Sp<T>& t = Bt.tmp_body[Bt.bi(w, Bt.b0[rr])]; for (int k = 0; k < 6; ++k) t.v[k] += dl * Bt.J0[rr].v[k]; }
    if (Bt.b1[rr] >= 0) { Sp<T>& t = Bt.tmp_body[Bt.bi(w, Bt.b1[rr])]; for (int k = 0; k < 6; ++k) t.v[k] += dl * Bt.J1[rr].v[k]; }
  }
  std::vector<Sp<T>> u0(Bt.B);
  std::vector<T> l0(Bt.Nc);
  for (int b = 0; b < Bt.B; ++b) {
    size_t k = Bt.bi(w, b);
    Sp<T> a;
    for (int i = 0; i < 6; ++i) a.v[i] = Bt.tmp_body[k].v[i] * Bt.dt - Bt.h_d[k].v[i];
    Bt.du[k] = minv(Bt, w, b, a);
    u0[b] = Bt.u[k];
  }
  for (int i = 0; i < Bt.Nc; ++i) l0[i] = Bt.lam[Bt.ri(w, i)];
  auto set_alpha = [&](T a) {
    for (int b = 0; b < Bt.B; ++b) {
      size_t k = Bt.bi(w, b);
      for (int i = 0; i < 6; ++i) Bt.u[k].v[i] = u0[b].v[i] + a * Bt.du[k].v[i];
      Bt.q[k] = integrate_pose(Bt.u[k], Bt.q_prev[k], Bt.com(w, b), Bt.dt);
    }
    for (int i = 0; i < Bt.Nc; ++i) Bt.lam[Bt.ri(w, i)] = l0[i] + a * Bt.dlam[Bt.ri(w, i)];
  };
  size_t best = 0;
  T best_m = 0;
  for (size_t si = 0; si < steps.size(); ++si) {
    set_alpha(T(steps[si]));
    residual(Bt, w);
    T m = merit_of(Bt, w);
    if (si == 0 || m < best_m) { best_m = m; best = si; }
  }
  set_alpha(T(steps[best]));
}

template <class T> void save_candidate(Batch<T>& Bt, int w, int idx) {  // history_group.py:12-79
  if (idx >= Bt.cfg.nr_max_iters) return;
  for (int b = 0; b < Bt.B; ++b) { size_t k = Bt.bi(w, b); Bt.cand_q[idx][k] = Bt.q[k]; Bt.cand_u[idx][k] = Bt.u[k]; }
  for (int i = 0; i < Bt.Nc; ++i) Bt.cand_lam[idx][Bt.ri(w, i)] = Bt.lam[Bt.ri(w, i)];
  Bt.cand_merit[idx][w] = Bt.merit[w];
}

// perform_backtracking (core/backtracking.py:12-95, reductions.py:28-62)
template <class T> void backtrack(Batch<T>& Bt, int w, int end_idx) {
  int min_iter = Bt.cfg.nr_backtrack_min_iter;
  int bi_;
  if (end_idx <= 0) bi_ = 0;
  else if (end_idx <= min_iter) bi_ = end_idx - 1;
  else {
    bi_ = min_iter;
    T mv_ = Bt.cand_merit[min_iter][w];
    for (int i = min_iter + 1; i < end_idx; ++i) {
      T v = Bt.cand_merit[i][w];
      if (v < mv_) { bi_ = i; mv_ = v; }
    }
  }
  Bt.best[w] = bi_;
  for (int b = 0; b < Bt.B; ++b) { size_t k = Bt.bi(w, b); Bt.q[k] = Bt.cand_q[bi_][k]; Bt.u[k] = Bt.cand_u[bi_][k]; }
  for (int i = 0; i < Bt.Nc; ++i) Bt.lam[Bt.ri(w, i)] = Bt.cand_lam[bi_][Bt.ri(w, i)];
  Bt.merit[w] = Bt.cand_merit[bi_][w];
}

template <class T> int run_step(const ofb_model* m, const ofb_config* cfg, const double* q_in, const double* qd_in,
                                const float* f_in, const float* tpos, const float* tvel, const ofb_contacts* contacts,
                                float dt, const double* lam_seed, double* q_out, double* qd_out, double* lam_out,
                                int32_t* nr_iters, int32_t* pcr_iters, double* merit_out, int32_t* best_out,
                                int32_t* flags_out, int32_t* ccount_out, double* merit_hist, double* res_out,
                                double* lam_prev_out) {
  Batch<T> Bt(*m, *cfg);
  Bt.dt = T(dt);
  const int W = Bt.W, B = Bt.B;
  pack_contacts(Bt, *contacts);
  reduce_contacts(Bt, q_in);
  for (int w = 0; w < W; ++w) {  // load_data (base_engine.py:289-335), engine.py:48-59
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      const double* x = q_in + k * 7;
      Bt.q_prev[k] = Tf<T>{v3<T>(T(x[0]), T(x[1]), T(x[2])), Qt<T>{T(x[3]), T(x[4]), T(x[5]), T(x[6])}};
      Bt.q[k] = Bt.q_prev[k];
      for (int i = 0; i < 6; ++i) {
        Bt.u_prev[k].v[i] = T(qd_in[k * 6 + i]);
        Bt.u[k].v[i] = Bt.u_prev[k].v[i];
        Bt.f_ext[k].v[i] = T(f_in[k * 6 + i]);
      }
    }
    for (int d = 0; d < Bt.D; ++d) { Bt.tpos[size_t(w) * Bt.D + d] = T(tpos[w * Bt.D + d]); Bt.tvel[size_t(w) * Bt.D + d] = T(tvel[w * Bt.D + d]); }
    for (int i = 0; i < Bt.Nc; ++i) {
      size_t rr = Bt.ri(w, i);
      Bt.lam[rr] = (cfg->seed_iterate && lam_seed) ? T(lam_seed[rr]) : T(0);
    }
  }
  const bool global = cfg->termination == OFB_TERM_GLOBAL;
  const T atol_sq = T(cfg->nr_atol) * T(cfg->nr_atol);
  const std::vector<float> ls_steps = cfg->linesearch ? linesearch_step_list(*cfg) : std::vector<float>();
  std::vector<char> running(W, 1);
  // one Newton iteration (_solve, base_engine.py:416-526) of the worlds in `act`;
  // returns whether any of them is above the tolerance
  auto iterate = [&](const std::vector<int>& act) {
    for (int w : act) {
      for (int i = 0; i < Bt.Nc; ++i) Bt.lam_prev[Bt.ri(w, i)] = Bt.lam[Bt.ri(w, i)];
      linear_system(Bt, w);
      if (cfg->preconditioner == OFB_PRECOND_PER_BODY_PAIR) pair_update(Bt, w);
      else jacobi_update(Bt, w);
      for (int i = 0; i < Bt.Nc; ++i) Bt.dlam[Bt.ri(w, i)] = 0;
    }
    pcr_solve(Bt, act, global);
    bool any = false;
    for (int w : act) {
      if (cfg->linesearch) newton_update_ls(Bt, w, ls_steps);
      else newton_update(Bt, w);
      residual(Bt, w);
      Bt.merit[w] = merit_of(Bt, w);
      if (merit_hist && Bt.nr_iters[w] < cfg->nr_max_iters) merit_hist[size_t(w) * cfg->nr_max_iters + Bt.nr_iters[w]] = Bt.merit[w];
      if (res_out) {  // residual [h_d | h_c] of the latest iterate
        size_t base = size_t(w) * (Bt.Nu + Bt.Nc);
        for (int b = 0; b < Bt.B; ++b) for (int i = 0; i < 6; ++i) res_out[base + b * 6 + i] = Bt.h_d[Bt.bi(w, b)].v[i];
        for (int i = 0; i < Bt.Nc; ++i) res_out[base + Bt.Nu + i] = Bt.h_c[Bt.ri(w, i)];
      }
      save_candidate(Bt, w, Bt.nr_iters[w]);
      Bt.nr_iters[w] += 1;
      bool unconv = Bt.merit[w] > atol_sq;
      any |= unconv;
      if (!global) {
        int it = Bt.nr_iters[w];
        bool keep = unconv;
        if (it < cfg->nr_backtrack_min_iter) keep = true;
        if (it >= cfg->nr_max_iters) keep = false;
        running[w] = keep ? 1 : 0;
      }
    }
    return any;
  };
  if (global) {  // every world iterates until none is above the tolerance
    std::vector<int> all(W);
    for (int w = 0; w < W; ++w) all[w] = w;
    int iter = 0;
    while (true) {
      bool keep = iterate(all);
      ++iter;
      if (iter < cfg->nr_backtrack_min_iter) keep = true;
      if (iter >= cfg->nr_max_iters) keep = false;
      if (!keep) break;
    }
  } else {
    // per-world termination: each world's loop touches only its own rows and bodies, so
    // worlds run independently (OpenMP over worlds; the arithmetic of a world does not
    // depend on the thread count)
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < W; ++w) {
      const std::vector<int> one(1, w);
      while (running[w]) iterate(one);
    }
  }
  for (int w = 0; w < W; ++w) {
    if (Bt.nr_iters[w] >= cfg->nr_max_iters && Bt.merit[w] > atol_sq) Bt.flags[w] |= 1;
    backtrack(Bt, w, Bt.nr_iters[w]);
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      double* x = q_out + k * 7;
      x[0] = Bt.q[k].p.x; x[1] = Bt.q[k].p.y; x[2] = Bt.q[k].p.z;
      x[3] = Bt.q[k].q.x; x[4] = Bt.q[k].q.y; x[5] = Bt.q[k].q.z; x[6] = Bt.q[k].q.w;
      for (int i = 0; i < 6; ++i) qd_out[k * 6 + i] = Bt.u[k].v[i];
    }
    if (lam_out) for (int i = 0; i < Bt.Nc; ++i) lam_out[Bt.ri(w, i)] = Bt.lam[Bt.ri(w, i)];
    // lambda_prev_iter as the step leaves it: the snapshot taken at the start of the last
    // Newton iteration (base_engine.py:441), which backtracking does not restore
    if (lam_prev_out) for (int i = 0; i < Bt.Nc; ++i) lam_prev_out[Bt.ri(w, i)] = Bt.lam_prev[Bt.ri(w, i)];
    if (nr_iters) nr_iters[w] = Bt.nr_iters[w];
    if (pcr_iters) pcr_iters[w] = Bt.pcr_iters[w];
    if (merit_out) merit_out[w] = Bt.merit[w];
    if (best_out) best_out[w] = Bt.best[w];
    if (flags_out) flags_out[w] = Bt.flags[w];
    if (ccount_out) ccount_out[w] = Bt.ccount[w];
  }
  return 0;
}

// ----------------------------------------------------------------------------------------
// Adjoint of one step (base_engine.py:528-743, adjoint/adjoint_utils.py:11-169,
// residual.py:176-406) at the state the forward step left: q+ = q_in, u+ = qd_in,
// lambda, lambda_prev_iter; q-, u-, f of the step start.
// ----------------------------------------------------------------------------------------
// driving force -(g_u + dt G(q)^T g_q) - [0; dt (R(q)(-com)) x g_p] (Gt_matvec,
// kinematic_mapping.py:208-279; compute_body_adjoint_init_kernel, adjoint_utils.py:12-79)
template <class T> Sp<T> adjoint_drive(const Tf<T>& q, V3<T> com, const T* gq, const T* gu, T dt) {
  T t2 = q.q.x, t3 = q.q.y, t4 = q.q.z, t1 = q.q.w;
  T d2 = gq[3], d3 = gq[4], d4 = gq[5], d1 = gq[6];
  T s = dt * T(0.5);
  Sp<T> o;
  for (int i = 0; i < 3; ++i) o.v[i] = gu[i] + gq[i] * dt;
  o.v[3] = gu[3] + (-t2 * d1 + t1 * d2 - t4 * d3 + t3 * d4) * s;
  o.v[4] = gu[4] + (-t3 * d1 + t4 * d2 + t1 * d3 - t2 * d4) * s;
  o.v[5] = gu[5] + (-t4 * d1 - t3 * d2 + t2 * d3 + t1 * d4) * s;
  for (int i = 0; i < 6; ++i) o.v[i] = -o.v[i];
  V3<T> lever = qrot(q.q, neg(com));
  V3<T> tq = cross(lever, v3<T>(gq[0], gq[1], gq[2]));
  o.v[3] -= tq.x * dt; o.v[4] -= tq.y * dt; o.v[5] -= tq.z * dt;
  return o;
}

// explicit carry dq+/dq- of a pose gradient (body_pose_prev_grad_kernel, residual.py:260-339)
template <class T> void adjoint_carry(const T* g, const Sp<T>& vel, T dt, T* out) {
  T w1 = vel.v[3], w2 = vel.v[4], w3 = vel.v[5], s = dt * T(0.5);
  T gx = g[3], gy = g[4], gz = g[5], gw = g[6];
  out[0] = g[0]; out[1] = g[1]; out[2] = g[2];
  out[3] = gx + s * (w3 * gy - w2 * gz - w1 * gw);
  out[4] = gy + s * (-w3 * gx + w1 * gz - w2 * gw);
  out[5] = gz + s * (w2 * gx - w1 * gy - w3 * gw);
  out[6] = gw + s * (w1 * gx + w2 * gy + w3 * gz);
}

// s = sum_rows w_lambda h_c + sum_bodies w_u . h_d at the current state of world w
template <class T> T adjoint_weighted_residual(Batch<T>& Bt, int w, const std::vector<T>& wl,
                                               const std::vector<Sp<T>>& wu) {
  residual(Bt, w);
  T s = 0;
  for (int i = 0; i < Bt.Nc; ++i) s += wl[Bt.ri(w, i)] * Bt.h_c[Bt.ri(w, i)];
  for (int b = 0; b < Bt.B; ++b) s += sdot(wu[Bt.bi(w, b)], Bt.h_d[Bt.bi(w, b)]);
  return s;
}

template <class T> T& pose_coord(Tf<T>& X, int k) {
  switch (k) {
    case 0: return X.p.x; case 1: return X.p.y; case 2: return X.p.z;
    case 3: return X.q.x; case 4: return X.q.y; case 5: return X.q.z;
    default: return X.q.w;
  }
}

// adjoint.friction_linearization = "frozen" (freeze_contact_mode_kernel and its soft
// variant, adjoint/adjoint_friction.py:20-158, launched in base_engine.py:547-590): each
// friction row of a contact with normal impulse >= 1e-6 gets the compliance of its contact
// mode -- sticking (|lambda_t| < 0.9 mu lambda_n): compliance.joint; sliding: 100 -- or,
// with soft blending, a sigmoid of (ratio - 0.9) / temperature between the two; rows of
// contacts without normal impulse leave the system.
template <class T> void freeze_friction(Batch<T>& Bt, int w, const ofb_config& cfg) {
  const int K = Bt.Cmax, on = Bt.Nj + Bt.Nctrl, of = on + K;
  const T stick = T(cfg.compliance_joint), slide = T(100);
  for (int c = 0; c < std::min(Bt.ccount[w], K); ++c) {
    size_t kc = size_t(w) * K + c;
    size_t r0 = Bt.ri(w, of + 2 * c), r1 = r0 + 1;
    T lt1 = Bt.lam[r0], lt2 = Bt.lam[r1], ln = Bt.lam[Bt.ri(w, on + c)];
    T mu = T(0.5) * (Bt.smu(w, Bt.cs0[kc]) + Bt.smu(w, Bt.cs1[kc]));
    if (ln < T(1e-6)) {
      Bt.active[r0] = T(0);
      Bt.active[r1] = T(0);
      continue;
    }
    T ratio = std::sqrt(lt1 * lt1 + lt2 * lt2 + T(1e-10)) / (mu * ln + T(1e-10));
    T ca;
    if (cfg.adjoint_soft_blending) {
      T e = -(ratio - T(0.9)) / std::max(T(cfg.adjoint_soft_temperature), T(1e-8));
      e = std::min(std::max(e, T(-20)), T(20));
      T sg = T(1) / (T(1) + std::exp(e));
      ca = (T(1) - sg) * stick + sg * slide;
    } else {
      ca = ratio > T(0.9) ? slide : stick;
    }
    Bt.C[r0] = ca;
    Bt.C[r1] = ca;
  }
}

template <class T> int run_adjoint(const ofb_model* m, const ofb_config* cfg, const double* q_prev_in,
                                   const double* qd_prev_in, const float* f_in, const double* q_in,
                                   const double* qd_in, const float* tpos, const float* tvel,
                                   const ofb_contacts* contacts, float dt, const double* lam_in,
                                   const double* lam_prev_in, const double* g_q, const double* g_u, double fd_eps,
                                   double* g_q_prev, double* g_u_prev, double* g_f, double* g_tpos, double* g_tvel,
                                   double* g_ke, double* g_kd, double* g_mu, double* w_lam_out, double* w_u_out,
                                   int32_t* pcr_iters, int32_t* flags_out) {
  Batch<T> Bt(*m, *cfg);
  Bt.dt = T(dt);
  const int W = Bt.W, B = Bt.B, D = Bt.D, Nc = Bt.Nc, SC = Bt.SC;
  pack_contacts(Bt, *contacts);
  reduce_contacts(Bt, q_prev_in);
  auto ld_tf = [](const double* x) {
    return Tf<T>{v3<T>(T(x[0]), T(x[1]), T(x[2])), Qt<T>{T(x[3]), T(x[4]), T(x[5]), T(x[6])}};
  };
  std::vector<T> wl(size_t(W) * Nc, T(0));
  std::vector<Sp<T>> wu(size_t(W) * B, spzero<T>());
  std::vector<int> all(W);
  for (int w = 0; w < W; ++w) {
    all[w] = w;
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      Bt.q_prev[k] = ld_tf(q_prev_in + k * 7);
      Bt.q[k] = ld_tf(q_in + k * 7);
      for (int i = 0; i < 6; ++i) {
        Bt.u_prev[k].v[i] = T(qd_prev_in[k * 6 + i]);
        Bt.u[k].v[i] = T(qd_in[k * 6 + i]);
        Bt.f_ext[k].v[i] = T(f_in[k * 6 + i]);
      }
    }
    for (int d = 0; d < D; ++d) { Bt.tpos[size_t(w) * D + d] = T(tpos[w * D + d]); Bt.tvel[size_t(w) * D + d] = T(tvel[w * D + d]); }
    for (int i = 0; i < Nc; ++i) { Bt.lam[Bt.ri(w, i)] = T(lam_in[Bt.ri(w, i)]); Bt.lam_prev[Bt.ri(w, i)] = T(lam_prev_in[Bt.ri(w, i)]); }
    // compute_linear_system at the converged state (base_engine.py:533-535)
    linear_system(Bt, w);
    if (cfg->adjoint_friction_frozen) freeze_friction(Bt, w, *cfg);
    // adjoint_regularize_compliance_kernel: C += gamma on active rows
    if (cfg->adjoint_regularization > 0)
      for (int i = 0; i < Nc; ++i) if (Bt.active[Bt.ri(w, i)] > T(0)) Bt.C[Bt.ri(w, i)] += T(cfg->adjoint_regularization);
    // w_u0 = M^-1 f_drive (compute_body_adjoint_init_kernel)
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      T gq[7], gu[6];
      for (int i = 0; i < 7; ++i) gq[i] = T(g_q[k * 7 + i]);
      for (int i = 0; i < 6; ++i) gu[i] = T(g_u[k * 6 + i]);
      wu[k] = minv(Bt, w, b, adjoint_drive(Bt.q[k], Bt.com(w, b), gq, gu, Bt.dt));
    }
    // b = J w_u0 (compute_adjoint_rhs_kernel)
    for (int i = 0; i < Nc; ++i) {
      size_t rr = Bt.ri(w, i);
      T v = 0;
      if (Bt.active[rr] > T(0)) {
        if (Bt.b0[rr] >= 0) v += sdot(Bt.J0[rr], wu[Bt.bi(w, Bt.b0[rr])]);
        if (Bt.b1[rr] >= 0) v += sdot(Bt.J1[rr], wu[Bt.bi(w, Bt.b1[rr])]);
      }
      Bt.rhs[rr] = v;
      Bt.dlam[rr] = 0;
    }
    if (cfg->preconditioner == OFB_PRECOND_PER_BODY_PAIR) pair_update(Bt, w);
    else jacobi_update(Bt, w);
  }
  // A w_lambda = b, same operator and PCR (base_engine.py:633-644)
  pcr_solve(Bt, all, cfg->termination == OFB_TERM_GLOBAL);
  for (int w = 0; w < W; ++w) {
    // w_u -= M^-1 J^T w_lambda (subtract_constraint_feedback_kernel)
    for (int i = 0; i < Nc; ++i) {
      size_t rr = Bt.ri(w, i);
      wl[rr] = Bt.dlam[rr];
      if (Bt.active[rr] == T(0)) continue;
      for (int s_ = 0; s_ < 2; ++s_) {
        int b = s_ == 0 ? Bt.b0[rr] : Bt.b1[rr];
        if (b < 0) continue;
        Sp<T> a = minv(Bt, w, b, sscale(s_ == 0 ? Bt.J0[rr] : Bt.J1[rr], wl[rr]));
        for (int k = 0; k < 6; ++k) wu[Bt.bi(w, b)].v[k] -= a.v[k];
      }
    }
    int flags = 0;
    // closed-form parts of w^T dR/dxi
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      M3<T> Iw = world_inertia(Bt.q[k].q, Bt.inertia(w, b));
      Sp<T> gp = spatial_momentum(-Bt.mass(w, b), Iw, wu[k]);  // dynamics: d h_d / d u- = -M
      for (int i = 3; i < 6; ++i) gp.v[i] = -gp.v[i];
      for (int i = 0; i < 6; ++i) {
        if (g_u_prev) g_u_prev[k * 6 + i] = double(gp.v[i]);
        if (g_f) g_f[k * 6 + i] = double(-Bt.dt * wu[k].v[i]);  // d h_d / d f = -dt
      }
    }
    for (int d = 0; d < D; ++d) {
      size_t kd = size_t(w) * D + d;
      if (g_tpos) g_tpos[kd] = 0;
      if (g_tvel) g_tvel[kd] = 0;
      if (g_ke) g_ke[kd] = 0;
      if (g_kd) g_kd[kd] = 0;
    }
    for (int j = 0; j < Bt.J; ++j) {  // control_target_grad_kernel (residual.py:201-257)
      if (Bt.ctrl_off[j] < 0) continue;
      size_t rr = Bt.ri(w, Bt.Nj + Bt.ctrl_off[j]);
      if (Bt.active[rr] == T(0)) continue;
      int qs = m->joint_qd_start[w * Bt.J + j];
      int mode = m->joint_dof_mode[w * D + qs];
      size_t kd = size_t(w) * D + qs;
      T wlam = wl[rr], lam = Bt.lam[rr];
      T ke = T(m->joint_target_ke[w * D + qs]), kdg = T(m->joint_target_kd[w * D + qs]);
      if (mode == 1) {
        if (g_tpos) g_tpos[kd] += double(wlam * (T(-1) / Bt.dt));
        T den = Bt.dt * Bt.dt * ke + Bt.dt * kdg;
        if (den > T(1e-6)) {
          T inv2 = T(1) / (den * den);
          if (g_ke) g_ke[kd] += double(wlam * (-Bt.dt * Bt.dt * inv2) * lam * Bt.dt);
          if (g_kd) g_kd[kd] += double(wlam * (-Bt.dt * inv2) * lam * Bt.dt);
        }
      } else if (mode == 2) {
        if (g_tvel) g_tvel[kd] += double(-wlam);
        T den = Bt.dt * ke;
        if (den > T(1e-6)) {
          T inv2 = T(1) / (den * den);
          if (g_ke) g_ke[kd] += double(wlam * (-Bt.dt * inv2) * lam * Bt.dt);
        }
      }
    }
    // frozen friction: the friction rows are not seeded in w^T dR/dxi (base_engine.py:
    // 704-711), so neither the mu gradient nor the pose terms see them
    std::vector<T> wl_seed(wl);
    if (cfg->adjoint_friction_frozen)
      for (int i = Bt.Nj + Bt.Nctrl + Bt.Cmax; i < Nc; ++i) wl_seed[Bt.ri(w, i)] = T(0);
    // friction coefficient: h_f = v_t + w(mu) lambda on both tangent rows
    residual(Bt, w);  // refreshes dw_dmu at the converged state
    if (g_mu) for (int s_ = 0; s_ < SC; ++s_) g_mu[size_t(w) * SC + s_] = 0;
    for (int c = 0; c < std::min(Bt.ccount[w], Bt.Cmax); ++c) {
      size_t kc = size_t(w) * Bt.Cmax + c;
      size_t r0 = Bt.ri(w, Bt.Nj + Bt.Nctrl + Bt.Cmax + 2 * c), r1 = r0 + 1;
      T dw = Bt.dw_dmu[kc];
      T contrib = wl_seed[r0] * Bt.lam[r0] + wl_seed[r1] * Bt.lam[r1];
      if (std::isnan(dw)) { if (contrib != T(0)) flags |= 16; continue; }
      if (dw == T(0) || !g_mu) continue;
      int s0 = Bt.cs0[kc], s1 = Bt.cs1[kc];
      g_mu[size_t(w) * SC + s0] += double(T(0.5) * dw * contrib);
      g_mu[size_t(w) * SC + s1] += double(T(0.5) * dw * contrib);
    }
    // pose pull-back w^T dR/dq+ and w^T dR/dq- by central differences of the residual
    std::vector<T> gplus(size_t(B) * 7, T(0)), gminus(size_t(B) * 7, T(0));
    if (cfg->adjoint_pose_vjp) {
      T h = T(fd_eps);
      for (int b = 0; b < B; ++b) {
        for (int side = 0; side < 2; ++side) {
          Tf<T>& X = side == 0 ? Bt.q[Bt.bi(w, b)] : Bt.q_prev[Bt.bi(w, b)];
          for (int kk = 0; kk < 7; ++kk) {
            T& v = pose_coord(X, kk);
            T orig = v;
            v = orig + h;
            T sp_ = adjoint_weighted_residual(Bt, w, wl_seed, wu);
            v = orig - h;
            T sm_ = adjoint_weighted_residual(Bt, w, wl_seed, wu);
            v = orig;
            (side == 0 ? gplus : gminus)[b * 7 + kk] = (sp_ - sm_) / (T(2) * h);
          }
        }
      }
    }
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      T G[7], out[7];
      for (int i = 0; i < 7; ++i) G[i] = T(g_q[k * 7 + i]) + gplus[b * 7 + i];
      adjoint_carry(G, Bt.u[k], Bt.dt, out);
      if (g_q_prev) for (int i = 0; i < 7; ++i) g_q_prev[k * 7 + i] = double(out[i] + gminus[b * 7 + i]);
      if (w_u_out) for (int i = 0; i < 6; ++i) w_u_out[k * 6 + i] = double(wu[k].v[i]);
    }
    if (w_lam_out) for (int i = 0; i < Nc; ++i) w_lam_out[Bt.ri(w, i)] = double(wl[Bt.ri(w, i)]);
    if (pcr_iters) pcr_iters[w] = Bt.pcr_iters[w];
    if (flags_out) flags_out[w] = flags;
  }
  return 0;
}

// ----------------------------------------------------------------------------------------
// Cross-step contact warm start, position_match (collision/warm_start.py:84-519).  The
// snapshot of the previous step lives in caller arrays (per world: count; per slot: b0,
// b1, world midpoint, normal, lambda_n, lambda_t).  Both hooks use the start-of-step
// pose and velocity (body_pose_prev / body_vel_prev, warm_start.py:600-609).
// ----------------------------------------------------------------------------------------
struct WSPrev {
  int32_t* count;  // [W]
  int32_t* b0;     // [W*K]
  int32_t* b1;
  double* p;       // [W*K*3]
  double* n;       // [W*K*3]
  double* ln;      // [W*K]
  double* lt;      // [W*K*2]
};

constexpr double kWsAlpha = 0.1, kWsMinThreshold = 5.0e-3, kWsNDotUpMin = 0.3;  // warm_start.py:520-538

// midpoint of contact c in world frame (_world_midpoint, warm_start.py:51-82)
inline V3<double> ws_midpoint(const Batch<double>& Bt, int w, int c, const std::vector<Tf<double>>& pose, int b0,
                              int b1) {
  size_t k = size_t(w) * Bt.Cmax + c;
  V3<double> n = Bt.cvec(Bt.cn, w, c);
  Tf<double> pa = b0 >= 0 ? pose[Bt.bi(w, b0)] : tident<double>();
  Tf<double> pb = b1 >= 0 ? pose[Bt.bi(w, b1)] : tident<double>();
  V3<double> a = sub(tpoint(pa, Bt.cvec(Bt.cp0, w, c)), mul(n, Bt.ct0[k]));
  V3<double> b = add(tpoint(pb, Bt.cvec(Bt.cp1, w, c)), mul(n, Bt.ct1[k]));
  return mul(add(a, b), 0.5);
}

// apply(): _match_kernel (:84-297), or with pair_aggregate _pair_count_current_kernel and
// _pair_apply_kernel (:943-1010), then _cold_start_kernel (:300-462) -> seed [W*Nc]
void ws_apply(Batch<double>& Bt, const WSPrev& P, int method, int cold_gravity, int cold_impact,
              double v_t_threshold, double* seed) {
  const int W = Bt.W, K = Bt.Cmax, B = Bt.B;
  const double dt = Bt.dt;
  V3<double> g = v3<double>(Bt.M.gravity[0], Bt.M.gravity[1], Bt.M.gravity[2]);
  double gmag = std::sqrt(dot(g, g));
  V3<double> up = gmag > 1e-6 ? mul(g, -1.0 / gmag) : v3<double>(0, 0, 1);
  int on = Bt.Nj + Bt.Nctrl, of = on + K;
  for (int w = 0; w < W; ++w) {
    for (int i = 0; i < Bt.Nc; ++i) seed[Bt.ri(w, i)] = 0;
    std::vector<double> support(B, 0.0);
    std::vector<int> unmatched(B, 0), matched(K, 0);
    const int n_c = std::min(Bt.ccount[w], K);
    const int n_prev = std::min(P.count[w], K);
    for (int c = 0; c < n_c; ++c) {
      size_t k = size_t(w) * K + c;
      int s0 = Bt.cs0[k], s1 = Bt.cs1[k];
      if (s0 == s1) continue;
      int b0 = s0 >= 0 ? Bt.shape_body(w, s0) : -1, b1 = s1 >= 0 ? Bt.shape_body(w, s1) : -1;
      if (method == OFB_WS_PAIR_AGGREGATE) {
        // the pair's previous total normal impulse (_pair_snapshot_kernel, :910-940, summed
        // here in contact order) spread evenly over the pair's current contacts; lambda_t
        // stays 0 and the gravity share of the cold start sees no unmatched contacts
        double total = 0;
        int n_pp = 0, n_cur = 0;
        for (int j = 0; j < n_prev; ++j) {
          size_t kj = size_t(w) * K + j;
          if (P.b0[kj] == b0 && P.b1[kj] == b1) { total += P.ln[kj]; ++n_pp; }
        }
        for (int j = 0; j < n_c; ++j) {
          size_t kj = size_t(w) * K + j;
          if (Bt.cs0[kj] == Bt.cs1[kj]) continue;
          int c0 = Bt.cs0[kj] >= 0 ? Bt.shape_body(w, Bt.cs0[kj]) : -1;
          int c1 = Bt.cs1[kj] >= 0 ? Bt.shape_body(w, Bt.cs1[kj]) : -1;
          if (c0 == b0 && c1 == b1) ++n_cur;
        }
        if (n_pp > 0 && n_cur > 0) {
          seed[Bt.ri(w, on + c)] = total / double(n_cur);
          matched[c] = 1;
        }
        continue;
      }
      V3<double> pc = ws_midpoint(Bt, w, c, Bt.q_prev, b0, b1);
      V3<double> nc = Bt.cvec(Bt.cn, w, c);
      V3<double> vt = v3<double>(0, 0, 0);
      if (b0 >= 0) vt = add(vt, top(Bt.u_prev[Bt.bi(w, b0)]));
      if (b1 >= 0) vt = sub(vt, top(Bt.u_prev[Bt.bi(w, b1)]));
      vt = sub(vt, mul(nc, dot(vt, nc)));
      double thr = std::max(kWsMinThreshold, kWsAlpha * std::sqrt(dot(vt, vt)) * dt);
      double best_d = thr * thr;
      int best = -1;
      for (int j = 0; j < n_prev; ++j) {
        size_t kj = size_t(w) * K + j;
        if (P.b0[kj] != b0 || P.b1[kj] != b1) continue;
        V3<double> pp = add(v3<double>(P.p[kj * 3], P.p[kj * 3 + 1], P.p[kj * 3 + 2]), mul(vt, dt));
        V3<double> d = sub(pc, pp);
        double dsq = dot(d, d);
        if (dsq < best_d) { best_d = dsq; best = j; }
      }
      double ndu = dot(nc, up);
      bool sup0 = -ndu > kWsNDotUpMin, sup1 = ndu > kWsNDotUpMin;
      if (best < 0) {
        if (b0 >= 0 && sup0) unmatched[b0] += 1;
        if (b1 >= 0 && sup1) unmatched[b1] += 1;
        continue;
      }
      size_t kb = size_t(w) * K + best;
      V3<double> pn = v3<double>(P.n[kb * 3], P.n[kb * 3 + 1], P.n[kb * 3 + 2]);
      V3<double> pt1, pt2, ct1, ct2;
      orthogonal_basis(pn, pt1, pt2);
      orthogonal_basis(nc, ct1, ct2);
      V3<double> ltw = add(mul(pt1, P.lt[kb * 2]), mul(pt2, P.lt[kb * 2 + 1]));
      double lnm = P.ln[kb];
      seed[Bt.ri(w, on + c)] = lnm;
      seed[Bt.ri(w, of + 2 * c)] = dot(ltw, ct1);
      seed[Bt.ri(w, of + 2 * c + 1)] = dot(ltw, ct2);
      matched[c] = 1;
      if (b0 >= 0 && sup0) support[b0] += -ndu * lnm;
      if (b1 >= 0 && sup1) support[b1] += ndu * lnm;
    }
    for (int c = 0; c < n_c; ++c) {
      if (matched[c]) continue;
      size_t k = size_t(w) * K + c;
      int s0 = Bt.cs0[k], s1 = Bt.cs1[k];
      if (s0 == s1) continue;
      int b0 = s0 >= 0 ? Bt.shape_body(w, s0) : -1, b1 = s1 >= 0 ? Bt.shape_body(w, s1) : -1;
      V3<double> n = Bt.cvec(Bt.cn, w, c);
      V3<double> p = ws_midpoint(Bt, w, c, Bt.q_prev, b0, b1);
      V3<double> v0 = v3<double>(0, 0, 0), v1 = v0;
      if (b0 >= 0) {  // velocity about the body origin, as the reference computes it (:383-392)
        const Sp<double>& tw = Bt.u_prev[Bt.bi(w, b0)];
        v0 = add(top(tw), cross(bot(tw), sub(p, Bt.q_prev[Bt.bi(w, b0)].p)));
      }
      if (b1 >= 0) {
        const Sp<double>& tw = Bt.u_prev[Bt.bi(w, b1)];
        v1 = add(top(tw), cross(bot(tw), sub(p, Bt.q_prev[Bt.bi(w, b1)].p)));
      }
      V3<double> vr = sub(v1, v0);
      double vn = dot(vr, n);
      V3<double> vtt = sub(vr, mul(n, vn));
      double l_imp = 0;
      if (cold_impact && vn < 0) {
        double im = (b0 >= 0 ? Bt.inv_mass(w, b0) : 0.0) + (b1 >= 0 ? Bt.inv_mass(w, b1) : 0.0);
        if (im > 1e-9) l_imp = (1.0 / im) * (-vn) / dt;
      }
      double ndu = dot(n, up), l_grav = 0;
      if (cold_gravity) {
        if (b0 >= 0 && -ndu > kWsNDotUpMin && unmatched[b0] > 0) {
          double res = Bt.mass(w, b0) * gmag - support[b0];
          if (res > 0) l_grav = std::max(l_grav, res / (double(unmatched[b0]) * (-ndu)));
        }
        if (b1 >= 0 && ndu > kWsNDotUpMin && unmatched[b1] > 0) {
          double res = Bt.mass(w, b1) * gmag - support[b1];
          if (res > 0) l_grav = std::max(l_grav, res / (double(unmatched[b1]) * ndu));
        }
      }
      double ln = std::max(l_imp, l_grav);
      if (ln > 0) seed[Bt.ri(w, on + c)] = ln;
      if (v_t_threshold > 0 && ln > 0) {
        double vsq = dot(vtt, vtt);
        if (vsq > v_t_threshold * v_t_threshold) {
          V3<double> th = mul(vtt, -1.0 / std::sqrt(vsq));
          double mu = 0.5 * (Bt.smu(w, s0) + Bt.smu(w, s1));
          V3<double> ft = mul(th, mu * ln);
          V3<double> t1, t2;
          orthogonal_basis(n, t1, t2);
          seed[Bt.ri(w, of + 2 * c)] = dot(ft, t1);
          seed[Bt.ri(w, of + 2 * c + 1)] = dot(ft, t2);
        }
      }
    }
  }
}

// snapshot(): _snapshot_kernel (:465-519) with the step's lambda
void ws_snapshot(const Batch<double>& Bt, const double* lam, const WSPrev& P) {
  const int K = Bt.Cmax;
  int on = Bt.Nj + Bt.Nctrl, of = on + K;
  for (int w = 0; w < Bt.W; ++w) {
    const int n_c = std::min(Bt.ccount[w], K);
    P.count[w] = n_c;
    for (int c = 0; c < n_c; ++c) {
      size_t k = size_t(w) * K + c;
      int s0 = Bt.cs0[k], s1 = Bt.cs1[k];
      if (s0 == s1) { P.b0[k] = -2; P.b1[k] = -2; continue; }  // never matched
      int b0 = s0 >= 0 ? Bt.shape_body(w, s0) : -1, b1 = s1 >= 0 ? Bt.shape_body(w, s1) : -1;
      V3<double> m = ws_midpoint(Bt, w, c, Bt.q_prev, b0, b1);
      V3<double> n = Bt.cvec(Bt.cn, w, c);
      P.b0[k] = b0; P.b1[k] = b1;
      P.p[k * 3] = m.x; P.p[k * 3 + 1] = m.y; P.p[k * 3 + 2] = m.z;
      P.n[k * 3] = n.x; P.n[k * 3 + 1] = n.y; P.n[k * 3 + 2] = n.z;
      P.ln[k] = lam[Bt.ri(w, on + c)];
      P.lt[k * 2] = lam[Bt.ri(w, of + 2 * c)];
      P.lt[k * 2 + 1] = lam[Bt.ri(w, of + 2 * c + 1)];
    }
  }
}

}  // namespace

// Packed contacts after packing and the configured reduction (body poses q [W*B*7] of
// the step start): counts [W] and the slot arrays [W*Cmax(*3)].
template <class T> int run_pack(const ofb_model* m, const ofb_config* cfg, const double* q,
                                const ofb_contacts* contacts, int32_t* count, int32_t* shape0, int32_t* shape1,
                                double* point0, double* point1, double* normal, double* th0, double* th1) {
  Batch<T> Bt(*m, *cfg);
  pack_contacts(Bt, *contacts);
  reduce_contacts(Bt, q);
  const size_t n = size_t(Bt.W) * Bt.Cmax;
  for (int w = 0; w < Bt.W; ++w) count[w] = Bt.ccount[w];
  for (size_t k = 0; k < n; ++k) {
    shape0[k] = Bt.cs0[k]; shape1[k] = Bt.cs1[k];
    th0[k] = double(Bt.ct0[k]); th1[k] = double(Bt.ct1[k]);
    for (int d = 0; d < 3; ++d) {
      point0[k * 3 + d] = double(Bt.cp0[k * 3 + d]);
      point1[k * 3 + d] = double(Bt.cp1[k * 3 + d]);
      normal[k * 3 + d] = double(Bt.cn[k * 3 + d]);
    }
  }
  return 0;
}

extern "C" {

// threads of the per-world loops (OpenMP); returns the number in effect
int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

// Number of constraint rows of the model (core/engine_dims.py:56-61).
int oracle_dims(const ofb_model* m, const ofb_config* cfg, int32_t* n_c, int32_t* n_j, int32_t* n_ctrl) {
  Batch<double> probe(*m, *cfg);
  *n_c = probe.Nc; *n_j = probe.Nj; *n_ctrl = probe.Nctrl;
  return 0;
}

int oracle_pack(const ofb_model* m, const ofb_config* cfg, const double* q, const ofb_contacts* contacts,
                int precision, int32_t* count, int32_t* shape0, int32_t* shape1, double* point0, double* point1,
                double* normal, double* th0, double* th1) {
  if (precision == 64)
    return run_pack<double>(m, cfg, q, contacts, count, shape0, shape1, point0, point1, normal, th0, th1);
  return run_pack<float>(m, cfg, q, contacts, count, shape0, shape1, point0, point1, normal, th0, th1);
}

// One OstrichEngine.step in FP32 (precision 32) or FP64 (precision 64).  State is passed
// as double so FP64 rollouts are not rounded between steps; model and contacts are f32.
// Optional diagnostics: merit of every Newton iteration [W*nr_max_iters], the
// residual [W*(6B+N_c)] of the last iterate and lambda_prev_iter [W*N_c] as the step
// leaves it.
int oracle_step(const ofb_model* m, const ofb_config* cfg, const double* q_in, const double* qd_in,
                const float* f_in, const float* tpos, const float* tvel, const ofb_contacts* contacts,
                float dt, int precision, const double* lam_seed, double* q_out, double* qd_out,
                double* lam_out, int32_t* nr_iters, int32_t* pcr_iters, double* merit, int32_t* best,
                int32_t* flags, int32_t* ccount, double* merit_hist, double* res_out, double* lam_prev_out) {
  if (precision == 64)
    return run_step<double>(m, cfg, q_in, qd_in, f_in, tpos, tvel, contacts, dt, lam_seed, q_out, qd_out, lam_out,
                            nr_iters, pcr_iters, merit, best, flags, ccount, merit_hist, res_out, lam_prev_out);
  return run_step<float>(m, cfg, q_in, qd_in, f_in, tpos, tvel, contacts, dt, lam_seed, q_out, qd_out, lam_out,
                         nr_iters, pcr_iters, merit, best, flags, ccount, merit_hist, res_out, lam_prev_out);
}

// Scaled FB value and derivatives (complementarity.py:22-47), FP64, for pinning tests.
void oracle_fb_diff(double a, double b, double alpha, double beta, double eps_sq, double* out3) {
  fb_diff<double>(a, b, alpha, beta, eps_sq, out3[0], out3[1], out3[2]);
}

// Pose integration (pose_integration.py:6-38), FP64, for pinning tests.
void oracle_integrate_pose(const double* vel6, const double* pose7, const double* com3, double dt, double* out7) {
  Sp<double> v; for (int i = 0; i < 6; ++i) v.v[i] = vel6[i];
  Tf<double> pp{v3<double>(pose7[0], pose7[1], pose7[2]), Qt<double>{pose7[3], pose7[4], pose7[5], pose7[6]}};
  Tf<double> o = integrate_pose(v, pp, v3<double>(com3[0], com3[1], com3[2]), dt);
  out7[0] = o.p.x; out7[1] = o.p.y; out7[2] = o.p.z; out7[3] = o.q.x; out7[4] = o.q.y; out7[5] = o.q.z; out7[6] = o.q.w;
}

// Assembles the linearised system of NR iteration 0 (lambda = 0, q = q_in, u = qd_in)
// and exports the dense Schur matrix A = J M^-1 J^T + C + reg (active rows), the RHS,
// the Jacobi diagonal and the result of the matrix-free operator applied to `x`, for
// brute-force checks of the operator, preconditioner and PCR.  FP64.
int oracle_linear_system(const ofb_model* m, const ofb_config* cfg, const double* q_in, const double* qd_in,
                         const float* f_in, const float* tpos, const float* tvel, const ofb_contacts* contacts,
                         float dt, const double* x, double* A_dense, double* rhs, double* pinv, double* Ax,
                         double* active, double* pcr_x, int32_t* pcr_iters) {
  Batch<double> Bt(*m, *cfg);
  Bt.dt = dt;
  pack_contacts(Bt, *contacts);
  reduce_contacts(Bt, q_in);
  const int W = Bt.W, B = Bt.B, Nc = Bt.Nc;
  for (int w = 0; w < W; ++w) {
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      const double* xx = q_in + k * 7;
      Bt.q_prev[k] = Tf<double>{v3<double>(xx[0], xx[1], xx[2]), Qt<double>{xx[3], xx[4], xx[5], xx[6]}};
      Bt.q[k] = Bt.q_prev[k];
      for (int i = 0; i < 6; ++i) { Bt.u_prev[k].v[i] = qd_in[k * 6 + i]; Bt.u[k].v[i] = qd_in[k * 6 + i]; Bt.f_ext[k].v[i] = f_in[k * 6 + i]; }
    }
    for (int d = 0; d < Bt.D; ++d) { Bt.tpos[size_t(w) * Bt.D + d] = tpos[w * Bt.D + d]; Bt.tvel[size_t(w) * Bt.D + d] = tvel[w * Bt.D + d]; }
    linear_system(Bt, w);
    jacobi_update(Bt, w);
    std::vector<double> xv(Bt.lam.size(), 0), yv(Bt.lam.size(), 0), zv(Bt.lam.size(), 0);
    for (int i = 0; i < Nc; ++i) xv[Bt.ri(w, i)] = x[Bt.ri(w, i)];
    matvec(Bt, w, xv, yv, zv, 1.0, 0.0);
    for (int i = 0; i < Nc; ++i) {
      size_t rr = Bt.ri(w, i);
      rhs[rr] = Bt.rhs[rr]; pinv[rr] = Bt.active[rr] > 0 ? Bt.Pinv[rr] : 0; Ax[rr] = zv[rr]; active[rr] = Bt.active[rr];
      for (int j = 0; j < Nc; ++j) {
        size_t rj = Bt.ri(w, j);
        double val = 0;
        if (Bt.active[rr] > 0 && Bt.active[rj] > 0) {
          for (int s = 0; s < 2; ++s) {
            int bi_ = s == 0 ? Bt.b0[rr] : Bt.b1[rr];
            const Sp<double>& Ji = s == 0 ? Bt.J0[rr] : Bt.J1[rr];
            if (bi_ < 0) continue;
            for (int t = 0; t < 2; ++t) {
              int bj = t == 0 ? Bt.b0[rj] : Bt.b1[rj];
              const Sp<double>& Jj = t == 0 ? Bt.J0[rj] : Bt.J1[rj];
              if (bj != bi_) continue;
              val += sdot(Ji, minv(Bt, w, bi_, Jj));
            }
          }
          if (i == j) val += Bt.C[rr] + cfg->regularization;
        }
        A_dense[(size_t(w) * Nc + i) * Nc + j] = val;
      }
    }
  }
  if (pcr_x) {  // PCR on the assembled systems (config preconditioner and limits)
    std::vector<int> all(W);
    for (int w = 0; w < W; ++w) {
      all[w] = w;
      if (cfg->preconditioner == OFB_PRECOND_PER_BODY_PAIR) pair_update(Bt, w);
      for (int i = 0; i < Nc; ++i) Bt.dlam[Bt.ri(w, i)] = 0;
    }
    pcr_solve(Bt, all, cfg->termination == OFB_TERM_GLOBAL);
    for (int w = 0; w < W; ++w) {
      for (int i = 0; i < Nc; ++i) pcr_x[Bt.ri(w, i)] = Bt.dlam[Bt.ri(w, i)];
      if (pcr_iters) pcr_iters[w] = Bt.pcr_iters[w];
    }
  }
  return 0;
}

// s = sum_rows w_lambda h_c + sum_bodies w_u . h_d of R(q+, q-, u+, u-, f, targets, mu,
// lambda, lambda_prev) in FP64: the quantity whose gradient the adjoint's VJP returns,
// for finite-difference checks of every term.
double oracle_weighted_residual(const ofb_model* m, const ofb_config* cfg, const double* q_prev,
                                const double* qd_prev, const float* f_in, const double* q, const double* qd,
                                const float* tpos, const float* tvel, const ofb_contacts* contacts, float dt,
                                const double* lam, const double* lam_prev, const double* wl_in,
                                const double* wu_in) {
  Batch<double> Bt(*m, *cfg);
  Bt.dt = dt;
  pack_contacts(Bt, *contacts);
  reduce_contacts(Bt, q_prev);
  const int W = Bt.W, B = Bt.B, D = Bt.D, Nc = Bt.Nc;
  std::vector<double> wl(size_t(W) * Nc);
  std::vector<Sp<double>> wu(size_t(W) * B);
  double total = 0;
  for (int w = 0; w < W; ++w) {
    for (int b = 0; b < B; ++b) {
      size_t k = Bt.bi(w, b);
      const double* a = q_prev + k * 7;
      const double* c = q + k * 7;
      Bt.q_prev[k] = Tf<double>{v3<double>(a[0], a[1], a[2]), Qt<double>{a[3], a[4], a[5], a[6]}};
      Bt.q[k] = Tf<double>{v3<double>(c[0], c[1], c[2]), Qt<double>{c[3], c[4], c[5], c[6]}};
      for (int i = 0; i < 6; ++i) {
        Bt.u_prev[k].v[i] = qd_prev[k * 6 + i]; Bt.u[k].v[i] = qd[k * 6 + i]; Bt.f_ext[k].v[i] = f_in[k * 6 + i];
        wu[k].v[i] = wu_in[k * 6 + i];
      }
    }
    for (int d = 0; d < D; ++d) { Bt.tpos[size_t(w) * D + d] = tpos[w * D + d]; Bt.tvel[size_t(w) * D + d] = tvel[w * D + d]; }
    for (int i = 0; i < Nc; ++i) {
      size_t rr = Bt.ri(w, i);
      Bt.lam[rr] = lam[rr]; Bt.lam_prev[rr] = lam_prev[rr]; wl[rr] = wl_in[rr];
    }
    total += adjoint_weighted_residual(Bt, w, wl, wu);
  }
  return total;
}

// Adjoint of one step (OstrichEngine.step_backward), FP32 (precision 32) or FP64 (64).
// Gradients are written in double; any output pointer may be null.  flags bit 4: a
// friction row contributes but its mu derivative has no closed form here (anisotropic
// or Stribeck friction), so that part of g_mu is missing.
int oracle_adjoint(const ofb_model* m, const ofb_config* cfg, const double* q_prev, const double* qd_prev,
                   const float* f_in, const double* q, const double* qd, const float* tpos, const float* tvel,
                   const ofb_contacts* contacts, float dt, const double* lam, const double* lam_prev,
                   const double* g_q, const double* g_u, int precision, double fd_eps, double* g_q_prev,
                   double* g_u_prev, double* g_f, double* g_tpos, double* g_tvel, double* g_ke, double* g_kd,
                   double* g_mu, double* w_lam, double* w_u, int32_t* pcr_iters, int32_t* flags) {
  if (precision == 64)
    return run_adjoint<double>(m, cfg, q_prev, qd_prev, f_in, q, qd, tpos, tvel, contacts, dt, lam, lam_prev, g_q,
                               g_u, fd_eps, g_q_prev, g_u_prev, g_f, g_tpos, g_tvel, g_ke, g_kd, g_mu, w_lam, w_u,
                               pcr_iters, flags);
  return run_adjoint<float>(m, cfg, q_prev, qd_prev, f_in, q, qd, tpos, tvel, contacts, dt, lam, lam_prev, g_q, g_u,
                            fd_eps, g_q_prev, g_u_prev, g_f, g_tpos, g_tvel, g_ke, g_kd, g_mu, w_lam, w_u, pcr_iters,
                            flags);
}

// The merit line search's candidate steps of a config; returns their number.
int oracle_linesearch_steps(const ofb_config* cfg, float* out) {
  std::vector<float> v = linesearch_step_list(*cfg);
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  return (int)v.size();
}

// Warm start of one step: mode 0 = apply (writes seed [W*N_c] from the snapshot), mode 1 =
// snapshot (stores the step's contacts and lambda [W*N_c]).  State is the step-start
// state (q-, u-).  FP64.
int oracle_warm_start(const ofb_model* m, const ofb_config* cfg, const double* q_prev, const double* qd_prev,
                      const ofb_contacts* contacts, float dt, int mode, const double* lam, int32_t* prev_count,
                      int32_t* prev_b0, int32_t* prev_b1, double* prev_p, double* prev_n, double* prev_ln,
                      double* prev_lt, double* seed) {
  Batch<double> Bt(*m, *cfg);
  Bt.dt = dt;
  pack_contacts(Bt, *contacts);
  reduce_contacts(Bt, q_prev);
  for (int w = 0; w < Bt.W; ++w)
    for (int b = 0; b < Bt.B; ++b) {
      size_t k = Bt.bi(w, b);
      const double* x = q_prev + k * 7;
      Bt.q_prev[k] = Tf<double>{v3<double>(x[0], x[1], x[2]), Qt<double>{x[3], x[4], x[5], x[6]}};
      for (int i = 0; i < 6; ++i) Bt.u_prev[k].v[i] = qd_prev[k * 6 + i];
    }
  WSPrev P{prev_count, prev_b0, prev_b1, prev_p, prev_n, prev_ln, prev_lt};
  if (mode == 0) ws_apply(Bt, P, cfg->warm_start_method, cfg->warm_start_cold_gravity, cfg->warm_start_cold_impact,
                          cfg->warm_start_cold_friction_v, seed);
  else ws_snapshot(Bt, lam, P);
  return 0;
}

}  // extern "C"
