// fb_fused.hip -- fused persistent Fisher-Burmeister Newton solver for gfx950.
//
// One wavefront (64 lanes) owns one world for the whole step: the Newton loop
// (src/ostrich/core/base_engine.py:416-526), the linear-system assembly
// (core/linear_system.py:140-341), the Jacobi-preconditioned conjugate-residual solve
// (optim/pcr_solver.py:170-295), the Newton update (core/newton_step.py:22-54), the
// merit evaluation (core/residual.py:18-173) and best-iterate backtracking
// (core/backtracking.py:12-95) all run without leaving the kernel.
//
// Layout on chip:
//   * constraint rows of the world are compacted to
//     [joints | control | normals(n_c) | friction(2 n_c)] and dealt to lanes,
//     row i -> lane i % 64, slot i / 64; every per-row quantity (Jacobian pair, C,
//     Jacobi diagonal, the seven PCR vectors, lambda) lives in VGPRs, so PCR vector
//     updates are pure VALU and dot products are in-register partial sums followed by
//     one DPP cross-lane reduction;
//   * body state (pose, velocity, previous step, world inverse inertia, the J^T x
//     accumulator) lives in LDS, 96 floats per body, and couples the rows: the Schur
//     matvec J M^-1 J^T x is a scatter of J^T x into LDS (ds_add_f32 from one wave, so
//     the accumulation order is fixed), a per-body M^-1 and a gather;
//   * a world leaves the Newton loop on its own criterion (per-world termination), and
//     its wave retires, freeing the SIMD for the next world.
// Waves of one workgroup never synchronise with each other; every LDS exchange is
// wave-local.

#include <hip/hip_runtime.h>

#include "fb_math.h"
#include "fb_rows.h"
#include "fb_solver.h"

namespace ofb {

using namespace rows;

#ifdef OFB_PHASE_TIMING
// Profiling build only: per-phase cycle counts of the Newton loop, summed over worlds.
__device__ unsigned long long g_phase_cycles[8];
#define OFB_PHASE_INIT unsigned long long ph_[8] = {0, 0, 0, 0, 0, 0, 0, 0}; unsigned long long ph_t = clock64();
#define OFB_PHASE(k) { unsigned long long t_ = clock64(); ph_[k] += t_ - ph_t; ph_t = t_; }
#define OFB_PHASE_FLUSH if (lane == 0) for (int k_ = 0; k_ < 8; ++k_) atomicAdd(&g_phase_cycles[k_], ph_[k_]);
#else
#define OFB_PHASE_INIT
#define OFB_PHASE(k)
#define OFB_PHASE_FLUSH
#endif

namespace {

struct FusedParams {
  DeviceModel m;
  RowLayout L;
  ofb_config cfg;
  float dt;
  const float* body_q;
  const float* body_qd;
  const float* body_f;
  const float* target_pos;
  const float* target_vel;
  float* out_q;
  float* out_qd;
  const int32_t* ccount;
  const float* cf;
  const int32_t* ci;
  const int32_t* topo;  // [Nj] joint of each joint row, [Nj] row within joint, [Nctrl] joint of each control row
  const float* lambda_seed;
  float* lambda_out;
  int32_t *st_nr, *st_pcr, *st_best, *st_flags;
  float* st_merit;
  int lds_per_wave;  // floats
  const int32_t* wlist;   // worlds of this launch's row class
  const int32_t* wcount;  // number of them
  int bj;                 // per-body-pair block-Jacobi preconditioner
  int rows_cap;           // 64 R: row capacity of this launch's class
};

// ---------------------------------------------------------------------------------
// Per-body-pair block-Jacobi inside one wavefront (optim/per_body_pair_preconditioner.py:
// 62-410).  Rows that couple the same two bodies (static side = body B) form a pair;
// the pair's lowest row is its leader.  Once per Newton iteration each leader builds
// its dense block A_pair = sum_b J_a,b M_b^-1 J_c,b^T (+ C + reg on the diagonal) in
// LDS and factors it in place (column Cholesky, packed lower triangle); in every PCR
// iteration the leader runs both triangular solves for its members.  A pair that fails
// to factor uses the Jacobi diagonal for its rows, and rows ranked past 64 in a pair
// use the Jacobi diagonal too (the reference's fixed 64-member blocks,
// per_body_pair_preconditioner.py:59).
// LDS per wave after the bodies and lambda_prev (cap = 64 R rows):
//   pid[cap] i32, lsz[cap] i32, mlist[cap] i32, wbuf[cap], vbuf[cap],
//   rowd[16 cap] (J0 6, J1 6, b0, b1, C+reg, Jacobi 1/diag), Lbuf[R * 2080]
// ---------------------------------------------------------------------------------
constexpr int kPairCap = 64;
constexpr int kPairL = kPairCap * (kPairCap + 1) / 2;

__host__ __device__ inline int bj_words(int cap) { return 21 * cap + (cap / 64) * kPairL; }

struct BJView {
  int* pid;
  int* lsz;
  int* mlist;
  float* wbuf;
  float* vbuf;
  float* rowd;
  float* Lbuf;
  int cap;
};

__device__ __forceinline__ BJView bj_view(float* base, int cap) {
  BJView v;
  v.pid = reinterpret_cast<int*>(base);
  v.lsz = v.pid + cap;
  v.mlist = v.lsz + cap;
  v.wbuf = reinterpret_cast<float*>(v.mlist + cap);
  v.vbuf = v.wbuf + cap;
  v.rowd = v.vbuf + cap;
  v.Lbuf = v.rowd + 16 * cap;
  return v;
}

// J of row `row` on body `tb` (_jacobian_for_body, per_body_pair_preconditioner.py:146-184)
__device__ __forceinline__ sv bj_jac(const BJView& V, int row, int tb, int B) {
  const float* d = V.rowd + 16 * row;
  if (tb >= B) return svzero();
  if (__float_as_int(d[12]) == tb) return sv{mk(d[0], d[1], d[2]), mk(d[3], d[4], d[5])};
  if (__float_as_int(d[13]) == tb) return sv{mk(d[6], d[7], d[8]), mk(d[9], d[10], d[11])};
  return svzero();
}

// Per-row-slot block-Jacobi bookkeeping kept in registers.
template <int R>
struct BJRegs {
  int rank[R];   // rank of the row inside its pair
  int moff[R];   // offset of the pair's member list / solve workspace
  int lead[R];   // member count if the row leads its pair, else 0
  int loff[R];   // offset of the pair's factor (leaders)
  int fail[R];   // factorisation failed (leaders)
};

template <int R>
__device__ void bj_setup(const BJView& V, const float* sm, int lane, int n_r, int B, const bool (&act)[R],
                         const int (&b0)[R], const int (&b1)[R], const sv (&J0)[R], const sv (&J1)[R],
                         const float (&Cr)[R], const float (&Pi)[R], BJRegs<R>& G, int& flags) {
  int mypid[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    mypid[k] = -1;
    if (i < n_r) {
      if (act[k]) {  // _compute_pair_ids_kernel :62-109
        int e1 = b0[k] < 0 ? B : b0[k], e2 = b1[k] < 0 ? B : b1[k];
        int lo = min(e1, e2), hi = max(e1, e2);
        mypid[k] = lo * (B + 1) + hi;
      }
      V.pid[i] = mypid[k];
      float* d = V.rowd + 16 * i;
      d[0] = J0[k].a.x; d[1] = J0[k].a.y; d[2] = J0[k].a.z; d[3] = J0[k].b.x; d[4] = J0[k].b.y; d[5] = J0[k].b.z;
      d[6] = J1[k].a.x; d[7] = J1[k].a.y; d[8] = J1[k].a.z; d[9] = J1[k].b.x; d[10] = J1[k].b.y; d[11] = J1[k].b.z;
      d[12] = __int_as_float(b0[k]); d[13] = __int_as_float(b1[k]);
      d[14] = Cr[k]; d[15] = Pi[k];
    }
  }
  wsync();
  // rank and size of every row's pair; leaders publish their size
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    int rank = 0, size = 0;
    if (i < n_r && mypid[k] >= 0) {
      for (int j = 0; j < n_r; ++j) {
        int pj = V.pid[j];
        if (pj == mypid[k]) { size += 1; if (j < i) rank += 1; }
      }
    }
    G.rank[k] = rank;
    G.lead[k] = (i < n_r && mypid[k] >= 0 && rank == 0) ? size : 0;
    if (i < n_r) V.lsz[i] = G.lead[k];
    if (G.lead[k] > kPairCap) flags |= FLAG_PAIR_OVERFLOW;
  }
  wsync();
  // offsets of member lists and factors; leaders write their member lists
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    int mo = 0, lo = 0;
    if (i < n_r && mypid[k] >= 0) {
      // the pair's leader is the first row with this pair id
      int leader = i;
      for (int j = 0; j < i; ++j) if (V.pid[j] == mypid[k]) { leader = j; break; }
      for (int j = 0; j < leader; ++j) {
        int s = V.lsz[j];
        int sc = min(s, kPairCap);
        mo += sc;
        lo += sc * (sc + 1) / 2;
      }
    }
    G.moff[k] = mo;
    G.loff[k] = lo;
    if (G.lead[k] > 0) {
      int cnt = 0;
      for (int j = i; j < n_r && cnt < kPairCap; ++j)
        if (V.pid[j] == mypid[k]) V.mlist[mo + cnt++] = j;
    }
  }
  wsync();
  // leaders: dense block and in-place column Cholesky (:187-321)
#pragma unroll
  for (int k = 0; k < R; ++k) {
    G.fail[k] = 0;
    int m = min(G.lead[k], kPairCap);
    if (m > 0) {
      int i = lane + WAVE * k;
      int e1 = b0[k] < 0 ? B : b0[k], e2 = b1[k] < 0 ? B : b1[k];
      int blo = min(e1, e2), bhi = max(e1, e2);
      (void)i;
      float* Lp = V.Lbuf + G.loff[k];
      const int* mem = V.mlist + G.moff[k];
      for (int a = 0; a < m; ++a) {
        int ra = mem[a];
        for (int c = 0; c <= a; ++c) {
          int rc = mem[c];
          float val = 0.f;
          if (blo < B) val += sdot(bj_jac(V, ra, blo, B), apply_minv(sm + blo * BS, bj_jac(V, rc, blo, B)));
          if (bhi < B) val += sdot(bj_jac(V, ra, bhi, B), apply_minv(sm + bhi * BS, bj_jac(V, rc, bhi, B)));
          if (a == c) val += V.rowd[16 * ra + 14];
          Lp[a * (a + 1) / 2 + c] = val;
        }
      }
      for (int kk = 0; kk < m; ++kk) {
        float ds = 0.f;
        for (int q = 0; q < kk; ++q) { float l = Lp[kk * (kk + 1) / 2 + q]; ds += l * l; }
        float dv = Lp[kk * (kk + 1) / 2 + kk] - ds;
        if (dv <= 0.f) { G.fail[k] = 1; flags |= FLAG_FACTOR_FAILURE; break; }
        float Lkk = sqrtf(dv);
        Lp[kk * (kk + 1) / 2 + kk] = Lkk;
        float inv = 1.f / Lkk;
        for (int a = kk + 1; a < m; ++a) {
          float rs = 0.f;
          for (int q = 0; q < kk; ++q) rs += Lp[a * (a + 1) / 2 + q] * Lp[kk * (kk + 1) / 2 + q];
          Lp[a * (a + 1) / 2 + kk] = (Lp[a * (a + 1) / 2 + kk] - rs) * inv;
        }
      }
    }
  }
  wsync();
}

// out = P v (apply of per_body_pair_preconditioner.py:324-410)
template <int R>
__device__ void bj_apply(const BJView& V, int lane, int n_r, const bool (&act)[R], const float (&Pi)[R],
                         const BJRegs<R>& G, const float (&v)[R], float (&out)[R]) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    if (i < n_r) V.vbuf[i] = v[k];
  }
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int m = min(G.lead[k], kPairCap);
    if (m > 0) {
      const int* mem = V.mlist + G.moff[k];
      float* ws = V.wbuf + G.moff[k];
      if (G.fail[k]) {
        for (int a = 0; a < m; ++a) ws[a] = V.rowd[16 * mem[a] + 15] * V.vbuf[mem[a]];
      } else {
        const float* Lp = V.Lbuf + G.loff[k];
        for (int a = 0; a < m; ++a) {
          float s = V.vbuf[mem[a]];
          for (int c = 0; c < a; ++c) s -= Lp[a * (a + 1) / 2 + c] * ws[c];
          ws[a] = s / Lp[a * (a + 1) / 2 + a];
        }
        for (int aa = 0; aa < m; ++aa) {
          int a = m - 1 - aa;
          float s = ws[a];
          for (int c = a + 1; c < m; ++c) s -= Lp[c * (c + 1) / 2 + a] * ws[c];
          ws[a] = s / Lp[a * (a + 1) / 2 + a];
        }
      }
    }
  }
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float o = 0.f;
    if (act[k]) o = G.rank[k] < kPairCap ? V.wbuf[G.moff[k] + G.rank[k]] : Pi[k] * v[k];
    out[k] = o;
  }
  wsync();
}

// out = (J M^-1 J^T + C + reg) v over the rows of one world (optim/system_operator.py:36-124):
// scatter J^T v into the LDS body accumulator, apply M^-1 per body, gather.
template <int R>
__device__ __forceinline__ void schur_matvec(float* sm, int lane, int B, const bool (&act)[R], const int (&b0)[R],
                                             const int (&b1)[R], const sv (&J0)[R], const sv (&J1)[R],
                                             const float (&Cr)[R], const float (&v)[R], float (&out)[R]) {
  for (int e = lane; e < 6 * B; e += WAVE) sm[(e / 6) * BS + BACC + (e % 6)] = 0.f;
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (act[k]) {
      if (b0[k] >= 0) atomic_add_sv(sm + b0[k] * BS + BACC, J0[k] * v[k]);
      if (b1[k] >= 0) atomic_add_sv(sm + b1[k] * BS + BACC, J1[k] * v[k]);
    }
  }
  wsync();
  for (int b = lane; b < B; b += WAVE) {
    float* bb = sm + b * BS;
    stsv(bb + BVV, apply_minv(bb, ldsv(bb + BACC)));
  }
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float o = 0.f;
    if (act[k]) {
      if (b0[k] >= 0) o += sdot(J0[k], ldsv(sm + b0[k] * BS + BVV));
      if (b1[k] >= 0) o += sdot(J1[k], ldsv(sm + b1[k] * BS + BVV));
      o += Cr[k] * v[k];
    }
    out[k] = o;
  }
  wsync();
}

}  // namespace

// ------------------------------------------------------------------------------------
// The kernel.  R = row slots per lane (rows per world <= 64 R).
// ------------------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(256) fused_step_kernel(FusedParams P) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = threadIdx.x / WAVE;
  const int idx = blockIdx.x * (blockDim.x / WAVE) + wave;
  if (idx >= *P.wcount) return;
  const int w = P.wlist[idx];
  float* sm = lds + wave * P.lds_per_wave;
  const int B = P.m.B;
  const RowLayout& L = P.L;
  float* lamprev = sm + B * BS;  // compacted row index
  const float dt = P.dt;
  const BJView V = bj_view(lamprev + L.Nc + 1, P.rows_cap);
  BJRegs<R> G;
  int bj_flags = 0;

  // ---- load bodies (load_data, base_engine.py:303-308; engine.py:48-50) ----
  for (int b = lane; b < B; b += WAVE) {
    float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
#pragma unroll
    for (int i = 0; i < 7; ++i) { float v = P.body_q[k * 7 + i]; bb[BQ + i] = v; bb[BQP + i] = v; }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      float v = P.body_qd[k * 6 + i];
      bb[BU + i] = v; bb[BUP + i] = v;
      bb[BF + i] = P.body_f[k * 6 + i];
    }
    bb[BM] = P.m.body_mass[k];
    bb[BMINV] = P.m.body_inv_mass[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) bb[BCOM + i] = P.m.body_com[k * 3 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) { bb[BIB + i] = P.m.body_inertia[k * 9 + i]; bb[BIIB + i] = P.m.body_inv_inertia[k * 9 + i]; }
  }
  const int n_c = min(P.ccount[w], L.K);
  const int n_r = L.Nj + L.Nctrl + 3 * n_c;

  // ---- per-row registers ----
  float lam[R], lamb[R], x[R], r[R], z[R], p[R], Ap[R], y[R], Az[R], Cr[R], Pi[R], rhs[R];
  sv J0[R], J1[R];
  int b0[R], b1[R];
  bool act[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    float l0 = 0.f;
    if (P.cfg.seed_iterate && i < n_r) l0 = P.lambda_seed[(size_t)w * L.Nc + original_row(L, i, n_c)];
    lam[k] = l0; lamb[k] = l0;
    act[k] = false; b0[k] = -1; b1[k] = -1;
  }
  wsync();

  const float atol_sq = P.cfg.nr_atol * P.cfg.nr_atol;
  const float tol_sq = P.cfg.lin_tol >= 0.f ? P.cfg.lin_tol * P.cfg.lin_tol : -1.f;
  const float latol_sq = P.cfg.lin_atol >= 0.f ? P.cfg.lin_atol * P.cfg.lin_atol : -1.f;
  const float reg = P.cfg.regularization;
  int n_it = 0, pcr_total = 0, best_idx = -1;
  float best_merit = 0.f, merit = 0.f;

  OFB_PHASE_INIT
  while (true) {
    // ---- friction-lag snapshot (base_engine.py:441) ----
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + WAVE * k;
      if (i < n_r) lamprev[i] = lam[k];
    }
    dynamics_bodies(P, sm, lane, true);
    wsync();
    OFB_PHASE(0)
    // ---- assembly (linear_system.py:140-322) ----
    float hc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + WAVE * k;
      hc[k] = 0.f;
      if (i < n_r) {
        RowOut o = assemble_row(P, sm, lamprev, w, i, n_c, lam[k]);
        J0[k] = o.J0; J1[k] = o.J1; b0[k] = o.b0; b1[k] = o.b1; act[k] = o.active;
        hc[k] = o.hc;
        Cr[k] = o.C + reg;
        // inactive contact / friction rows force lambda := 0 (contact_constraint.py:289-296,
        // friction_constraint.py:713-787)
        if (!o.active && i >= L.Nj + L.Nctrl) lam[k] = 0.f;
      } else {
        act[k] = false;
      }
    }
    wsync();
    OFB_PHASE(1)
    // M^-1 h_d per body
    for (int b = lane; b < B; b += WAVE) {
      float* bb = sm + b * BS;
      stsv(bb + BVV, apply_minv(bb, ldsv(bb + BHD)));
    }
    wsync();
    // Schur RHS (linear_system.py:17-65) and Jacobi diagonal (preconditioner.py:9-50)
    float bsq = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      rhs[k] = 0.f; Pi[k] = 0.f;
      if (act[k]) {
        float s = 0.f, d = 0.f;
        if (b0[k] >= 0) {
          const float* bb = sm + b0[k] * BS;
          s += sdot(J0[k], ldsv(bb + BVV));
          d += sdot(J0[k], apply_minv(bb, J0[k]));
        }
        if (b1[k] >= 0) {
          const float* bb = sm + b1[k] * BS;
          s += sdot(J1[k], ldsv(bb + BVV));
          d += sdot(J1[k], apply_minv(bb, J1[k]));
        }
        rhs[k] = (s - hc[k]) / dt;
        Pi[k] = 1.f / (d + Cr[k] + 1e-6f);
      }
      bsq += rhs[k] * rhs[k];
    }
    bsq = wave_sum(bsq);
    if (P.bj) bj_setup<R>(V, sm, lane, n_r, B, act, b0, b1, J0, J1, Cr, Pi, G, bj_flags);

    OFB_PHASE(2)
    // ---- PCR (optim/pcr_solver.py:170-295), x0 = 0 => r = b ----
#pragma unroll
    for (int k = 0; k < R; ++k) { x[k] = 0.f; r[k] = rhs[k]; z[k] = Pi[k] * r[k]; }
    if (P.bj) bj_apply<R>(V, lane, n_r, act, Pi, G, r, z);
    schur_matvec<R>(sm, lane, B, act, b0, b1, J0, J1, Cr, z, Az);
    float zAz = 0.f, rsq = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) { p[k] = z[k]; Ap[k] = Az[k]; zAz += z[k] * Az[k]; rsq += r[k] * r[k]; }
    zAz = wave_sum(zAz);
    rsq = wave_sum(rsq);
    float target = latol_sq;
    if (bsq * tol_sq > target) target = bsq * tol_sq;
    int it = 0;
    while (true) {
      float zAz_old = zAz;
      float yAp = 0.f;
#pragma unroll
      for (int k = 0; k < R; ++k) y[k] = Pi[k] * Ap[k];
      if (P.bj) bj_apply<R>(V, lane, n_r, act, Pi, G, Ap, y);
#pragma unroll
      for (int k = 0; k < R; ++k) yAp += y[k] * Ap[k];
      yAp = wave_sum(yAp);
      float a = yAp > 0.f ? zAz_old / yAp : 0.f;
#pragma unroll
      for (int k = 0; k < R; ++k) { x[k] += a * p[k]; r[k] -= a * Ap[k]; z[k] -= a * y[k]; }
      schur_matvec<R>(sm, lane, B, act, b0, b1, J0, J1, Cr, z, Az);
      zAz = 0.f;
#pragma unroll
      for (int k = 0; k < R; ++k) zAz += z[k] * Az[k];
      zAz = wave_sum(zAz);
      float bt = zAz_old > 0.f ? zAz / zAz_old : 0.f;
      rsq = 0.f;
#pragma unroll
      for (int k = 0; k < R; ++k) { p[k] = z[k] + bt * p[k]; Ap[k] = Az[k] + bt * Ap[k]; rsq += r[k] * r[k]; }
      rsq = wave_sum(rsq);
      ++it;
      if (!(rsq > target) || it >= P.cfg.lin_max_iters) break;
    }
    pcr_total += it;
    OFB_PHASE(3)

    // ---- velocity recovery + Newton update (linear_system.py:344-377, newton_step.py:22-54) ----
    for (int e = lane; e < 6 * B; e += WAVE) sm[(e / 6) * BS + BACC + (e % 6)] = 0.f;
    wsync();
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (act[k]) {
        if (b0[k] >= 0) atomic_add_sv(sm + b0[k] * BS + BACC, J0[k] * x[k]);
        if (b1[k] >= 0) atomic_add_sv(sm + b1[k] * BS + BACC, J1[k] * x[k]);
      }
      lam[k] = lam[k] + x[k];
    }
    wsync();
    for (int b = lane; b < B; b += WAVE) {
      float* bb = sm + b * BS;
      sv acc = ldsv(bb + BACC), hd = ldsv(bb + BHD);
      sv du = apply_minv(bb, sv{acc.a * dt - hd.a, acc.b * dt - hd.b});
      sv u = ldsv(bb + BU);
      u = sv{u.a + du.a, u.b + du.b};
      stsv(bb + BU, u);
      sttf(bb + BQ, integrate(u, ldtf(bb + BQP), ld3(bb + BCOM), dt));
    }
    wsync();

    OFB_PHASE(4)
    // ---- residual at the new iterate (core/residual.py:18-173) ----
    dynamics_bodies(P, sm, lane, false);
    wsync();
    float msum = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + WAVE * k;
      if (i < n_r) {
        RowOut o = assemble_row(P, sm, lamprev, w, i, n_c, lam[k]);
        msum += o.hc * o.hc;
      }
    }
    wsync();
    for (int e = lane; e < 6 * B; e += WAVE) {
      float v = sm[(e / 6) * BS + BHD + (e % 6)];
      msum += v * v;
    }
    merit = wave_sum(msum);
    OFB_PHASE(5)

    // ---- candidate bookkeeping (history_group.py + backtracking.py:12-95) ----
    int kidx = n_it;
    ++n_it;
    bool take = (kidx == P.cfg.nr_backtrack_min_iter) || (kidx > P.cfg.nr_backtrack_min_iter && merit < best_merit);
    if (take) {
      best_idx = kidx;
      best_merit = merit;
      for (int e = lane; e < 13 * B; e += WAVE) {
        int b = e / 13, c = e % 13;
        float* bb = sm + b * BS;
        if (c < 7) bb[BBQ + c] = bb[BQ + c];
        else bb[BBU + c - 7] = bb[BU + c - 7];
      }
#pragma unroll
      for (int k = 0; k < R; ++k) lamb[k] = lam[k];
    }
    // ---- per-world stop rule (base_engine.py:42-80) ----
    bool keep = merit > atol_sq;
    if (n_it < P.cfg.nr_backtrack_min_iter) keep = true;
    if (n_it >= P.cfg.nr_max_iters) keep = false;
    OFB_PHASE(6)
    if (!keep) break;
    wsync();
  }
  wsync();
  OFB_PHASE_FLUSH
  // ---- pick the candidate ----
  int picked = n_it - 1;
  float picked_merit = merit;
  if (n_it > P.cfg.nr_backtrack_min_iter && best_idx >= 0) {
    picked = best_idx;
    picked_merit = best_merit;
    for (int e = lane; e < 13 * B; e += WAVE) {
      int b = e / 13, c = e % 13;
      float* bb = sm + b * BS;
      if (c < 7) bb[BQ + c] = bb[BBQ + c];
      else bb[BU + c - 7] = bb[BBU + c - 7];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) lam[k] = lamb[k];
  }
  wsync();
  // ---- outputs (engine.py:68-69) ----
  for (int e = lane; e < 13 * B; e += WAVE) {
    int b = e / 13, c = e % 13;
    const float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
    if (c < 7) P.out_q[k * 7 + c] = bb[BQ + c];
    else P.out_qd[k * 6 + c - 7] = bb[BU + c - 7];
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + WAVE * k;
    if (i < n_r) P.lambda_out[(size_t)w * L.Nc + original_row(L, i, n_c)] = lam[k];
  }
  // contact slots beyond the world's contact count hold lambda = 0
  for (int c = n_c + lane; c < L.K; c += WAVE) {
    P.lambda_out[(size_t)w * L.Nc + L.off_n + c] = 0.f;
    P.lambda_out[(size_t)w * L.Nc + L.off_f + 2 * c] = 0.f;
    P.lambda_out[(size_t)w * L.Nc + L.off_f + 2 * c + 1] = 0.f;
  }
  int bj_any = 0;
  if (__any(bj_flags & FLAG_FACTOR_FAILURE)) bj_any |= FLAG_FACTOR_FAILURE;
  if (__any(bj_flags & FLAG_PAIR_OVERFLOW)) bj_any |= FLAG_PAIR_OVERFLOW;
  if (lane == 0) {
    P.st_nr[w] = n_it;
    P.st_pcr[w] = pcr_total;
    P.st_merit[w] = picked_merit;
    P.st_best[w] = picked;
    int fl = (P.st_flags[w] & FLAG_CONTACT_OVERFLOW) | bj_any;
    if (n_it >= P.cfg.nr_max_iters && merit > atol_sq) fl |= FLAG_NR_CAP;
    P.st_flags[w] = fl;
  }
}

// Row classes: a world with n_r active rows runs the instance with R = the smallest
// class >= ceil(n_r / 64) slots per lane, so small worlds do not pay for the padded
// row count's registers.  Worlds are bucketed on the device every step.
constexpr int kNumClasses = 11;
__constant__ int kClassR[kNumClasses] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16};
static const int kClassRHost[kNumClasses] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16};

__global__ void reset_classes_kernel(int32_t* counts) {
  if (threadIdx.x < kNumClasses) counts[threadIdx.x] = 0;
}

__global__ void classify_worlds_kernel(int W, const int32_t* __restrict__ ccount, int K, int base_rows,
                                       int32_t* __restrict__ counts, int32_t* __restrict__ lists) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  int n_r = base_rows + 3 * min(ccount[w], K);
  int r = max(1, (n_r + WAVE - 1) / WAVE);
  int c = 0;
  while (c < kNumClasses - 1 && kClassR[c] < r) ++c;
  int slot = atomicAdd(counts + c, 1);
  lists[(size_t)c * W + slot] = w;
}

// LDS floats per wave of the class with R row slots per lane
static int lds_words(const ofb_solver* s, int R) {
  int base = s->m.B * BS + s->L.Nc + 1;
  return base + (s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR ? bj_words(WAVE * R) : 0);
}

template <int R>
static void launch_fused(FusedParams P, const ofb_solver* s, int cls, hipStream_t stream) {
  P.wlist = s->cls_lists + (size_t)cls * P.m.W;
  P.wcount = s->cls_counts + cls;
  P.rows_cap = WAVE * R;
  P.lds_per_wave = lds_words(s, R);
  // worlds (waves) per workgroup: as many as fit in 64 KB of LDS, at most the tuned value
  int wpb = s->fused_worlds_per_block;
  while (wpb > 1 && (size_t)wpb * P.lds_per_wave * sizeof(float) > 65536) --wpb;
  size_t lds_bytes = (size_t)wpb * P.lds_per_wave * sizeof(float);
  int blocks = (P.m.W + wpb - 1) / wpb;
  hipLaunchKernelGGL(fused_step_kernel<R>, dim3(blocks), dim3(wpb * WAVE), lds_bytes, stream, P);
}

// lets a class kernel use more than 64 KB of LDS when one world needs it
template <int R>
static int allow_lds(const ofb_solver* s) {
  size_t bytes = (size_t)lds_words(s, R) * sizeof(float);
  if (bytes <= 65536) return 0;
  if (hipFuncSetAttribute((const void*)fused_step_kernel<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)bytes) != hipSuccess) {
    set_error("fused kernel needs %zu bytes of LDS per world", bytes);
    return -4;
  }
  return 0;
}

#ifdef OFB_PHASE_TIMING
extern "C" int ofb_debug_phase_cycles(unsigned long long* out, int reset) {
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase_cycles), sizeof(unsigned long long) * 8) != hipSuccess) return -4;
  if (reset) {
    unsigned long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_phase_cycles), z, sizeof(z)) != hipSuccess) return -4;
  }
  return 0;
}
#endif

int fused_classes_alloc(ofb_solver* s) {
  int rc = 0;
  const int r_max = (s->L.Nc + WAVE - 1) / WAVE;
  for (int c = 0; c < kNumClasses; ++c) {
    if (c > 0 && kClassRHost[c - 1] >= r_max) break;
    switch (kClassRHost[c]) {
#define CASE(RR) case RR: rc = allow_lds<RR>(s); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(10) CASE(12) CASE(16)
#undef CASE
    }
    if (rc) return rc;
  }
  if (hipMalloc((void**)&s->cls_counts, kNumClasses * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&s->cls_lists, (size_t)kNumClasses * s->m.W * sizeof(int32_t)) != hipSuccess) {
    set_error("hipMalloc failed (world classes)");
    return -4;
  }
  return 0;
}

int fused_step(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl, float dt,
               hipStream_t stream) {
  FusedParams P;
  P.m = s->m;
  P.L = s->L;
  P.cfg = s->cfg;
  P.dt = dt;
  P.body_q = in->body_q; P.body_qd = in->body_qd; P.body_f = in->body_f;
  P.target_pos = ctl->joint_target_pos; P.target_vel = ctl->joint_target_vel;
  P.out_q = out->body_q; P.out_qd = out->body_qd;
  P.ccount = s->ccount; P.cf = s->cf; P.ci = s->ci;
  P.topo = s->topo;
  P.lambda_seed = s->lambda_seed;
  P.lambda_out = s->lambda;
  P.st_nr = s->st_nr; P.st_pcr = s->st_pcr; P.st_best = s->st_best; P.st_flags = s->st_flags; P.st_merit = s->st_merit;
  P.bj = s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR;
  const int W = s->m.W;
  const int r_max = (s->L.Nc + WAVE - 1) / WAVE;
  if (r_max > kClassRHost[kNumClasses - 1]) {
    set_error("fused path supports at most %d constraint rows per world (got %d)",
              WAVE * kClassRHost[kNumClasses - 1], s->L.Nc);
    return -3;
  }
  hipLaunchKernelGGL(reset_classes_kernel, dim3(1), dim3(64), 0, stream, s->cls_counts);
  hipLaunchKernelGGL(classify_worlds_kernel, dim3((W + 255) / 256), dim3(256), 0, stream, W, s->ccount, s->L.K,
                     s->L.Nj + s->L.Nctrl, s->cls_counts, s->cls_lists);
  for (int c = 0; c < kNumClasses; ++c) {
    if (c > 0 && kClassRHost[c - 1] >= r_max) break;  // no world can need this class
    switch (kClassRHost[c]) {
#define CASE(RR) case RR: launch_fused<RR>(P, s, c, stream); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(10) CASE(12) CASE(16)
#undef CASE
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("fused kernel launch failed: %s", hipGetErrorString(e)); return -4; }
  return 0;
}

}  // namespace ofb
