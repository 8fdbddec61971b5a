// fb_plain.hip -- plain per-kernel execution path.
//
// The reference's kernel decomposition (core/base_engine.py:424-488, optim/pcr_solver.py:
// 170-295) on gfx950: one kernel per phase, one thread per (world, row) or (world, body),
// every intermediate in HBM, J^T x scattered with global atomics, per-world dot products
// by one workgroup per world.  Newton iterations are driven from the host with one flag
// read-back per iteration, like the reference's eager loop (base_engine.py:502-507); the
// PCR iterations are a fixed unroll whose kernels skip worlds that have stopped.  Both
// termination rules are supported: per-world, and the reference's batch-global flags.
//
// Body data uses the same block layout as the fused kernel's LDS (fb_rows.h), here in
// HBM: bodies[(w*B + b)*BS + field].  Row arrays are world-major, rows compacted to
// [joints | control | normals(n_c) | friction(2 n_c)].

#include <algorithm>
#include <cstring>

#include <hip/hip_runtime.h>

#include "fb_math.h"
#include "fb_rows.h"
#include "fb_solver.h"

namespace ofb {

using namespace rows;

struct PlainBuffers {
  float *bodies, *lam, *lamprev, *lamb, *J, *Cr, *Pi, *rhs, *hc, *x, *r, *z, *p, *Ap, *y, *Az;
  int32_t *b0, *b1, *act;
  float *bsq, *zAz, *zAz_old, *yAp, *rsq, *merit, *best_merit;
  int32_t *nr_it, *pcr_it, *pcr_total, *best_idx, *nr_run, *pcr_run;
  int32_t* flags;    // [0] nr keep, [1] pcr keep, [2] global nr iter, [3] global pcr iter
  int32_t* host_keep;// pinned
};

namespace {

struct PlainParams {
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
  const int32_t* topo;
  const float* lambda_seed;
  float* lambda_out;
  int32_t *st_nr, *st_pcr, *st_best, *st_flags;
  float* st_merit;
  PlainBuffers b;
  int global_term;
};

constexpr int RT = 128;  // threads per block of the row kernels
constexpr int DT_ = 256; // threads per block of the per-world reduction kernels

__device__ __forceinline__ int n_c_of(const PlainParams& P, int w) { return min(P.ccount[w], P.L.K); }
__device__ __forceinline__ int n_r_of(const PlainParams& P, int w) { return P.L.Nj + P.L.Nctrl + 3 * n_c_of(P, w); }
__device__ __forceinline__ float* body_block(const PlainParams& P, int w) { return P.b.bodies + (size_t)w * P.m.B * BS; }
__device__ __forceinline__ size_t ri(const PlainParams& P, int w, int i) { return (size_t)w * P.L.Nc + i; }
__device__ __forceinline__ sv ldJ(const PlainParams& P, int w, int i, int side) {
  const float* j = P.b.J + (size_t)w * 12 * P.L.Nc + (size_t)(side * 6) * P.L.Nc + i;
  const int N = P.L.Nc;
  return sv{mk(j[0], j[N], j[2 * N]), mk(j[3 * N], j[4 * N], j[5 * N])};
}
__device__ __forceinline__ void stJ(const PlainParams& P, int w, int i, int side, const sv& v) {
  float* j = P.b.J + (size_t)w * 12 * P.L.Nc + (size_t)(side * 6) * P.L.Nc + i;
  const int N = P.L.Nc;
  j[0] = v.a.x; j[N] = v.a.y; j[2 * N] = v.a.z; j[3 * N] = v.b.x; j[4 * N] = v.b.y; j[5 * N] = v.b.z;
}

// block sum over DT_ threads; every thread returns the total
__device__ float block_sum(float v, float* sh) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wid = threadIdx.x / 64, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  float t = 0.f;
  for (int k = 0; k < DT_ / 64; ++k) t += sh[k];
  return t;
}

#define ROW_PROLOGUE(RUNFLAG)                                  \
  const int w = blockIdx.y;                                    \
  const int i = blockIdx.x * blockDim.x + threadIdx.x;         \
  if (!P.b.RUNFLAG[w]) return;                                 \
  const int n_r = n_r_of(P, w);                                \
  if (i >= n_r) return;

// ---- step setup: load bodies, lambda := 0 or seed, flags ----
__global__ void k_load(PlainParams P) {
  const int w = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int B = P.m.B;
  float* sm = body_block(P, w);
  if (t < B) {
    float* bb = sm + t * BS;
    size_t k = (size_t)w * B + t;
    for (int i = 0; i < 7; ++i) { float v = P.body_q[k * 7 + i]; bb[BQ + i] = v; bb[BQP + i] = v; }
    for (int i = 0; i < 6; ++i) {
      float v = P.body_qd[k * 6 + i];
      bb[BU + i] = v; bb[BUP + i] = v; bb[BF + i] = P.body_f[k * 6 + i];
    }
    bb[BM] = P.m.body_mass[k];
    bb[BMINV] = P.m.body_inv_mass[k];
    for (int i = 0; i < 3; ++i) bb[BCOM + i] = P.m.body_com[k * 3 + i];
    for (int i = 0; i < 9; ++i) { bb[BIB + i] = P.m.body_inertia[k * 9 + i]; bb[BIIB + i] = P.m.body_inv_inertia[k * 9 + i]; }
  }
  const int n_c = n_c_of(P, w);
  const int n_r = P.L.Nj + P.L.Nctrl + 3 * n_c;
  for (int i = t; i < n_r; i += gridDim.x * blockDim.x) {
    float l0 = 0.f;
    if (P.cfg.seed_iterate) {
      int orig = i < P.L.Nj + P.L.Nctrl ? i : (i - P.L.Nj - P.L.Nctrl < n_c ? P.L.off_n + (i - P.L.Nj - P.L.Nctrl)
                                                                          : P.L.off_f + (i - P.L.Nj - P.L.Nctrl - n_c));
      l0 = P.lambda_seed[(size_t)w * P.L.Nc + orig];
    }
    P.b.lam[ri(P, w, i)] = l0;
    P.b.lamb[ri(P, w, i)] = l0;
  }
  if (t == 0) {
    P.b.nr_it[w] = 0; P.b.pcr_total[w] = 0; P.b.best_idx[w] = -1; P.b.best_merit[w] = 0.f;
    P.b.nr_run[w] = 1;
    if (w == 0) { P.b.flags[0] = 1; P.b.flags[2] = 0; }
  }
}

// ---- lambda_prev := lambda (base_engine.py:441) ----
__global__ void k_lamprev(PlainParams P) {
  ROW_PROLOGUE(nr_run)
  P.b.lamprev[ri(P, w, i)] = P.b.lam[ri(P, w, i)];
}

// ---- bodies: world inverse inertia + dynamics residual (linear_system.py:120-177) ----
__global__ void k_bodies_dyn(PlainParams P, int with_inv) {
  const int w = blockIdx.x * blockDim.y + threadIdx.y;
  if (w >= P.m.W || !P.b.nr_run[w]) return;
  dynamics_bodies(P, body_block(P, w), threadIdx.x, with_inv != 0);
}

// ---- rows: assembly (linear_system.py:179-322) ----
__global__ void k_rows_assemble(PlainParams P) {
  ROW_PROLOGUE(nr_run)
  float* sm = body_block(P, w);
  const float* lp = P.b.lamprev + (size_t)w * P.L.Nc;
  size_t k = ri(P, w, i);
  RowOut o = assemble_row(P, sm, lp, w, i, n_c_of(P, w), P.b.lam[k]);
  stJ(P, w, i, 0, o.J0);
  stJ(P, w, i, 1, o.J1);
  P.b.b0[k] = o.b0; P.b.b1[k] = o.b1; P.b.act[k] = o.active ? 1 : 0;
  P.b.hc[k] = o.hc;
  P.b.Cr[k] = o.C + P.cfg.regularization;
  if (!o.active && i >= P.L.Nj + P.L.Nctrl) P.b.lam[k] = 0.f;
}

// ---- bodies: M^-1 (field SRC) -> BVV ----
__global__ void k_bodies_minv(PlainParams P, int src, int runsel) {
  const int w = blockIdx.x * blockDim.y + threadIdx.y;
  if (w >= P.m.W) return;
  if (!(runsel ? P.b.pcr_run[w] : P.b.nr_run[w])) return;
  for (int b = threadIdx.x; b < P.m.B; b += blockDim.x) {
    float* bb = body_block(P, w) + b * BS;
    stsv(bb + BVV, apply_minv(bb, ldsv(bb + src)));
  }
}

// ---- rows: Schur RHS and Jacobi diagonal (linear_system.py:17-65, preconditioner.py:9-50) ----
__global__ void k_rows_rhs(PlainParams P) {
  ROW_PROLOGUE(nr_run)
  size_t k = ri(P, w, i);
  float rh = 0.f, pinv = 0.f;
  if (P.b.act[k]) {
    const float* sm = body_block(P, w);
    sv J0 = ldJ(P, w, i, 0), J1 = ldJ(P, w, i, 1);
    int a = P.b.b0[k], c = P.b.b1[k];
    float s = 0.f, d = 0.f;
    if (a >= 0) { s += sdot(J0, ldsv(sm + a * BS + BVV)); d += sdot(J0, apply_minv(sm + a * BS, J0)); }
    if (c >= 0) { s += sdot(J1, ldsv(sm + c * BS + BVV)); d += sdot(J1, apply_minv(sm + c * BS, J1)); }
    rh = (s - P.b.hc[k]) / P.dt;
    pinv = 1.f / (d + P.b.Cr[k] + 1e-6f);
  }
  P.b.rhs[k] = rh;
  P.b.Pi[k] = pinv;
  // PCR initialisation, x0 = 0 => r = b, z = P r
  P.b.x[k] = 0.f;
  P.b.r[k] = rh;
  P.b.z[k] = pinv * rh;
}

// ---- matvec pieces (system_operator.py:36-124) ----
__global__ void k_zero_acc(PlainParams P, int runsel) {
  const int w = blockIdx.x * blockDim.y + threadIdx.y;
  if (w >= P.m.W) return;
  if (!(runsel ? P.b.pcr_run[w] : P.b.nr_run[w])) return;
  for (int e = threadIdx.x; e < 6 * P.m.B; e += blockDim.x) body_block(P, w)[(e / 6) * BS + BACC + (e % 6)] = 0.f;
}

__global__ void k_scatter(PlainParams P, const float* __restrict__ v, int runsel) {
  const int w = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!(runsel ? P.b.pcr_run[w] : P.b.nr_run[w])) return;
  if (i >= n_r_of(P, w)) return;
  size_t k = ri(P, w, i);
  if (!P.b.act[k]) return;
  float* sm = body_block(P, w);
  float xv = v[k];
  int a = P.b.b0[k], c = P.b.b1[k];
  if (a >= 0) atomic_add_sv(sm + a * BS + BACC, ldJ(P, w, i, 0) * xv);
  if (c >= 0) atomic_add_sv(sm + c * BS + BACC, ldJ(P, w, i, 1) * xv);
}

__global__ void k_gather(PlainParams P, const float* __restrict__ v, float* __restrict__ out, int runsel) {
  const int w = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!(runsel ? P.b.pcr_run[w] : P.b.nr_run[w])) return;
  if (i >= n_r_of(P, w)) return;
  size_t k = ri(P, w, i);
  float o = 0.f;
  if (P.b.act[k]) {
    const float* sm = body_block(P, w);
    int a = P.b.b0[k], c = P.b.b1[k];
    if (a >= 0) o += sdot(ldJ(P, w, i, 0), ldsv(sm + a * BS + BVV));
    if (c >= 0) o += sdot(ldJ(P, w, i, 1), ldsv(sm + c * BS + BVV));
    o += P.b.Cr[k] * v[k];
  }
  out[k] = o;
}

// ---- PCR init scalars: b.b, z.Az, r.r (pcr_solver.py:206-228) ----
__global__ void k_pcr_init_dots(PlainParams P) {
  __shared__ float sh[DT_ / 64];
  const int w = blockIdx.x;
  if (!P.b.nr_run[w]) { if (threadIdx.x == 0) P.b.pcr_run[w] = 0; return; }
  const int n_r = n_r_of(P, w);
  float bb = 0.f, za = 0.f, rr = 0.f;
  for (int i = threadIdx.x; i < n_r; i += DT_) {
    size_t k = ri(P, w, i);
    P.b.p[k] = P.b.z[k];
    P.b.Ap[k] = P.b.Az[k];
    bb += P.b.rhs[k] * P.b.rhs[k];
    za += P.b.z[k] * P.b.Az[k];
    rr += P.b.r[k] * P.b.r[k];
  }
  bb = block_sum(bb, sh);
  za = block_sum(za, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    P.b.bsq[w] = bb; P.b.zAz[w] = za; P.b.rsq[w] = rr;
    P.b.pcr_run[w] = 1; P.b.pcr_it[w] = 0;
    if (w == 0) { P.b.flags[1] = 1; P.b.flags[3] = 0; }
  }
}

// ---- PCR body (pcr_solver.py:232-292) ----
__global__ void k_pcr_y(PlainParams P) {  // y = P Ap; zAz_old = zAz
  ROW_PROLOGUE(pcr_run)
  size_t k = ri(P, w, i);
  P.b.y[k] = P.b.act[k] ? P.b.Pi[k] * P.b.Ap[k] : 0.f;
}

__global__ void k_pcr_dot_yAp(PlainParams P) {
  __shared__ float sh[DT_ / 64];
  const int w = blockIdx.x;
  if (!P.b.pcr_run[w]) return;
  const int n_r = n_r_of(P, w);
  float s = 0.f;
  for (int i = threadIdx.x; i < n_r; i += DT_) s += P.b.y[ri(P, w, i)] * P.b.Ap[ri(P, w, i)];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) { P.b.yAp[w] = s; P.b.zAz_old[w] = P.b.zAz[w]; }
}

__global__ void k_pcr_xrz(PlainParams P) {
  ROW_PROLOGUE(pcr_run)
  size_t k = ri(P, w, i);
  float yAp = P.b.yAp[w];
  float a = yAp > 0.f ? P.b.zAz_old[w] / yAp : 0.f;
  P.b.x[k] += a * P.b.p[k];
  P.b.r[k] -= a * P.b.Ap[k];
  P.b.z[k] -= a * P.b.y[k];
}

__global__ void k_pcr_dot_zAz(PlainParams P) {
  __shared__ float sh[DT_ / 64];
  const int w = blockIdx.x;
  if (!P.b.pcr_run[w]) return;
  const int n_r = n_r_of(P, w);
  float s = 0.f;
  for (int i = threadIdx.x; i < n_r; i += DT_) s += P.b.z[ri(P, w, i)] * P.b.Az[ri(P, w, i)];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) P.b.zAz[w] = s;
}

__global__ void k_pcr_pAp(PlainParams P) {
  ROW_PROLOGUE(pcr_run)
  size_t k = ri(P, w, i);
  float zo = P.b.zAz_old[w];
  float bt = zo > 0.f ? P.b.zAz[w] / zo : 0.f;
  P.b.p[k] = P.b.z[k] + bt * P.b.p[k];
  P.b.Ap[k] = P.b.Az[k] + bt * P.b.Ap[k];
}

// r.r and the per-world stop test; the global flag collects unconverged worlds
__global__ void k_pcr_check(PlainParams P) {
  __shared__ float sh[DT_ / 64];
  const int w = blockIdx.x;
  if (!P.b.pcr_run[w]) return;
  const int n_r = n_r_of(P, w);
  float s = 0.f;
  for (int i = threadIdx.x; i < n_r; i += DT_) s += P.b.r[ri(P, w, i)] * P.b.r[ri(P, w, i)];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    P.b.rsq[w] = s;
    int it = ++P.b.pcr_it[w];
    P.b.pcr_total[w] += 1;
    float tol_sq = P.cfg.lin_tol >= 0.f ? P.cfg.lin_tol * P.cfg.lin_tol : -1.f;
    float atol_sq = P.cfg.lin_atol >= 0.f ? P.cfg.lin_atol * P.cfg.lin_atol : -1.f;
    float target = fmaxf(atol_sq, P.b.bsq[w] * tol_sq);
    bool unconv = s > target;
    if (unconv) atomicOr(P.b.flags + 1, 2);  // bit 1: some world unconverged this iteration
    if (!P.global_term) P.b.pcr_run[w] = (unconv && it < P.cfg.lin_max_iters) ? 1 : 0;
  }
}

// global PCR flag (pcr_solver.py:84-117): one thread
__global__ void k_pcr_global(PlainParams P) {
  int f = P.b.flags[1];
  int it = ++P.b.flags[3];
  int keep = (f & 2) ? 1 : 0;
  if (it >= P.cfg.lin_max_iters) keep = 0;
  P.b.flags[1] = keep;
}

__global__ void k_pcr_apply_global(PlainParams P) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= P.m.W || !P.b.nr_run[w]) return;
  P.b.pcr_run[w] = P.b.flags[1] & 1;
}

// ---- velocity recovery + Newton update (linear_system.py:344-377, newton_step.py:22-54) ----
__global__ void k_bodies_update(PlainParams P) {
  const int w = blockIdx.x * blockDim.y + threadIdx.y;
  if (w >= P.m.W || !P.b.nr_run[w]) return;
  for (int b = threadIdx.x; b < P.m.B; b += blockDim.x) {
    float* bb = body_block(P, w) + b * BS;
    sv acc = ldsv(bb + BACC), hd = ldsv(bb + BHD);
    sv du = apply_minv(bb, sv{acc.a * P.dt - hd.a, acc.b * P.dt - hd.b});
    sv u = ldsv(bb + BU);
    u = sv{u.a + du.a, u.b + du.b};
    stsv(bb + BU, u);
    sttf(bb + BQ, integrate(u, ldtf(bb + BQP), ld3(bb + BCOM), P.dt));
  }
}

__global__ void k_lam_update(PlainParams P) {
  ROW_PROLOGUE(nr_run)
  P.b.lam[ri(P, w, i)] += P.b.x[ri(P, w, i)];
}

// ---- residual rows (core/residual.py:18-173) ----
__global__ void k_rows_residual(PlainParams P) {
  ROW_PROLOGUE(nr_run)
  float* sm = body_block(P, w);
  const float* lp = P.b.lamprev + (size_t)w * P.L.Nc;
  size_t k = ri(P, w, i);
  RowOut o = assemble_row(P, sm, lp, w, i, n_c_of(P, w), P.b.lam[k]);
  P.b.hc[k] = o.hc;
}

// merit, candidate bookkeeping and the Newton stop rule (base_engine.py:42-80,
// backtracking.py:12-95); one workgroup per world
__global__ void k_nr_finish(PlainParams P) {
  __shared__ float sh[DT_ / 64];
  __shared__ int take_s;
  const int w = blockIdx.x;
  if (!P.b.nr_run[w]) return;
  const int n_r = n_r_of(P, w);
  const int B = P.m.B;
  float* sm = body_block(P, w);
  float s = 0.f;
  for (int i = threadIdx.x; i < n_r; i += DT_) { float v = P.b.hc[ri(P, w, i)]; s += v * v; }
  for (int e = threadIdx.x; e < 6 * B; e += DT_) { float v = sm[(e / 6) * BS + BHD + (e % 6)]; s += v * v; }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    P.b.merit[w] = s;
    int kidx = P.b.nr_it[w];
    P.b.nr_it[w] = kidx + 1;
    int take = (kidx == P.cfg.nr_backtrack_min_iter) ||
               (kidx > P.cfg.nr_backtrack_min_iter && s < P.b.best_merit[w]);
    if (take) { P.b.best_idx[w] = kidx; P.b.best_merit[w] = s; }
    take_s = take;
    float atol_sq = P.cfg.nr_atol * P.cfg.nr_atol;
    if (P.global_term) {
      if (s > atol_sq) atomicOr(P.b.flags + 0, 2);  // some world unconverged
    } else {
      int n = kidx + 1;
      bool keep = s > atol_sq;
      if (n < P.cfg.nr_backtrack_min_iter) keep = true;
      if (n >= P.cfg.nr_max_iters) keep = false;
      P.b.nr_run[w] = keep ? 1 : 0;
      if (keep) atomicOr(P.b.flags + 0, 4);  // some world keeps iterating
    }
  }
  __syncthreads();
  if (take_s) {
    for (int e = threadIdx.x; e < 13 * B; e += DT_) {
      int b = e / 13, c = e % 13;
      float* bb = sm + b * BS;
      if (c < 7) bb[BBQ + c] = bb[BQ + c];
      else bb[BBU + c - 7] = bb[BU + c - 7];
    }
    for (int i = threadIdx.x; i < n_r; i += DT_) P.b.lamb[ri(P, w, i)] = P.b.lam[ri(P, w, i)];
  }
}

// Newton keep flag for the host (and the global rule, base_engine.py:42-80); one thread
__global__ void k_nr_global(PlainParams P, int32_t* host_keep) {
  int f = P.b.flags[0];
  int it = ++P.b.flags[2];
  int keep;
  if (P.global_term) {
    keep = (f & 2) ? 1 : 0;
    if (it < P.cfg.nr_backtrack_min_iter) keep = 1;
    if (it >= P.cfg.nr_max_iters) keep = 0;
  } else {
    keep = (f & 4) ? 1 : 0;
  }
  P.b.flags[0] = keep;
  *host_keep = keep;
}

__global__ void k_nr_apply_global(PlainParams P) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= P.m.W) return;
  P.b.nr_run[w] = P.b.flags[0] & 1;
}

__global__ void k_nr_reset_flags(PlainParams P) { P.b.flags[0] = 0; }

// ---- pick the candidate and write the outputs ----
__global__ void k_finish(PlainParams P) {
  const int w = blockIdx.x;
  const int B = P.m.B;
  float* sm = body_block(P, w);
  const int n_it = P.b.nr_it[w];
  const bool restore = n_it > P.cfg.nr_backtrack_min_iter && P.b.best_idx[w] >= 0;
  const int n_c = n_c_of(P, w);
  const int n_r = n_r_of(P, w);
  for (int e = threadIdx.x; e < 13 * B; e += blockDim.x) {
    int b = e / 13, c = e % 13;
    const float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
    if (c < 7) P.out_q[k * 7 + c] = restore ? bb[BBQ + c] : bb[BQ + c];
    else P.out_qd[k * 6 + c - 7] = restore ? bb[BBU + c - 7] : bb[BU + c - 7];
  }
  const RowLayout& L = P.L;
  for (int i = threadIdx.x; i < n_r; i += blockDim.x) {
    int orig = i < L.Nj + L.Nctrl ? i : (i - L.Nj - L.Nctrl < n_c ? L.off_n + (i - L.Nj - L.Nctrl)
                                                                 : L.off_f + (i - L.Nj - L.Nctrl - n_c));
    P.lambda_out[(size_t)w * L.Nc + orig] = restore ? P.b.lamb[ri(P, w, i)] : P.b.lam[ri(P, w, i)];
  }
  for (int c = n_c + threadIdx.x; c < L.K; c += blockDim.x) {
    P.lambda_out[(size_t)w * L.Nc + L.off_n + c] = 0.f;
    P.lambda_out[(size_t)w * L.Nc + L.off_f + 2 * c] = 0.f;
    P.lambda_out[(size_t)w * L.Nc + L.off_f + 2 * c + 1] = 0.f;
  }
  if (threadIdx.x == 0) {
    P.st_nr[w] = n_it;
    P.st_pcr[w] = P.b.pcr_total[w];
    P.st_merit[w] = restore ? P.b.best_merit[w] : P.b.merit[w];
    P.st_best[w] = restore ? P.b.best_idx[w] : n_it - 1;
    int fl = P.st_flags[w] & FLAG_CONTACT_OVERFLOW;
    if (n_it >= P.cfg.nr_max_iters && P.b.merit[w] > P.cfg.nr_atol * P.cfg.nr_atol) fl |= FLAG_NR_CAP;
    P.st_flags[w] = fl;
  }
}

template <class T>
int palloc(T** p, size_t n) {
  if (hipMalloc((void**)p, (n ? n : 1) * sizeof(T)) != hipSuccess) { set_error("hipMalloc failed (plain path)"); return -4; }
  if (hipMemset(*p, 0, (n ? n : 1) * sizeof(T)) != hipSuccess) { set_error("hipMemset failed"); return -4; }
  return 0;
}

}  // namespace

int plain_create(ofb_solver* s) {
  s->plain = new PlainBuffers;
  PlainBuffers& b = *s->plain;
  std::memset(&b, 0, sizeof(b));
  const size_t W = s->m.W, N = (size_t)W * s->L.Nc;
  int rc = 0;
#define PA(f, n) if ((rc = palloc(&b.f, n)) != 0) return rc;
  PA(bodies, W * s->m.B * BS);
  PA(lam, N); PA(lamprev, N); PA(lamb, N); PA(J, N * 12); PA(Cr, N); PA(Pi, N); PA(rhs, N); PA(hc, N);
  PA(x, N); PA(r, N); PA(z, N); PA(p, N); PA(Ap, N); PA(y, N); PA(Az, N);
  PA(b0, N); PA(b1, N); PA(act, N);
  PA(bsq, W); PA(zAz, W); PA(zAz_old, W); PA(yAp, W); PA(rsq, W); PA(merit, W); PA(best_merit, W);
  PA(nr_it, W); PA(pcr_it, W); PA(pcr_total, W); PA(best_idx, W); PA(nr_run, W); PA(pcr_run, W);
  PA(flags, 8);
#undef PA
  if (hipHostMalloc((void**)&b.host_keep, sizeof(int32_t)) != hipSuccess) { set_error("hipHostMalloc failed"); return -4; }
  return 0;
}

void plain_destroy(ofb_solver* s) {
  if (!s->plain) return;
  PlainBuffers& b = *s->plain;
  void* ptrs[] = {b.bodies, b.lam, b.lamprev, b.lamb, b.J, b.Cr, b.Pi, b.rhs, b.hc, b.x, b.r, b.z, b.p, b.Ap, b.y,
                  b.Az, b.b0, b.b1, b.act, b.bsq, b.zAz, b.zAz_old, b.yAp, b.rsq, b.merit, b.best_merit, b.nr_it,
                  b.pcr_it, b.pcr_total, b.best_idx, b.nr_run, b.pcr_run, b.flags};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b.host_keep) (void)hipHostFree(b.host_keep);
  delete s->plain;
  s->plain = nullptr;
}

int plain_step(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl, float dt,
               hipStream_t st) {
  PlainParams P;
  P.m = s->m; P.L = s->L; P.cfg = s->cfg; P.dt = dt;
  P.body_q = in->body_q; P.body_qd = in->body_qd; P.body_f = in->body_f;
  P.target_pos = ctl->joint_target_pos; P.target_vel = ctl->joint_target_vel;
  P.out_q = out->body_q; P.out_qd = out->body_qd;
  P.ccount = s->ccount; P.cf = s->cf; P.ci = s->ci; P.topo = s->topo;
  P.lambda_seed = s->lambda_seed; P.lambda_out = s->lambda;
  P.st_nr = s->st_nr; P.st_pcr = s->st_pcr; P.st_best = s->st_best; P.st_flags = s->st_flags; P.st_merit = s->st_merit;
  P.b = *s->plain;
  P.global_term = s->cfg.termination == OFB_TERM_GLOBAL;
  const int W = s->m.W;
  const dim3 rowgrid((s->L.Nc + RT - 1) / RT, W), rowblk(RT);
  const int wpb = 4;  // worlds per block for per-body kernels
  const dim3 bgrid((W + wpb - 1) / wpb), bblk(64, wpb);
  const dim3 wgrid(W), wblk(DT_);
  const dim3 fgrid((W + 255) / 256), fblk(256);
  hipLaunchKernelGGL(k_load, dim3(std::max(1, (s->L.Nc + RT - 1) / RT), W), rowblk, 0, st, P);
  hipLaunchKernelGGL(k_nr_reset_flags, dim3(1), dim3(1), 0, st, P);
  for (int it = 0; it < s->cfg.nr_max_iters; ++it) {
    hipLaunchKernelGGL(k_lamprev, rowgrid, rowblk, 0, st, P);
    hipLaunchKernelGGL(k_bodies_dyn, bgrid, bblk, 0, st, P, 1);
    hipLaunchKernelGGL(k_rows_assemble, rowgrid, rowblk, 0, st, P);
    hipLaunchKernelGGL(k_bodies_minv, bgrid, bblk, 0, st, P, (int)BHD, 0);
    hipLaunchKernelGGL(k_rows_rhs, rowgrid, rowblk, 0, st, P);
    // Az = A z for the initial z
    hipLaunchKernelGGL(k_zero_acc, bgrid, bblk, 0, st, P, 0);
    hipLaunchKernelGGL(k_scatter, rowgrid, rowblk, 0, st, P, (const float*)P.b.z, 0);
    hipLaunchKernelGGL(k_bodies_minv, bgrid, bblk, 0, st, P, (int)BACC, 0);
    hipLaunchKernelGGL(k_gather, rowgrid, rowblk, 0, st, P, (const float*)P.b.z, P.b.Az, 0);
    hipLaunchKernelGGL(k_pcr_init_dots, wgrid, wblk, 0, st, P);  // p = z, Ap = Az, scalars, pcr_run
    for (int j = 0; j < s->cfg.lin_max_iters; ++j) {
      hipLaunchKernelGGL(k_pcr_y, rowgrid, rowblk, 0, st, P);
      hipLaunchKernelGGL(k_pcr_dot_yAp, wgrid, wblk, 0, st, P);
      hipLaunchKernelGGL(k_pcr_xrz, rowgrid, rowblk, 0, st, P);
      hipLaunchKernelGGL(k_zero_acc, bgrid, bblk, 0, st, P, 1);
      hipLaunchKernelGGL(k_scatter, rowgrid, rowblk, 0, st, P, (const float*)P.b.z, 1);
      hipLaunchKernelGGL(k_bodies_minv, bgrid, bblk, 0, st, P, (int)BACC, 1);
      hipLaunchKernelGGL(k_gather, rowgrid, rowblk, 0, st, P, (const float*)P.b.z, P.b.Az, 1);
      hipLaunchKernelGGL(k_pcr_dot_zAz, wgrid, wblk, 0, st, P);
      hipLaunchKernelGGL(k_pcr_pAp, rowgrid, rowblk, 0, st, P);
      hipLaunchKernelGGL(k_pcr_check, wgrid, wblk, 0, st, P);
      if (P.global_term) {
        hipLaunchKernelGGL(k_pcr_global, dim3(1), dim3(1), 0, st, P);
        hipLaunchKernelGGL(k_pcr_apply_global, fgrid, fblk, 0, st, P);
      }
    }
    hipLaunchKernelGGL(k_zero_acc, bgrid, bblk, 0, st, P, 0);
    hipLaunchKernelGGL(k_scatter, rowgrid, rowblk, 0, st, P, (const float*)P.b.x, 0);
    hipLaunchKernelGGL(k_bodies_update, bgrid, bblk, 0, st, P);
    hipLaunchKernelGGL(k_lam_update, rowgrid, rowblk, 0, st, P);
    hipLaunchKernelGGL(k_bodies_dyn, bgrid, bblk, 0, st, P, 0);
    hipLaunchKernelGGL(k_rows_residual, rowgrid, rowblk, 0, st, P);
    hipLaunchKernelGGL(k_nr_finish, wgrid, wblk, 0, st, P);
    hipLaunchKernelGGL(k_nr_global, dim3(1), dim3(1), 0, st, P, P.b.host_keep);
    if (P.global_term) hipLaunchKernelGGL(k_nr_apply_global, fgrid, fblk, 0, st, P);
    hipLaunchKernelGGL(k_nr_reset_flags, dim3(1), dim3(1), 0, st, P);
    if (hipGetLastError() != hipSuccess) { set_error("plain path launch failed"); return -4; }
    // host read-back of the Newton keep flag, as the reference's eager loop does
    if (hipStreamSynchronize(st) != hipSuccess) { set_error("plain path stream sync failed"); return -4; }
    if (*P.b.host_keep == 0) break;
  }
  hipLaunchKernelGGL(k_finish, wgrid, dim3(256), 0, st, P);
  if (hipGetLastError() != hipSuccess) { set_error("plain path launch failed"); return -4; }
  return 0;
}

}  // namespace ofb
