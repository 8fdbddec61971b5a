// fb_fused_kernel.h -- fused persistent Fisher-Burmeister Newton solver for gfx950.
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

#pragma once

#include <hip/hip_runtime.h>

#include "fb_math.h"
#include "fb_rows.h"
#include "fb_solver.h"

namespace ofb {

using namespace rows;

#ifdef OFB_PHASE_TIMING
// Profiling build only: per-phase cycle counts of the Newton loop, summed over worlds,
// per instance (row 0: paired, 1: dense, R: matrix-free with R slots per lane, 13-15:
// worlds of 2, 4, 8 waves with two slots per lane, 9-11: with four, 8: with three);
// columns 8 and 9 count Newton and PCR iterations.
__device__ unsigned long long g_phase_cycles[17][10];
#define OFB_PHASE_INIT unsigned long long ph_[8] = {0, 0, 0, 0, 0, 0, 0, 0}; unsigned long long ph_t = clock64();
#define OFB_PHASE(k) { unsigned long long t_ = clock64(); ph_[k] += t_ - ph_t; ph_t = t_; }
#define OFB_PHASE_FLUSH                                                                       \
  if (lane == 0) {                                                                            \
    const int ci_ = DENSE ? (GW == 32 ? 0 : 1) : GW > WAVE ? (R == 3 ? 8 : (GW == 128 ? 13 : GW == 256 ? 14 : 15) - (R - 2) * 2) : R; \
    for (int k_ = 0; k_ < 8; ++k_) atomicAdd(&g_phase_cycles[ci_][k_], ph_[k_]);              \
    atomicAdd(&g_phase_cycles[ci_][8], (unsigned long long)n_it);                             \
    atomicAdd(&g_phase_cycles[ci_][9], (unsigned long long)*pcr_count);                       \
  }
#else
#define OFB_PHASE_INIT
#define OFB_PHASE(k)
#define OFB_PHASE_FLUSH
#endif

// LDS layout of one world (floats; regions that are read as quads start on 16-byte
// boundaries): bodies, lambda_prev (rows_cap + 1, compacted row index), one static record
// per joint, one per contact slot of the class (cs_cap), the block-Jacobi region, the
// dense Schur operator or the lane-private body sums, and for a world of several waves
// the per-wave body-sum partials and the cross-wave reduction slots.  `words` is a
// multiple of 4, so every world's base stays aligned.
struct FusedLayout {
  int lamprev, js, cs, bj, dense, rs, part, red, hist, words;
};
// Per-Newton-iteration history of a world (||h||^2 and the running PCR count of the first
// kHistIters iterations) kept in LDS and written to the caller's buffers once per step.
constexpr int kHistIters = 32;
constexpr int kHistWords = 2 * kHistIters;
__host__ __device__ inline int up4(int x) { return (x + 3) & ~3; }

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
  const float* lambda_seed2;  // learned seed (null: none); the lower initial residual wins
  const float* ls_steps;      // merit line-search candidates (ls_count 0: full Newton step)
  int ls_count;
  float* lambda_out;
  float* lambda_prev_out;  // lambda_prev_iter as the step leaves it (input of the adjoint); null: not kept
  int32_t *st_nr, *st_pcr, *st_best, *st_flags;
  float* st_merit;
  // per-Newton-iteration history [hist_cap][W] (null: off): ||h||^2 after the iteration
  // (candidates_res_norm_sq) and its PCR iterations (pcr_history_iter_count)
  float* hist_merit;
  int32_t* hist_pcr;
  int hist_cap;
  FusedLayout lay;        // LDS regions of one world (floats)
  const int32_t* wlist;   // worlds of this launch's row class
  const int32_t* wcount;  // number of them
  int rows_cap;           // G R: row capacity of this launch's class
};

// ---------------------------------------------------------------------------------
// Sums over the lanes of one world, which every lane of the world holds bit for bit (the
// stop rules and the PCR scalars are evaluated by each lane).  A world of one wave or half
// of one reduces inside the wave (group_sum / group_sum2).  A world of NW waves reduces
// each wave, lane 0 of each wave leaves its total in LDS, and after the workgroup barrier
// every lane adds the NW totals in wave order.  The totals alternate between two LDS
// slots: slot s is read right after the barrier of reduction n, and written again only by
// reduction n + 2, which no wave reaches before every wave has passed the barrier of
// reduction n + 1, i.e. finished its reads of reduction n.  All lanes of the world must
// call every reduction (world-uniform control flow).
// ---------------------------------------------------------------------------------
template <int GW>
struct WorldSum {
  static constexpr int NW = GW > WAVE ? GW / WAVE : 1;
  float* red;  // 2 slots x NW waves x 2 values (worlds of several waves)
  int slot;

  __device__ __forceinline__ float sum(float v) {
    if constexpr (GW <= WAVE) {
      return group_sum<GW>(v);
    } else {
      const float s = wave_sum(v);
      float* r = red + slot * 2 * NW;
      if ((threadIdx.x & (WAVE - 1)) == 0) r[threadIdx.x / WAVE] = s;
      __syncthreads();
      float t = r[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) t = t + r[w];
      slot ^= 1;
      return t;
    }
  }
  __device__ __forceinline__ void sum2(float& a, float& b) {
    if constexpr (GW <= WAVE) {
      group_sum2<GW>(a, b);
    } else {
      group_sum2<WAVE>(a, b);
      float* r = red + slot * 2 * NW;
      if ((threadIdx.x & (WAVE - 1)) == 0) {
        r[threadIdx.x / WAVE] = a;
        r[NW + threadIdx.x / WAVE] = b;
      }
      __syncthreads();
      float ta = r[0], tb = r[NW];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        ta = ta + r[w];
        tb = tb + r[NW + w];
      }
      a = ta;
      b = tb;
      slot ^= 1;
    }
  }
};

// ---------------------------------------------------------------------------------
// Per-body-pair block-Jacobi inside one wavefront (optim/per_body_pair_preconditioner.py:
// 62-410).  Rows that couple the same two bodies (static side = body B) form a pair; a
// row's rank is its position among the pair's rows in row order, and rows ranked 64 or
// more fall back to the Jacobi diagonal (the reference's fixed 64-member blocks, :59).
//
// Work is spread over the lanes by member slot (pair offset + rank; a pair's members
// hold consecutive slots), so a 64-member block of a large world occupies one slot of
// every lane instead of a few lanes in each of its row slots:
//  * structure (once per step: row activity and bodies do not change inside a step):
//    pair ids, ranks, sizes, the pair's slot range and packed-block offset, kept per
//    member slot in LDS;
//  * factor (once per Newton iteration): member s builds row s of its pair's packed lower
//    triangle of A_pair = sum_b J_a,b M_b^-1 J_c,b^T (+ C + reg on the diagonal) from
//    the staged member records; the wave factors it in place with a right-looking
//    Cholesky, pivot by pivot (the reference's factor, same failure test pivot <= 0,
//    :296-321; a failed pair uses the Jacobi diagonal, :374-380), then inverts L in
//    place, one column per member;
//  * apply (twice per PCR iteration): y = L^-1 v and z = L^-T y as two dense triangular
//    matvecs of at most 64 LDS reads per member with no serial dependence, where the
//    reference runs two triangular solves (:384-410).  An explicit A_pair^-1 would be
//    cheaper still, but in FP32 it loses an order of magnitude more accuracy on the
//    near-singular blocks of coplanar contacts (friction rows with C ~ 1e-6).
// LDS per wave after the bodies and lambda_prev (cap = 64 R rows):
//   pid[cap] i32, lsz[cap] i32 (structure), mrank/mm/mpoff/mfail[cap] i32 (per member
//   slot: rank, block size, packed offset, failed pivot), colbuf[cap], vbuf[cap],
//   rowd[16 cap] (J0 6, J1 6, b0 bits, b1 bits, C + reg, pad) per member slot,
//   X[R * 2080] packed blocks.
// ---------------------------------------------------------------------------------
constexpr int kPairCap = 64;
constexpr int kPairL = kPairCap * (kPairCap + 1) / 2;

// packed blocks: at most cap / 64 full 64-row pairs (cap a multiple of 64), or one pair
// holding every row (cap < 64)
__host__ __device__ inline int bj_words(int cap) {
  return 24 * cap + (cap >= kPairCap ? (cap / kPairCap) * kPairL : cap * (cap + 1) / 2);
}

struct BJView {
  int* pid;
  int* lsz;
  int* mrank;
  int* mm;
  int* mpoff;
  int* mfail;
  float* colbuf;
  float* vbuf;
  float* rowd;
  float* X;
};

__device__ __forceinline__ BJView bj_view(float* base, int cap) {
  BJView v;
  v.pid = reinterpret_cast<int*>(base);
  v.lsz = v.pid + cap;
  v.mrank = v.lsz + cap;
  v.mm = v.mrank + cap;
  v.mpoff = v.mm + cap;
  v.mfail = v.mpoff + cap;
  v.colbuf = reinterpret_cast<float*>(v.mfail + cap);
  v.vbuf = v.colbuf + cap;
  v.rowd = v.vbuf + cap;
  v.X = v.rowd + 16 * cap;
  return v;
}

// Per-row-slot block-Jacobi bookkeeping kept in registers.
template <int R>
struct BJRegs {
  int slot[R];  // member slot (pair offset + rank), -1 if the row is in no block
  int rank[R];  // rank of the row inside its pair
  int m[R];     // rows in the pair's block (<= 64)
  int poff[R];  // offset of the pair's packed block in X
  int fail[R];  // the pair's block has a pivot <= 0 this Newton iteration
  int mmax;     // largest block of the wave
  int nmem;     // member slots in use (every block's rows)
};

// pair ids (_compute_pair_ids_kernel :62-109), member ranks and sizes (:112-144)
template <int R, int GW = WAVE>
__device__ __forceinline__ void bj_structure(const BJView& V, int lane, int n_r, int B, const bool (&act)[R],
                                             const int (&b0)[R], const int (&b1)[R], BJRegs<R>& G, int& flags) {
  int mypid[R], leader[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + GW * k;
    mypid[k] = -1;
    if (i < n_r) {
      if (act[k]) {
        int e1 = b0[k] < 0 ? B : b0[k], e2 = b1[k] < 0 ? B : b1[k];
        mypid[k] = min(e1, e2) * (B + 1) + max(e1, e2);
      }
      V.pid[i] = mypid[k];
    }
  }
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + GW * k;
    int rank = 0, size = 0, lead = -1;
    if (i < n_r && mypid[k] >= 0) {
      for (int j = 0; j < n_r; ++j) {
        if (V.pid[j] == mypid[k]) {
          if (lead < 0) lead = j;
          size += 1;
          if (j < i) rank += 1;
        }
      }
    }
    G.rank[k] = rank;
    G.m[k] = min(size, kPairCap);
    leader[k] = lead;
    if (i < n_r) V.lsz[i] = (lead == i) ? min(size, kPairCap) : 0;
    if (lead == i && size > kPairCap) flags |= FLAG_PAIR_OVERFLOW;
  }
  wsync();
  int mmx = 0, nm = 0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int mo = 0, po = 0;
    for (int j = 0; j < leader[k]; ++j) {
      int s = V.lsz[j];
      mo += s;
      po += s * (s + 1) / 2;
    }
    G.slot[k] = (leader[k] >= 0 && G.rank[k] < kPairCap) ? mo + G.rank[k] : -1;
    G.poff[k] = po;
    if (G.slot[k] >= 0) {
      mmx = max(mmx, G.m[k]);
      nm = max(nm, G.slot[k] + 1);
      V.mrank[G.slot[k]] = G.rank[k];
      V.mm[G.slot[k]] = G.m[k];
      V.mpoff[G.slot[k]] = po;
    }
  }
  G.mmax = group_max_int<GW>(mmx);
  G.nmem = group_max_int<GW>(nm);
  wsync();
}

// dense blocks (_extract_pair_blocks_kernel :187-264) and their inverses; the member
// loops run over member slots s = lane + 64 j (j < R: a world has at most 64 R members)
template <int R>
__device__ __forceinline__ void bj_factor(const BJView& V, const float* sm, int lane, const int (&b0)[R],
                                          const int (&b1)[R], const sv (&J0)[R], const sv (&J1)[R],
                                          const float (&Cr)[R], BJRegs<R>& G, int& flags) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (G.slot[k] >= 0) {
      float* d = V.rowd + 16 * G.slot[k];
      d[0] = J0[k].a.x; d[1] = J0[k].a.y; d[2] = J0[k].a.z; d[3] = J0[k].b.x; d[4] = J0[k].b.y; d[5] = J0[k].b.z;
      d[6] = J1[k].a.x; d[7] = J1[k].a.y; d[8] = J1[k].a.z; d[9] = J1[k].b.x; d[10] = J1[k].b.y; d[11] = J1[k].b.z;
      d[12] = __int_as_float(b0[k]); d[13] = __int_as_float(b1[k]); d[14] = Cr[k];
    }
  }
  wsync();
  // member registers: rank, block size, packed offset, failed pivot of member slot
  // lane + 64 j (m = 0: no member)
  int mr[R], msz[R], mpo[R], mfl[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = lane + WAVE * j;
    mr[j] = 0; msz[j] = 0; mpo[j] = 0; mfl[j] = 0;
    if (s < G.nmem) { mr[j] = V.mrank[s]; msz[j] = V.mm[s]; mpo[j] = V.mpoff[s]; }
  }
  // block row of member s: entries c <= rank from the members' staged records
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (msz[j] == 0) continue;
    const int s = lane + WAVE * j, rank = mr[j], base = s - rank;
    const float4* ds = reinterpret_cast<const float4*>(V.rowd + 16 * s);
    const float4 e0 = ds[0], e1 = ds[1], e2 = ds[2], e3 = ds[3];
    const sv Ja{mk(e0.x, e0.y, e0.z), mk(e0.w, e1.x, e1.y)};
    const sv Jb{mk(e1.z, e1.w, e2.x), mk(e2.y, e2.z, e2.w)};
    const int ba = __float_as_int(e3.x), bb = __float_as_int(e3.y);
    const float crr = e3.z;
    sv m0 = svzero(), m1 = svzero();
    if (ba >= 0) m0 = apply_minv(sm + ba * BS, Ja);
    if (bb >= 0) m1 = apply_minv(sm + bb * BS, Jb);
    float* Xr = V.X + mpo[j] + rank * (rank + 1) / 2;
    // columns in batches of four: the batch's member records are read before any entry
    // is stored, so the LDS reads of a batch overlap
    for (int c0 = 0; c0 <= rank; c0 += 4) {
      float4 q[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4* dc = reinterpret_cast<const float4*>(V.rowd + 16 * (base + min(c0 + u, rank)));
        q[u][0] = dc[0]; q[u][1] = dc[1]; q[u][2] = dc[2]; q[u][3] = dc[3];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = c0 + u;
        if (c > rank) break;
        sv Jc0{mk(q[u][0].x, q[u][0].y, q[u][0].z), mk(q[u][0].w, q[u][1].x, q[u][1].y)};
        sv Jc1{mk(q[u][1].z, q[u][1].w, q[u][2].x), mk(q[u][2].y, q[u][2].z, q[u][2].w)};
        int cb0 = __float_as_int(q[u][3].x), cb1 = __float_as_int(q[u][3].y);
        float val = 0.f;
        if (ba >= 0) { if (cb0 == ba) val += sdot(Jc0, m0); else if (cb1 == ba) val += sdot(Jc1, m0); }
        if (bb >= 0) { if (cb0 == bb) val += sdot(Jc0, m1); else if (cb1 == bb) val += sdot(Jc1, m1); }
        if (c == rank) val += crr;
        Xr[c] = val;
      }
    }
  }
  wsync();
  // right-looking Cholesky in place, pivot by pivot (the reference's column Cholesky,
  // :296-321): column k is snapshotted, L_kk = sqrt(X_kk), L_ak = X_ak / L_kk and the
  // trailing rows update X_ac -= L_ak L_ck.  Every member reads the same pivot, so the
  // pair agrees on a failure.
  for (int piv = 0; piv < G.mmax; ++piv) {
    float colv[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      colv[j] = 0.f;
      if (piv < msz[j] && !mfl[j] && mr[j] >= piv) {
        const int s = lane + WAVE * j, rank = mr[j];
        colv[j] = V.X[mpo[j] + rank * (rank + 1) / 2 + piv];
        V.colbuf[s] = colv[j];
      }
    }
    wsync();
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (piv >= msz[j] || mfl[j]) continue;
      const int s = lane + WAVE * j, rank = mr[j], base = s - rank;
      const float dpiv = V.colbuf[base + piv];
      if (dpiv <= 0.f) {
        mfl[j] = 1;
        flags |= FLAG_FACTOR_FAILURE;
        continue;
      }
      if (rank < piv) continue;
      const float lkk = sqrtf(dpiv), inv = 1.f / lkk;
      float* Xr = V.X + mpo[j] + rank * (rank + 1) / 2;
      if (rank == piv) {
        Xr[piv] = lkk;
      } else {
        const float lak = colv[j] * inv;
        Xr[piv] = lak;
        // trailing entries in batches of eight: loads first, then the stores
        for (int c0 = piv + 1; c0 <= rank; c0 += 8) {
          float xc[8], cb[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int c = min(c0 + u, rank);
            xc[u] = Xr[c];
            cb[u] = V.colbuf[base + c];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (c0 + u <= rank) Xr[c0 + u] = xc[u] - lak * (cb[u] * inv);
        }
      }
    }
    wsync();
  }
  // X <- L^-1 in place, one column per member (member of rank c owns column c): at step
  // a row a of L is copied out, then Z_ac = -(sum_{q=c}^{a-1} L_aq Z_qc) / L_aa
  for (int a = 0; a < G.mmax; ++a) {
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (a < msz[j] && !mfl[j] && mr[j] <= a)
        V.colbuf[lane + WAVE * j] = V.X[mpo[j] + a * (a + 1) / 2 + mr[j]];
    wsync();
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (a >= msz[j] || mfl[j] || mr[j] > a) continue;
      const int s = lane + WAVE * j, c = mr[j], base = s - c;
      const float* Xp = V.X + mpo[j];
      const float laa = V.colbuf[base + a];
      float z;
      if (c == a) {
        z = 1.f / laa;
      } else {
        // the sum in q order; its operands are read eight at a time
        float acc = 0.f;
        for (int q0 = c; q0 < a; q0 += 8) {
          float lb[8], xz[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int q = min(q0 + u, a - 1);
            lb[u] = V.colbuf[base + q];
            xz[u] = Xp[q * (q + 1) / 2 + c];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (q0 + u < a) acc += lb[u] * xz[u];
        }
        z = -acc / laa;
      }
      V.X[mpo[j] + a * (a + 1) / 2 + c] = z;
    }
    wsync();
  }
  // failed pairs for the rows (apply's Jacobi fallback)
#pragma unroll
  for (int j = 0; j < R; ++j)
    if (msz[j] > 0) V.mfail[lane + WAVE * j] = mfl[j];
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) G.fail[k] = G.slot[k] >= 0 ? V.mfail[G.slot[k]] : 0;
}

// out = P v = L^-T (L^-1 v) over each pair (apply of per_body_pair_preconditioner.py:
// 324-410, whose two triangular solves become two dense triangular matvecs with the
// stored L^-1, computed per member slot); failed pairs and rows outside any block use
// the Jacobi diagonal
template <int R>
__device__ __forceinline__ void bj_apply(const BJView& V, int lane, const bool (&act)[R], const float (&Pi)[R],
                                         const BJRegs<R>& G, const float (&v)[R], float (&out)[R]) {
#pragma unroll
  for (int k = 0; k < R; ++k)
    if (G.slot[k] >= 0 && !G.fail[k]) V.vbuf[G.slot[k]] = v[k];
  wsync();
  // y = L^-1 v: row `rank` of each member's packed triangle
  float yv[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = lane + WAVE * j;
    yv[j] = 0.f;
    if (s >= G.nmem || V.mfail[s]) continue;
    const int rank = V.mrank[s], base = s - rank;
    const float* Zr = V.X + V.mpoff[s] + rank * (rank + 1) / 2;
    float y = 0.f;
    for (int c0 = 0; c0 <= rank; c0 += 8) {  // in c order, operands read eight at a time
      float zr[8], vb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = min(c0 + u, rank);
        zr[u] = Zr[c];
        vb[u] = V.vbuf[base + c];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (c0 + u <= rank) y += zr[u] * vb[u];
    }
    yv[j] = y;
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = lane + WAVE * j;
    if (s < G.nmem) V.colbuf[s] = yv[j];
  }
  wsync();
  // z = L^-T y: column `rank` of each member's packed triangle, into vbuf
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = lane + WAVE * j;
    if (s >= G.nmem || V.mfail[s]) continue;
    const int rank = V.mrank[s], base = s - rank, m = V.mm[s];
    const float* Xp = V.X + V.mpoff[s];
    float z = 0.f;
    for (int c0 = rank; c0 < m; c0 += 8) {  // in c order, operands read eight at a time
      float xp[8], cb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = min(c0 + u, m - 1);
        xp[u] = Xp[c * (c + 1) / 2 + rank];
        cb[u] = V.colbuf[base + c];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (c0 + u < m) z += xp[u] * cb[u];
    }
    yv[j] = z;
  }
  wsync();
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = lane + WAVE * j;
    if (s < G.nmem) V.vbuf[s] = yv[j];
  }
  wsync();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float o = 0.f;
    if (act[k]) o = (G.slot[k] >= 0 && !G.fail[k]) ? V.vbuf[G.slot[k]] : Pi[k] * v[k];
    out[k] = o;
  }
  wsync();
}

// out = (J M^-1 J^T + C + reg) v over the rows of one world (optim/system_operator.py:36-124):
// sum J^T v per body into the LDS accumulator, apply M^-1 per body, gather.
template <int R, int GW, class BS_>
__device__ __forceinline__ void schur_matvec(float* sm, int lane, int B, int n_r, const BS_& S, const bool (&act)[R],
                                             const int (&b0)[R], const int (&b1)[R], const sv (&J0)[R],
                                             const sv (&J1)[R], const float (&Cr)[R], const float (&v)[R],
                                             float (&out)[R]) {
  S.begin(BACC, true);
#pragma unroll
  for (int k = 0; k < R; ++k) S.add(BACC, lane + GW * k, act[k], b0[k], b1[k], J0[k] * v[k], J1[k] * v[k]);
  S.end(BACC, true, n_r);
  // A world of several waves with lane-private sums: the body totals were written by the
  // first wave's lanes (6B <= 60 values) after the workgroup barrier inside S.end, and
  // the same wave's lanes apply M^-1 below, so a wave barrier suffices here.
  if (GW > WAVE && S.rs) wsync();
  else gsync<GW>();
  for (int b = lane; b < B; b += GW) {
    float* bb = sm + b * BS;
    stsv(bb + BVV, apply_minv(bb, ldsv(bb + BACC)));
  }
  gsync<GW>();
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
  // No trailing barrier for a world of several waves: the next writes of BACC and BVV
  // come after the workgroup barrier inside the next S.end, which no wave passes before
  // every wave has finished the gather above (and after the workgroup barriers of the
  // PCR dot products in between); the lane-private copies are only read by their own wave.
  if constexpr (GW <= WAVE) gsync<GW>();
}

// ---------------------------------------------------------------------------------
// Dense Schur operator for worlds of at most LD rows (LD = 32: two worlds per wave; 64:
// one world per wave; one row per lane, R = 1).  Once per Newton iteration the group
// assembles A = J M^-1 J^T + diag(C + reg) (the operator of optim/system_operator.py:
// 36-124 written out), lane i holding row i in LD VGPRs, so a PCR matvec is one LDS round
// trip for the broadcast of v followed by LD FMAs: no LDS atomics and no per-body pass
// inside the Krylov loop, and the LDS pipe, which the worlds of a CU otherwise saturate,
// carries only v.  Row i is built by lane i from every row j's M^-1 J_j^T (two bodies,
// staged in LDS) and body indices; rows that share no body give exact zeros, inactive
// rows and the padding up to LD give zero rows and columns (their PCR entries are zero,
// as in the matrix-free operator).  Columns are handled in blocks of 8, and blocks past
// the world's rows are skipped.
// LDS: rowd[16 LD] (M^-1 J0^T 6, M^-1 J1^T 6, b0 bits, b1 bits, pad), v[LD].
//
// Worlds of at most kBodyMajorMax bodies build A body-major instead: row j stages
// M_b^-1 J_j,b^T for every body b (zeros where row j does not touch b) at wb[b][j], so
// lane i forms A_ij = J_i,m0 . wb[m0][j] + J_i,m1 . wb[m1][j] from its own two bodies:
// two 6-dots and no body matching per entry (the same sum: a row couples two distinct
// bodies, so at most one staged vector matches each side).  wb aliases rowd and v
// (neither is live while A is built): B (8 LD + 4) floats, the + 4 staggering the bodies
// over LDS banks.
// ---------------------------------------------------------------------------------
#ifndef OFB_BODY_MAJOR_MAX
#define OFB_BODY_MAJOR_MAX 8
#endif
constexpr int kBodyMajorMax = OFB_BODY_MAJOR_MAX;
__host__ __device__ inline bool dense_body_major(int B) { return B <= kBodyMajorMax; }
__host__ __device__ inline int dense_wb_stride(int ld) { return 8 * ld + 4; }
__host__ __device__ inline int dense_words(int ld, int B) {
  const int w = 17 * ld;
  return dense_body_major(B) && B * dense_wb_stride(ld) > w ? B * dense_wb_stride(ld) : w;
}

struct DenseView {
  float* rowd;
  float* v;
  float* wb;  // body-major staging (aliases rowd), null for the row-major build
  int n_r;
};

template <int LD>
__device__ __forceinline__ DenseView dense_view(float* base, int n_r, int B) {
  return DenseView{base, base + 16 * LD, dense_body_major(B) ? base : nullptr, n_r};
}

// row i of A, in registers
template <int LD>
struct DenseRow {
  float a[LD];
};

// one entry A_ij from row j's staged data (x0..x3) and row i's Jacobian and bodies
__device__ __forceinline__ float dense_entry(const float4& x0, const float4& x1, const float4& x2, const float4& x3,
                                             int m0, int m1, const sv& J0, const sv& J1) {
  const sv w0{mk(x0.x, x0.y, x0.z), mk(x0.w, x1.x, x1.y)};
  const sv w1{mk(x1.z, x1.w, x2.x), mk(x2.y, x2.z, x2.w)};
  const int c0 = __float_as_int(x3.x), c1 = __float_as_int(x3.y);
  float a = 0.f;
  if (m0 == c0) a += sdot(J0, w0);
  if (m0 == c1) a += sdot(J0, w1);
  if (m1 == c0) a += sdot(J1, w0);
  if (m1 == c1) a += sdot(J1, w1);
  return a;
}

__device__ __forceinline__ void dense_stage(float* __restrict__ rowd, int lane, bool on, int b0, int b1,
                                            const sv& W0, const sv& W1) {
  sv w0 = svzero(), w1 = svzero();
  int c0 = -2, c1 = -2;  // an inactive row (or a lane past the world's rows) matches no body
  if (on && b0 >= 0) { w0 = W0; c0 = b0; }
  if (on && b1 >= 0) { w1 = W1; c1 = b1; }
  float4* r = reinterpret_cast<float4*>(rowd + 16 * lane);
  r[0] = make_float4(w0.a.x, w0.a.y, w0.a.z, w0.b.x);
  r[1] = make_float4(w0.b.y, w0.b.z, w1.a.x, w1.a.y);
  r[2] = make_float4(w1.a.z, w1.b.x, w1.b.y, w1.b.z);
  r[3] = make_float4(__int_as_float(c0), __int_as_float(c1), 0.f, 0.f);
}

// W0, W1: M^-1 J0^T, M^-1 J1^T of the lane's row (as computed for the Jacobi diagonal).
// Every lane of the group stages its row, so rows past n_r inside a block carry bodies
// that match nothing.
template <int LD>
__device__ __forceinline__ void dense_build_body_major(const DenseView& D, DenseRow<LD>& A, int lane, int B, bool on,
                                                       int b0, int b1, const sv& J0, const sv& J1, const sv& W0,
                                                       const sv& W1, float Cr, int n_hi) {
  const int stride = dense_wb_stride(LD);
  for (int b = 0; b < B; ++b) {
    sv w = svzero();
    if (on && b == b0) w = W0;
    else if (on && b == b1) w = W1;
    float4* r = reinterpret_cast<float4*>(D.wb + b * stride + 8 * lane);
    r[0] = make_float4(w.a.x, w.a.y, w.a.z, w.b.x);
    r[1] = make_float4(w.b.y, w.b.z, 0.f, 0.f);
  }
  wsync();
  const bool s0 = on && b0 >= 0, s1 = on && b1 >= 0;  // a static side contributes nothing
  const sv Ja = s0 ? J0 : svzero(), Jb = s1 ? J1 : svzero();
  const float* p0 = D.wb + (s0 ? b0 : 0) * stride;
  const float* p1 = D.wb + (s1 ? b1 : 0) * stride;
#pragma unroll
  for (int j0 = 0; j0 < LD; j0 += 8) {
    if (j0 < n_hi) {
#pragma unroll
      for (int j = j0; j < j0 + 8; ++j) {
        const float4 x0 = *reinterpret_cast<const float4*>(p0 + 8 * j);
        const float2 y0 = *reinterpret_cast<const float2*>(p0 + 8 * j + 4);
        const float4 x1 = *reinterpret_cast<const float4*>(p1 + 8 * j);
        const float2 y1 = *reinterpret_cast<const float2*>(p1 + 8 * j + 4);
        float a = sdot(Ja, sv{mk(x0.x, x0.y, x0.z), mk(x0.w, y0.x, y0.y)});
        a += sdot(Jb, sv{mk(x1.x, x1.y, x1.z), mk(x1.w, y1.x, y1.y)});
        if (j == lane && on) a += Cr;
        A.a[j] = a;
      }
    } else {
#pragma unroll
      for (int j = j0; j < j0 + 8; ++j) A.a[j] = 0.f;
    }
  }
  wsync();
}

template <int LD>
__device__ __forceinline__ void dense_build(const DenseView& D, DenseRow<LD>& A, int lane, int B, bool act, int b0,
                                            int b1, const sv& J0, const sv& J1, const sv& W0, const sv& W1,
                                            float Cr) {
  const bool on = lane < D.n_r && act;
  // column blocks past every group's rows are skipped on a wave-uniform bound (the
  // larger row count of the two groups of a wave: a retired group's stale count only
  // widens it), so the branch never splits the wave; blocks below the bound but past a
  // group's own rows read staged rows that match no body and give zeros
  const int n_hi = LD == 32 ? max(__builtin_amdgcn_readlane(D.n_r, 0), __builtin_amdgcn_readlane(D.n_r, 32))
                            : __builtin_amdgcn_readfirstlane(D.n_r);
  if (D.wb) {
    dense_build_body_major<LD>(D, A, lane, B, on, b0, b1, J0, J1, W0, W1, Cr, n_hi);
    return;
  }
  dense_stage(D.rowd, lane, on, b0, b1, W0, W1);
  wsync();
  const int m0 = (on && b0 >= 0) ? b0 : -3, m1 = (on && b1 >= 0) ? b1 : -3;
  const float4* r = reinterpret_cast<const float4*>(D.rowd);
#pragma unroll
  for (int j0 = 0; j0 < LD; j0 += 8) {
    if (j0 < n_hi) {
#pragma unroll
      for (int j = j0; j < j0 + 8; ++j) {
        float a = dense_entry(r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3], m0, m1, J0, J1);
        if (j == lane && on) a += Cr;
        A.a[j] = a;
      }
    } else {
#pragma unroll
      for (int j = j0; j < j0 + 8; ++j) A.a[j] = 0.f;
    }
  }
  wsync();
}

// out_i = sum_j A_ij v_j: v goes through LDS once and its broadcast quads are read in
// batches of eight with no branch between them, so one LDS latency is paid per batch
// (one batch for LD = 32).  Columns past the world's rows hold zeros in A and in v, so
// reading every column leaves the sum unchanged.
template <int LD>
__device__ __forceinline__ float dense_matvec(const DenseView& D, const DenseRow<LD>& A, int lane, bool act,
                                              float v) {
  static_assert(LD % 32 == 0, "dense rows are read in batches of 32 columns");
  const bool on = lane < D.n_r && act;
  D.v[lane] = on ? v : 0.f;
  wsync();
  const float4* v4 = reinterpret_cast<const float4*>(D.v);
  float o[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j0 = 0; j0 < LD; j0 += 32) {
    float4 x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = v4[j0 / 4 + q];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      o[0] += A.a[j0 + 4 * q] * x[q].x;
      o[1] += A.a[j0 + 4 * q + 1] * x[q].y;
      o[2] += A.a[j0 + 4 * q + 2] * x[q].z;
      o[3] += A.a[j0 + 4 * q + 3] * x[q].w;
    }
  }
  wsync();
  return on ? (o[0] + o[1]) + (o[2] + o[3]) : 0.f;
}

// ---------------------------------------------------------------------------------
// Row -> body sums without LDS atomics (dense instances).  Every row couples at most two
// bodies, and many rows share one (all joint rows of a robot touch its chassis), so
// ds_add_f32 from the lanes of a wave serialises on a few addresses and stalls the LDS
// pipe for every wave of the CU.  Instead each row stages its two body indices and
// contributions in a 16-float record (reusing the dense view's staging rows, which are
// free outside dense_build), and lanes over (body, component) sum the records in row
// order: a fixed order, the one of a sequential loop over rows.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void stage_contrib(float* __restrict__ rec, int i, bool on, int b0, int b1, const sv& c0,
                                              const sv& c1) {
  float4* r = reinterpret_cast<float4*>(rec + 16 * i);
  r[0] = make_float4(__int_as_float(on ? b0 : -1), __int_as_float(on ? b1 : -1), 0.f, 0.f);
  r[1] = make_float4(c0.a.x, c0.a.y, c0.a.z, c0.b.x);
  r[2] = make_float4(c0.b.y, c0.b.z, c1.a.x, c1.a.y);
  r[3] = make_float4(c1.a.z, c1.b.x, c1.b.y, c1.b.z);
}

// field[b] (6 floats at body offset `field`) = (zero ? 0 : field[b]) + the records' sums,
// rows 0..n_r-1 in order; a body index < 0 (static side, inactive row) matches nothing
template <int GW>
__device__ __forceinline__ void gather_contrib(float* sm, const float* __restrict__ rec, int lane, int n_r, int B,
                                               int field, bool zero) {
  for (int e = lane; e < 6 * B; e += GW) {
    const int b = e / 6, k = e % 6;
    float acc = zero ? 0.f : sm[b * BS + field + k];
    const float* r = rec;
#pragma unroll 4
    for (int j = 0; j < n_r; ++j, r += 16) {
      const float2 bb = *reinterpret_cast<const float2*>(r);
      const float v0 = r[4 + k], v1 = r[10 + k];
      if (__float_as_int(bb.x) == b) acc += v0;
      if (__float_as_int(bb.y) == b) acc += v1;
    }
    sm[b * BS + field + k] = acc;
  }
}

// ---------------------------------------------------------------------------------
// Row -> body sums through lane-private copies (worlds of at most kRsMaxBodies bodies).
// Each lane owns a copy of the 6 B body accumulators in LDS and adds its rows'
// contributions to it with plain reads and writes of its own words.  The copies are
// stored value-major: value j of lane t at area[j * (G + 4) + t] (G lanes per world), so
// a lane's six words of one body are six loads at immediate offsets from one address,
// the lanes of a wave touch consecutive words, and the lane that sums value j reads
// consecutive copies four at a time (128-bit loads, 16-byte aligned rows; the 4-float
// pad spreads the rows of different values over the banks).  LDS float atomics (ds_add_f32) are not
// used: on gfx950 one of them costs about as much as twenty plain read-modify-writes
// (measured: 24 private ds_add_f32 per lane 4.6 k cycles, the same adds as reads and
// writes 0.6 k), whether or not lanes share addresses.  The copies are then summed per
// (body, component) by lanes that walk the copies in lane order (four interleaved
// chains): a fixed order, the same for every launch shape.  Larger worlds keep the
// staged records (dense instances) or direct LDS atomics.
// ---------------------------------------------------------------------------------
constexpr int kRsMaxBodies = 10;
__host__ __device__ inline bool rs_enabled(int B) { return B <= kRsMaxBodies; }
// floats between the rows of two values of the copies of a world of `lanes` lanes
__host__ __device__ constexpr int rs_row(int lanes) { return lanes + 4; }
__host__ __device__ inline int rs_words(int lanes, int B) { return rs_enabled(B) ? 6 * B * rs_row(lanes) : 0; }
// floats per wave of the per-wave body sums of a world of several waves
__host__ __device__ inline int part_stride(int B) { return 6 * B > 64 ? (6 * B + 3) & ~3 : 64; }

template <int GW, bool DENSE>
struct BodySum {
  float* sm;    // the world's body blocks
  float* area;  // lane-private copies, or the staged records of the dense fallback
  int lane, B;
  bool rs;
  float* part;  // a world of several waves: each wave's sums [NW][part_stride(B)]

  // begin a pass that sums into body field `field`: with `zero` the pass's sums replace
  // the field, else they are added to it
  static constexpr int VS = rs_row(GW);  // floats between the rows of two values

  __device__ __forceinline__ void begin(int field, bool zero) const {
    if (rs) {
      float* p = area + lane;
#pragma unroll
      for (int b = 0; b < kRsMaxBodies; ++b) {
        if (b < B) {
#pragma unroll
          for (int c = 0; c < 6; ++c) p[(6 * b + c) * VS] = 0.f;
        }
      }
    } else if constexpr (!DENSE && GW > WAVE) {
      // worlds of several waves and more than kRsMaxBodies bodies: each wave adds its
      // rows into its own accumulators with LDS atomics (a fixed order inside one wave),
      // and the waves' accumulators are summed in wave order at the end, so the sums do
      // not depend on the order in which the waves run
      float* p = part + (lane / WAVE) * part_stride(B);
      for (int e = lane & (WAVE - 1); e < 6 * B; e += WAVE) p[e] = 0.f;
      wsync();
    } else if (!DENSE && zero) {
      for (int e = lane; e < 6 * B; e += GW) sm[(e / 6) * BS + field + (e % 6)] = 0.f;
      gsync<GW>();
    }
  }
  // contributions c0 to body b0 and c1 to body b1 of row i (none when !on; a body < 0 is
  // a static side)
  __device__ __forceinline__ void add(int field, int i, bool on, int b0, int b1, const sv& c0, const sv& c1) const {
    if (rs) {
      if (!on) return;
      if (b0 >= 0) add_sv(area + 6 * b0 * VS + lane, c0);
      if (b1 >= 0) add_sv(area + 6 * b1 * VS + lane, c1);
    } else if (DENSE) {
      stage_contrib(area, i, on, b0, b1, c0, c1);
    } else if (on) {
      float* base = sm;
      int st = BS, off = field;
      if constexpr (GW > WAVE) {
        base = part + (lane / WAVE) * part_stride(B);
        st = 6;
        off = 0;
      }
      if (b0 >= 0) atomic_add_sv(base + b0 * st + off, c0);
      if (b1 >= 0) atomic_add_sv(base + b1 * st + off, c1);
    }
  }
  // p[c VS] += x_c (c = 0..5) on this lane's own words of one body
  __device__ __forceinline__ static void add_sv(float* p, const sv& x) {
    const float a0 = p[0], a1 = p[VS], a2 = p[2 * VS], a3 = p[3 * VS], a4 = p[4 * VS], a5 = p[5 * VS];
    p[0] = a0 + x.a.x; p[VS] = a1 + x.a.y; p[2 * VS] = a2 + x.a.z;
    p[3 * VS] = a3 + x.b.x; p[4 * VS] = a4 + x.b.y; p[5 * VS] = a5 + x.b.z;
  }
  // completes the pass (rows 0..n_r-1 staged for the record fallback); the caller
  // synchronises before reading the field
  __device__ __forceinline__ void end(int field, bool zero, int n_r) const {
    wsync();
    if (rs) {
      sums(field, zero);
    } else if constexpr (DENSE) {
      gather_contrib<GW>(sm, area, lane, n_r, B, field, zero);
    } else if constexpr (GW > WAVE) {
      __syncthreads();
      for (int e = lane; e < 6 * B; e += GW) {
        float t = part[e];
#pragma unroll
        for (int v = 1; v < GW / WAVE; ++v) t = t + part[v * part_stride(B) + e];
        float* f = sm + (e / 6) * BS + field + (e % 6);
        *f = zero ? t : *f + t;
      }
    }
  }

  // the sum of value j over the copies of lanes t0 .. t0 + nt - 1 (nt is 32 or 64): four
  // copies per 128-bit load, summed in four interleaved chains (copy t0 + u into chain
  // u % 4)
  __device__ __forceinline__ float copies_sum(int j, int t0, int nt) const {
    const float4* q = reinterpret_cast<const float4*>(area + j * VS + t0);
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    for (int t = 0; t < nt; t += 16, q += 4) {
      float4 x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = q[u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a0 += x[u].x;
        a1 += x[u].y;
        a2 += x[u].z;
        a3 += x[u].w;
      }
    }
    return (a0 + a1) + (a2 + a3);
  }
  __device__ __forceinline__ void sums(int field, bool zero) const {
    const int nv = 6 * B;
    if constexpr (GW > WAVE) {
      // each wave sums the copies of its own 64 lanes (as a one-wave world does), then
      // lanes j < 6B add the waves' sums of value j in wave order
      constexpr int NW = GW / WAVE;
      const int wl = lane & (WAVE - 1), wv = lane / WAVE;
      if (nv <= 32) {
        const int j = wl & 31;
        float s = j < nv ? copies_sum(j, wv * WAVE + (wl & 32), 32) : 0.f;
        s = wave_sum_from_halves(s);
        if (wl < nv) part[wv * part_stride(B) + wl] = s;
      } else {
        for (int j = wl; j < nv; j += WAVE) part[wv * part_stride(B) + j] = copies_sum(j, wv * WAVE, WAVE);
      }
      __syncthreads();
      if (lane < nv) {
        float t = part[lane];
#pragma unroll
        for (int v = 1; v < NW; ++v) t = t + part[v * part_stride(B) + lane];
        float* f = sm + (lane / 6) * BS + field + (lane % 6);
        *f = zero ? t : *f + t;
      }
      return;
    }
    if (GW == 64 && nv <= 32) {
      // lanes j and j + 32 sum the two halves of the copies of value j; the swap of
      // the halves gives both the same total
      const int j = lane & 31;
      float s = j < nv ? copies_sum(j, lane & 32, 32) : 0.f;
      s = wave_sum_from_halves(s);
      if (lane < nv) {
        float* f = sm + (lane / 6) * BS + field + (lane % 6);
        *f = zero ? s : *f + s;
      }
    } else {
      for (int j = lane; j < nv; j += GW) {
        const float s = copies_sum(j, 0, GW);
        float* f = sm + (j / 6) * BS + field + (j % 6);
        *f = zero ? s : *f + s;
      }
    }
  }
};

// ||[h_d | h_c]||^2 at the step-start state with lambda = l (and lambda_prev = l, as the
// first Newton iteration snapshots it): the gate between two Newton seeds.
template <int R, int GW, bool DENSE>
__device__ __forceinline__ float initial_merit(const FusedParams& P, float* sm, float* lamprev, const float* js,
                                               const float* cs, const int (&rj)[R], const int (&rk)[R], int lane,
                                               int n_r, int n_c, const float (&l)[R], const BodySum<GW, DENSE>& S,
                                               WorldSum<GW>& WS) {
  const int B = P.m.B;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + GW * k;
    if (i < n_r) lamprev[i] = l[k];
  }
  dynamics_bodies(P, sm, lane, false, GW);
  gsync<GW>();
  S.begin(BHD, false);
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + GW * k;
    if (i < n_r) {
      RowOut o = assemble_row_cached(P, sm, lamprev, js, cs, i, n_c, rj[k], rk[k], l[k], false);
      s += o.hc * o.hc;
      S.add(BHD, i, o.active, o.b0, o.b1, o.J0 * (-l[k] * P.dt), o.J1 * (-l[k] * P.dt));
    }
  }
  S.end(BHD, false, n_r);
  gsync<GW>();
  for (int e = lane; e < 6 * B; e += GW) {
    float v = sm[(e / 6) * BS + BHD + (e % 6)];
    s += v * v;
  }
  return WS.sum(s);
}

// PCR on A x = rhs from x0 = 0 (optim/pcr_solver.py:170-295): the seven vectors live in
// VGPRs, dots are DPP group reductions, A is the LDS matrix-free operator with its body
// sums S (or, DENSE, the assembled matrix of the dense view).  Returns the iteration count.
template <int R, bool BJ, int GW, bool DENSE, class BS_>
__device__ __forceinline__ int pcr_solve(float* sm, int lane, int B, int n_r, const BS_& S, const BJView& V,
                                         const BJRegs<R>& G, const bool (&act)[R], const int (&b0)[R],
                                         const int (&b1)[R], const sv (&J0)[R], const sv (&J1)[R],
                                         const float (&Cr)[R], const float (&Pi)[R], const float (&rhs)[R],
                                         float bsq, float tol_sq, float latol_sq, int max_iters, float (&x)[R],
                                         WorldSum<GW>& WS, const DenseView& D = DenseView{},
                                         const DenseRow<DENSE ? GW : 1>& Ad = DenseRow<DENSE ? GW : 1>{}) {
  static_assert(!DENSE || R == 1, "the dense operator holds one row per lane");
  float r[R], z[R], p[R], Ap[R], y[R], Az[R];
  auto matvec = [&](const float (&v)[R], float (&out)[R]) {
    if constexpr (DENSE) out[0] = dense_matvec(D, Ad, lane, act[0], v[0]);
    else schur_matvec<R, GW>(sm, lane, B, n_r, S, act, b0, b1, J0, J1, Cr, v, out);
  };
#pragma unroll
  for (int k = 0; k < R; ++k) { x[k] = 0.f; r[k] = rhs[k]; z[k] = Pi[k] * r[k]; }
  if (BJ) bj_apply<R>(V, lane, act, Pi, G, r, z);
  matvec(z, Az);
  float zAz = 0.f, rsq = 0.f;
#pragma unroll
  for (int k = 0; k < R; ++k) { p[k] = z[k]; Ap[k] = Az[k]; zAz += z[k] * Az[k]; rsq += r[k] * r[k]; }
  WS.sum2(zAz, rsq);
  float target = latol_sq;
  if (bsq * tol_sq > target) target = bsq * tol_sq;
  int it = 0;
  while (true) {
    float zAz_old = zAz;
    float yAp = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) y[k] = Pi[k] * Ap[k];
    if (BJ) bj_apply<R>(V, lane, act, Pi, G, Ap, y);
#pragma unroll
    for (int k = 0; k < R; ++k) yAp += y[k] * Ap[k];
    yAp = WS.sum(yAp);
    float a = yAp > 0.f ? zAz_old / yAp : 0.f;
    rsq = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      x[k] += a * p[k]; r[k] -= a * Ap[k]; z[k] -= a * y[k];
      rsq += r[k] * r[k];
    }
    matvec(z, Az);
    zAz = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) zAz += z[k] * Az[k];
    WS.sum2(zAz, rsq);  // ||r||^2 of the updated residual rides along
    float bt = zAz_old > 0.f ? zAz / zAz_old : 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) { p[k] = z[k] + bt * p[k]; Ap[k] = Az[k] + bt * Ap[k]; }
    ++it;
    if (!(rsq > target) || it >= max_iters) break;
  }
  return it;
}

// lanes: lanes per world (32, 64, or a multiple of 64 for a world of several waves)
__host__ __device__ inline FusedLayout fused_layout(int B, int J, int rows_cap, int cs_cap, bool bj, bool dense,
                                                    int lanes) {
  FusedLayout l;
  l.lamprev = B * BS;
  l.js = up4(l.lamprev + rows_cap + 1);
  l.cs = l.js + JS_WORDS * J;
  l.bj = l.cs + CS_WORDS * cs_cap;
  l.dense = up4(l.bj + (bj ? bj_words(rows_cap) : 0));
  // the dense instances keep the lane-private body sums in their dense region (free
  // outside the Schur build); the matrix-free ones get a region of their own
  const int rs = rs_words(lanes, B);
  if (dense) {
    const int dw = dense_words(rows_cap, B);
    l.rs = l.dense;
    l.hist = up4(l.dense + (dw > rs ? dw : rs));
    l.part = l.red = l.hist;
  } else {
    l.rs = l.dense;
    const int nw = lanes > WAVE ? lanes / WAVE : 0;  // waves of a multi-wave world
    l.part = up4(l.dense + rs);
    l.red = l.part + nw * part_stride(B);
    l.hist = up4(l.red + 4 * nw);
  }
  l.words = l.hist + kHistWords + 4;  // + the world's running PCR count
  return l;
}

// ------------------------------------------------------------------------------------
// The kernel.  GW = lanes per world (64: one world per wave; 32: two worlds per wave, one
// per half, for worlds of at most 32 rows); R = row slots per lane (rows per world <=
// GW R); BJ = per-body-pair block-Jacobi preconditioner (else Jacobi), a template flag
// so that the Jacobi kernel carries none of the block code's registers; DENSE = the PCR
// operator is the assembled Schur matrix (R = 1 only, a matrix row of GW floats per
// lane).  The dense instances hold the matrix row in VGPRs and are allowed 256 of them
// (two waves per SIMD, no spills); a matrix-free R = 1 instance would be held to 128 so
// that four waves share a SIMD.  GW = 128, 256, 512: one world spread over GW / 64 waves
// of its own workgroup (matrix-free, Jacobi), whose lanes synchronise with the
// workgroup barrier and reduce across waves through LDS: the large worlds' Newton solves
// take fewer dependent steps per lane than on one wave with many row slots, and are the
// ones that set a step's duration.
// ------------------------------------------------------------------------------------
#ifndef OFB_R1_WAVES_PER_EU
#define OFB_R1_WAVES_PER_EU 4
#endif
#ifndef OFB_DENSE_WAVES_PER_EU
#define OFB_DENSE_WAVES_PER_EU 2
#endif
// The paired Jacobi instance sits at the register edge of three waves per SIMD (168
// VGPRs); at two, its launch no longer fits the chip in one round at 4096 worlds (0.52 ->
// 0.72 ms).  The block-Jacobi one would spill heavily at three and keeps two.
#ifndef OFB_PAIR_WAVES_PER_EU
#define OFB_PAIR_WAVES_PER_EU 3
#endif
#ifndef OFB_MW_WAVES_PER_EU
#define OFB_MW_WAVES_PER_EU 2
#endif
#ifndef OFB_RN_WAVES_PER_EU
#define OFB_RN_WAVES_PER_EU 1
#endif
// threads of a workgroup of one-wave (or two-worlds-per-wave) worlds at most: the
// instances are compiled for it, and the launches keep to it
constexpr int kGroupThreads = 256;
template <int R, bool BJ, int GW, bool DENSE>
__global__ void __launch_bounds__(GW > kGroupThreads ? GW : kGroupThreads)
__attribute__((amdgpu_waves_per_eu(GW > WAVE ? OFB_MW_WAVES_PER_EU
                                   : R == 1 ? (DENSE ? (GW == 32 && !BJ ? OFB_PAIR_WAVES_PER_EU : OFB_DENSE_WAVES_PER_EU)
                                                     : OFB_R1_WAVES_PER_EU)
                                            : OFB_RN_WAVES_PER_EU)))
fused_step_kernel(FusedParams P) {
  extern __shared__ __align__(16) float lds[];
  static_assert(GW == 64 || (GW == 32 && R == 1) ||
                    ((GW == 128 || GW == 256 || GW == 512) && R >= 2 && R <= 4 && !DENSE && !BJ),
                "two worlds per wave hold one row slot per lane; worlds of several waves are matrix-free, Jacobi, "
                "with two to four row slots per lane");
  const int lane = threadIdx.x & (GW - 1);
  const int grp = threadIdx.x / GW;
  const int idx = blockIdx.x * (blockDim.x / GW) + grp;
  if (idx >= *P.wcount) return;
  const int w = P.wlist[idx];
  // issue priority of the wave on its SIMD: the classes of larger worlds, whose Newton
  // solves are the longest and set the step's duration, ahead of the small ones that
  // fill the gaps
  __builtin_amdgcn_s_setprio(DENSE ? (GW == 32 ? 0 : 1) : (R >= 3 || GW > WAVE ? 3 : 2));
  float* sm = lds + grp * P.lay.words;
  const int B = P.m.B;
  const RowLayout& L = P.L;
  float* lamprev = sm + P.lay.lamprev;  // compacted row index, rows_cap + 1 entries
  float* js = sm + P.lay.js;            // static record of each joint
  float* cs = sm + P.lay.cs;            // static record of each contact
  const float dt = P.dt;
  const BJView V = bj_view(sm + P.lay.bj, P.rows_cap);
  BJRegs<R> G;
  int bj_flags = 0;

  // ---- load bodies (load_data, base_engine.py:303-308; engine.py:48-50) ----
  for (int b = lane; b < B; b += GW) {
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
  const DenseView D = dense_view<GW>(sm + P.lay.dense, n_r, B);
  DenseRow<DENSE ? GW : 1> Adense;  // row of the assembled Schur matrix (DENSE)
  // row -> body sums: lane-private copies (worlds of at most kRsMaxBodies bodies), else
  // staged records (DENSE) or LDS atomics
  const BodySum<GW, DENSE> S{sm, sm + P.lay.rs, lane, B, rs_enabled(B), sm + P.lay.part};
  WorldSum<GW> WS{sm + P.lay.red, 0};
  gsync<GW>();
  // ---- step-constant row data (model, control, contacts, Jacobians at the step-start
  // poses) into LDS, once per step ----
  for (int j = lane; j < P.m.J; j += GW) store_joint_static(js + JS_WORDS * j, joint_static(P, w, j));
  for (int c = lane; c < n_c; c += GW)
    store_contact_static(cs + CS_WORDS * c, normal_static(P, sm, w, c), friction_static(P, sm, w, c));

  // ---- per-row registers ----
  float lam[R], lamb[R], x[R], Cr[R], Pi[R], rhs[R];
  sv J0[R], J1[R];
  int b0[R], b1[R];
  int rj[R], rk[R];  // joint of a joint / control row, and the row within the joint
  bool act[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + GW * k;
    float l0 = 0.f;
    if (P.cfg.seed_iterate && i < n_r) l0 = P.lambda_seed[(size_t)w * L.Nc + original_row(L, i, n_c)];
    lam[k] = l0; lamb[k] = l0;
    act[k] = false; b0[k] = -1; b1[k] = -1;
    J0[k] = svzero(); J1[k] = svzero(); Cr[k] = 0.f;
    rj[k] = 0; rk[k] = 0;
    if (i < L.Nj) { rj[k] = P.topo[i]; rk[k] = P.topo[L.Nj + i]; }
    else if (i < L.Nj + L.Nctrl) rj[k] = P.topo[2 * L.Nj + (i - L.Nj)];
  }
  gsync<GW>();
  // learned warm start: keep whichever seed starts from the lower residual, so the
  // network never replaces a better previous-step seed
  int seed_flag = 0;
  if (P.cfg.seed_iterate && P.lambda_seed2) {
    float alt[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + GW * k;
      alt[k] = i < n_r ? P.lambda_seed2[(size_t)w * L.Nc + original_row(L, i, n_c)] : 0.f;
    }
    const float ma = initial_merit<R, GW, DENSE>(P, sm, lamprev, js, cs, rj, rk, lane, n_r, n_c, lam, S, WS);
    const float mb = initial_merit<R, GW, DENSE>(P, sm, lamprev, js, cs, rj, rk, lane, n_r, n_c, alt, S, WS);
    if (mb < ma) {
      seed_flag = FLAG_LEARNED_SEED;
#pragma unroll
      for (int k = 0; k < R; ++k) { lam[k] = alt[k]; lamb[k] = alt[k]; }
    }
    gsync<GW>();
  }

  const float atol_sq = P.cfg.nr_atol * P.cfg.nr_atol;
  const float tol_sq = P.cfg.lin_tol >= 0.f ? P.cfg.lin_tol * P.cfg.lin_tol : -1.f;
  const float latol_sq = P.cfg.lin_atol >= 0.f ? P.cfg.lin_atol * P.cfg.lin_atol : -1.f;
  const float reg = P.cfg.regularization;
  int n_it = 0, best_idx = -1;
  // running PCR count of the step, kept by lane 0 in the world's LDS (a register copy came
  // out wrong in the two largest instances, which run at the register file's edge)
  int* const pcr_count = reinterpret_cast<int*>(sm + P.lay.hist + kHistWords);
  if (lane == 0) *pcr_count = 0;
  float best_merit = 0.f, merit = 0.f;

  // rows of the last residual pass (the dense instances have the registers for them)
  constexpr bool kRowCache = DENSE;
  RowOut rcache[R];
  bool rvalid = false;

  OFB_PHASE_INIT
  while (true) {
    // ---- friction-lag snapshot (base_engine.py:441) ----
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + GW * k;
      if (i < n_r) lamprev[i] = lam[k];
    }
    dynamics_bodies(P, sm, lane, true, GW);
    gsync<GW>();
    OFB_PHASE(0)
    // ---- assembly (linear_system.py:140-322) ----
    S.begin(BHD, false);
    float hc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + GW * k;
      hc[k] = 0.f;
      if (i < n_r) {
        RowOut o;
        if (kRowCache && rvalid && i < L.Nj + L.Nctrl + n_c) {
          // joint, control and normal rows do not read lambda_prev: the residual pass of
          // the previous iteration evaluated them at this same (q, u, lambda); its
          // -J^T lambda dt goes into h_d again
          o = rcache[k];
        } else {
          o = assemble_row_cached(P, sm, lamprev, js, cs, i, n_c, rj[k], rk[k], lam[k], false);
        }
        S.add(BHD, i, o.active, o.b0, o.b1, o.J0 * (-lam[k] * dt), o.J1 * (-lam[k] * dt));
        J0[k] = o.J0; J1[k] = o.J1; b0[k] = o.b0; b1[k] = o.b1; act[k] = o.active;
        hc[k] = o.hc;
        Cr[k] = o.C + reg;
        // inactive contact / friction rows force lambda := 0 (contact_constraint.py:289-296,
        // friction_constraint.py:713-787)
        if (!o.active && i >= L.Nj + L.Nctrl) lam[k] = 0.f;
      } else {  // a slot past the world's rows: no row, zero Jacobians (read by the sums)
        act[k] = false;
        J0[k] = svzero(); J1[k] = svzero(); b0[k] = -1; b1[k] = -1; Cr[k] = 0.f;
      }
    }
    S.end(BHD, false, n_r);
    gsync<GW>();
    OFB_PHASE(1)
    // M^-1 h_d per body
    for (int b = lane; b < B; b += GW) {
      float* bb = sm + b * BS;
      stsv(bb + BVV, apply_minv(bb, ldsv(bb + BHD)));
    }
    gsync<GW>();
    // Schur RHS (linear_system.py:17-65) and Jacobi diagonal (preconditioner.py:9-50)
    float bsq = 0.f;
    sv W0[R], W1[R];  // M^-1 J^T of each row's two sides
#pragma unroll
    for (int k = 0; k < R; ++k) {
      rhs[k] = 0.f; Pi[k] = 0.f;
      W0[k] = svzero(); W1[k] = svzero();
      if (act[k]) {
        float s = 0.f, d = 0.f;
        if (b0[k] >= 0) {
          const float* bb = sm + b0[k] * BS;
          s += sdot(J0[k], ldsv(bb + BVV));
          W0[k] = apply_minv(bb, J0[k]);
          d += sdot(J0[k], W0[k]);
        }
        if (b1[k] >= 0) {
          const float* bb = sm + b1[k] * BS;
          s += sdot(J1[k], ldsv(bb + BVV));
          W1[k] = apply_minv(bb, J1[k]);
          d += sdot(J1[k], W1[k]);
        }
        rhs[k] = (s - hc[k]) / dt;
        Pi[k] = 1.f / (d + Cr[k] + 1e-6f);
      }
      bsq += rhs[k] * rhs[k];
    }
    bsq = WS.sum(bsq);
    if (BJ) {
      if (n_it == 0) bj_structure<R, GW>(V, lane, n_r, B, act, b0, b1, G, bj_flags);
      bj_factor<R>(V, sm, lane, b0, b1, J0, J1, Cr, G, bj_flags);
    }

    OFB_PHASE(2)
    if constexpr (DENSE) dense_build(D, Adense, lane, B, act[0], b0[0], b1[0], J0[0], J1[0], W0[0], W1[0], Cr[0]);
    OFB_PHASE(7)
    const int pcr_it = pcr_solve<R, BJ, GW, DENSE>(sm, lane, B, n_r, S, V, G, act, b0, b1, J0, J1, Cr, Pi, rhs,
                                                   bsq, tol_sq, latol_sq, P.cfg.lin_max_iters, x, WS, D, Adense);
    if (lane == 0) *pcr_count += pcr_it;
    OFB_PHASE(3)

    // ---- velocity recovery + Newton update (linear_system.py:344-377, newton_step.py:22-54) ----
    S.begin(BACC, true);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      S.add(BACC, lane + GW * k, act[k], b0[k], b1[k], J0[k] * x[k], J1[k] * x[k]);
      if (P.ls_count == 0) lam[k] = lam[k] + x[k];
    }
    S.end(BACC, true, n_r);
    gsync<GW>();
    if (P.ls_count == 0) {
      for (int b = lane; b < B; b += GW) {
        float* bb = sm + b * BS;
        sv acc = ldsv(bb + BACC), hd = ldsv(bb + BHD);
        sv du = apply_minv(bb, sv{acc.a * dt - hd.a, acc.b * dt - hd.b});
        sv u = ldsv(bb + BU);
        u = sv{u.a + du.a, u.b + du.b};
        stsv(bb + BU, u);
        sttf(bb + BQ, integrate(u, ldtf(bb + BQP), ld3(bb + BCOM), dt));
      }
      gsync<GW>();
    } else {
      // merit line search (core/linesearch.py:330-339): every candidate alpha gives
      // u = u0 + alpha du, q = integrate(u, q_prev), lambda = lambda0 + alpha dlambda, with
      // lambda_prev unchanged (:222-223); the first argmin of ||h||^2 is kept
      // (reductions.py batched_argmin)
      for (int b = lane; b < B; b += GW) {
        float* bb = sm + b * BS;
        sv acc = ldsv(bb + BACC), hd = ldsv(bb + BHD);
        stsv(bb + BVV, apply_minv(bb, sv{acc.a * dt - hd.a, acc.b * dt - hd.b}));
        stsv(bb + BU0, ldsv(bb + BU));
      }
      gsync<GW>();
      float best_a = 0.f, best_m = 0.f;
      for (int si = 0; si < P.ls_count; ++si) {
        const float a = P.ls_steps[si];
        for (int b = lane; b < B; b += GW) {
          float* bb = sm + b * BS;
          sv u0 = ldsv(bb + BU0), du = ldsv(bb + BVV);
          sv u = sv{u0.a + du.a * a, u0.b + du.b * a};
          stsv(bb + BU, u);
          sttf(bb + BQ, integrate(u, ldtf(bb + BQP), ld3(bb + BCOM), dt));
        }
        gsync<GW>();
        float la[R];
#pragma unroll
        for (int k = 0; k < R; ++k) la[k] = lam[k] + a * x[k];
        dynamics_bodies(P, sm, lane, false, GW);
        gsync<GW>();
        S.begin(BHD, false);
        float ms = 0.f;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          int i = lane + GW * k;
          if (i < n_r) {
            RowOut o = assemble_row_cached(P, sm, lamprev, js, cs, i, n_c, rj[k], rk[k], la[k], false);
            ms += o.hc * o.hc;
            S.add(BHD, i, o.active, o.b0, o.b1, o.J0 * (-la[k] * dt), o.J1 * (-la[k] * dt));
          }
        }
        S.end(BHD, false, n_r);
        gsync<GW>();
        for (int e = lane; e < 6 * B; e += GW) {
          float v = sm[(e / 6) * BS + BHD + (e % 6)];
          ms += v * v;
        }
        ms = WS.sum(ms);
        if (si == 0 || ms < best_m) { best_m = ms; best_a = a; }
        gsync<GW>();
      }
#pragma unroll
      for (int k = 0; k < R; ++k) lam[k] += best_a * x[k];
      for (int b = lane; b < B; b += GW) {
        float* bb = sm + b * BS;
        sv u0 = ldsv(bb + BU0), du = ldsv(bb + BVV);
        sv u = sv{u0.a + du.a * best_a, u0.b + du.b * best_a};
        stsv(bb + BU, u);
        sttf(bb + BQ, integrate(u, ldtf(bb + BQP), ld3(bb + BCOM), dt));
      }
      gsync<GW>();
    }

    OFB_PHASE(4)
    // ---- residual at the new iterate (core/residual.py:18-173) ----
    dynamics_bodies(P, sm, lane, false, GW);
    gsync<GW>();
    S.begin(BHD, false);
    float msum = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int i = lane + GW * k;
      if (i < n_r) {
        RowOut o = assemble_row_cached(P, sm, lamprev, js, cs, i, n_c, rj[k], rk[k], lam[k], false);
        msum += o.hc * o.hc;
        if (kRowCache) rcache[k] = o;
        S.add(BHD, i, o.active, o.b0, o.b1, o.J0 * (-lam[k] * dt), o.J1 * (-lam[k] * dt));
      }
    }
    rvalid = true;
    S.end(BHD, false, n_r);
    gsync<GW>();
    for (int e = lane; e < 6 * B; e += GW) {
      float v = sm[(e / 6) * BS + BHD + (e % 6)];
      msum += v * v;
    }
    merit = WS.sum(msum);
    if (P.hist_merit && lane == 0 && n_it < kHistIters) {  // ||h||^2 and the PCR iterations so far
      sm[P.lay.hist + n_it] = merit;
      sm[P.lay.hist + kHistIters + n_it] = __int_as_float(*pcr_count);
    }
    OFB_PHASE(5)

    // ---- candidate bookkeeping (history_group.py + backtracking.py:12-95) ----
    int kidx = n_it;
    ++n_it;
    bool take = (kidx == P.cfg.nr_backtrack_min_iter) || (kidx > P.cfg.nr_backtrack_min_iter && merit < best_merit);
    if (take) {
      best_idx = kidx;
      best_merit = merit;
      for (int e = lane; e < 13 * B; e += GW) {
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
    gsync<GW>();
  }
  gsync<GW>();
  OFB_PHASE_FLUSH
  // ---- pick the candidate ----
  int picked = n_it - 1;
  float picked_merit = merit;
  if (n_it > P.cfg.nr_backtrack_min_iter && best_idx >= 0) {
    picked = best_idx;
    picked_merit = best_merit;
    for (int e = lane; e < 13 * B; e += GW) {
      int b = e / 13, c = e % 13;
      float* bb = sm + b * BS;
      if (c < 7) bb[BQ + c] = bb[BBQ + c];
      else bb[BU + c - 7] = bb[BBU + c - 7];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) lam[k] = lamb[k];
  }
  gsync<GW>();
  // ---- outputs (engine.py:68-69) ----
  for (int e = lane; e < 13 * B; e += GW) {
    int b = e / 13, c = e % 13;
    const float* bb = sm + b * BS;
    size_t k = (size_t)w * B + b;
    if (c < 7) P.out_q[k * 7 + c] = bb[BQ + c];
    else P.out_qd[k * 6 + c - 7] = bb[BU + c - 7];
  }
  // lambda and lambda_prev_iter (the snapshot of the last Newton iteration, which
  // backtracking does not restore, base_engine.py:441) in the reference row layout
#pragma unroll
  for (int k = 0; k < R; ++k) {
    int i = lane + GW * k;
    if (i < n_r) {
      size_t o = (size_t)w * L.Nc + original_row(L, i, n_c);
      P.lambda_out[o] = lam[k];
      if (P.lambda_prev_out) P.lambda_prev_out[o] = lamprev[i];
    }
  }
  // contact slots beyond the world's contact count hold lambda = 0
  for (int c = n_c + lane; c < L.K; c += GW) {
    size_t o = (size_t)w * L.Nc;
    P.lambda_out[o + L.off_n + c] = 0.f;
    P.lambda_out[o + L.off_f + 2 * c] = 0.f;
    P.lambda_out[o + L.off_f + 2 * c + 1] = 0.f;
    if (P.lambda_prev_out) {
      P.lambda_prev_out[o + L.off_n + c] = 0.f;
      P.lambda_prev_out[o + L.off_f + 2 * c] = 0.f;
      P.lambda_prev_out[o + L.off_f + 2 * c + 1] = 0.f;
    }
  }
  if (P.hist_merit) {  // NaN / -1: iterations the world did not run (or past kHistIters)
    gsync<GW>();
    const float* hs = sm + P.lay.hist;
    for (int k = lane; k < P.hist_cap; k += GW) {
      const bool ran = k < n_it && k < kHistIters;
      const int before = (ran && k > 0) ? __float_as_int(hs[kHistIters + k - 1]) : 0;
      P.hist_merit[(size_t)k * P.m.W + w] = ran ? hs[k] : __builtin_nanf("");
      P.hist_pcr[(size_t)k * P.m.W + w] = ran ? __float_as_int(hs[kHistIters + k]) - before : -1;
    }
  }
  int bj_any = 0;
  if constexpr (BJ) {
    if (group_any<GW>(bj_flags & FLAG_FACTOR_FAILURE)) bj_any |= FLAG_FACTOR_FAILURE;
    if (group_any<GW>(bj_flags & FLAG_PAIR_OVERFLOW)) bj_any |= FLAG_PAIR_OVERFLOW;
  }
  if (lane == 0) {
    P.st_nr[w] = n_it;
    P.st_pcr[w] = *pcr_count;
    P.st_merit[w] = picked_merit;
    P.st_best[w] = picked;
    int fl = (P.st_flags[w] & FLAG_CONTACT_OVERFLOW) | bj_any | seed_flag;
    if (n_it >= P.cfg.nr_max_iters && merit > atol_sq) fl |= FLAG_NR_CAP;
    P.st_flags[w] = fl;
  }
}

// Class code of the instance for worlds spread over NW waves with R row slots per lane:
// kMultiWave + NW + kMultiWaveR (R - 2) (NW = 2, 4, 8; R = 2, or 4 for the wide layout
// of large worlds).  (One slot per lane on twice the waves was slower: 1.70 against 1.57
// ms per replayed step on the uneven skid-steer batch, 5.2 against 4.5 ms on the tracked
// robots; the extra barriers outweigh the shorter lanes.)
constexpr int kMultiWave = 100;
constexpr int kMultiWaveR = 32;
__host__ __device__ constexpr int multiwave_code(int nw, int r) { return kMultiWave + nw + kMultiWaveR * (r - 2); }

// Host side of one class instance: op 0 launches it, op 1 raises its dynamic-LDS limit
// to `lds` bytes.  Instances are compiled in several translation units
// (fb_fused_classes.hip) so that the build runs in parallel.
template <int R, bool BJ, int GW, bool DENSE>
int fused_class_op(int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t stream) {
  if (op == 1)
    return hipFuncSetAttribute((const void*)fused_step_kernel<R, BJ, GW, DENSE>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess ? 0 : -4;
  hipLaunchKernelGGL((fused_step_kernel<R, BJ, GW, DENSE>), dim3(blocks), dim3(threads), lds, stream, P);
  return 0;
}

// defined in fb_fused_classes.hip, one per (preconditioner, class set) translation unit;
// return -1 when the class code is not in that unit's set.  Class code 0 is the
// two-worlds-per-wave dense instance (at most 32 rows), code 1 the one-world-per-wave
// dense instance (at most 64 rows), code R >= 2 the one-world-per-wave matrix-free
// instance with R row slots per lane, code multiwave_code(NW, R) the instance whose world
// spans NW waves (Jacobi only).
int fused_class_op_bj0_set0(int R, int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int fused_class_op_bj0_set1(int R, int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int fused_class_op_bj0_set2(int R, int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int fused_class_op_bj1_set0(int R, int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t st);
int fused_class_op_bj1_set1(int R, int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t st);

}  // namespace ofb
