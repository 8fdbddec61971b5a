// fb_warmstart.hip -- cross-step contact warm start (collision/warm_start.py, methods
// position_match :84-519 and pair_aggregate :896-1010) for both solver paths.
//
// apply (before the solve, one wavefront per world): every current contact looks for the
// previous step's contact of the same body pair whose start-of-step midpoint, moved by
// the pair's tangential velocity * dt, lies closest within max(5 mm, 0.1 |v_t| dt); a
// match copies lambda_n and re-expresses lambda_t in the new tangent basis; unmatched
// contacts get the cold-start seeds (residual gravity share, impact impulse, sliding
// friction).  The result is the Newton seed lambda_seed [W*N_c] (reference layout; joint
// and control rows 0).  snapshot (after the solve): the step's contacts, start-of-step
// midpoints and lambda.  Both hooks read the start-of-step pose and velocity
// (body_pose_prev / body_vel_prev, warm_start.py:600-609).
//
// The cold-start kernel of the reference takes the force on body 0 as -n lambda, the
// convention from before its contact packing flipped Newton's normals
// (core/contacts.py:80-83), while its contact rows push body 0 along +n; it is restated
// as is, so for physical contacts its gravity and impact terms stay silent.

#include <hip/hip_runtime.h>

#include "fb_math.h"
#include "fb_solver.h"

namespace ofb {

namespace {

constexpr float kWsAlpha = 0.1f, kWsMinThreshold = 5.0e-3f, kWsNDotUpMin = 0.3f;  // warm_start.py:520-538

struct WsParams {
  DeviceModel m;
  RowLayout L;
  float dt;
  int method;  // OFB_WS_*
  int cold_gravity, cold_impact;
  float v_t_threshold;
  const float* q_prev;   // [W*B*7] start of step
  const float* qd_prev;  // [W*B*6]
  const int32_t* ccount;
  const float* cf;
  const int32_t* ci;
  const float* lambda;   // [W*Nc] of the finished step (snapshot)
  int32_t* prev_count;   // [W]
  int32_t *prev_b0, *prev_b1;  // [W*K]
  float *prev_p, *prev_n;      // [W*K*3]
  float *prev_ln, *prev_lt;    // [W*K], [W*K*2]
  float* seed;                 // [W*Nc]
  float* feat;                 // [W*K*12] learned warm-start features (may be null)
  float* fscale;               // [W*K] m_eff |g| of each contact
};

__device__ __forceinline__ tf body_pose(const WsParams& P, int w, int b) {
  if (b < 0) return tident();
  const float* x = P.q_prev + ((size_t)w * P.m.B + b) * 7;
  return tf{mk(x[0], x[1], x[2]), qt{x[3], x[4], x[5], x[6]}};
}

__device__ __forceinline__ v3 cvec(const WsParams& P, int w, int c, int k0) {
  const int K = P.L.K;
  const float* f = P.cf + (size_t)w * CF * K + c;
  return mk(f[k0 * K], f[(k0 + 1) * K], f[(k0 + 2) * K]);
}

// _world_midpoint (warm_start.py:51-82)
__device__ __forceinline__ v3 midpoint(const WsParams& P, int w, int c, int b0, int b1, v3 n) {
  const int K = P.L.K;
  const float* f = P.cf + (size_t)w * CF * K + c;
  v3 a = tpoint(body_pose(P, w, b0), cvec(P, w, c, CF_P0X)) - n * f[CF_TH0 * K];
  v3 b = tpoint(body_pose(P, w, b1), cvec(P, w, c, CF_P1X)) + n * f[CF_TH1 * K];
  return (a + b) * 0.5f;
}

__device__ __forceinline__ v3 lin_vel(const WsParams& P, int w, int b) {
  const float* u = P.qd_prev + ((size_t)w * P.m.B + b) * 6;
  return mk(u[0], u[1], u[2]);
}

__device__ __forceinline__ v3 ang_vel(const WsParams& P, int w, int b) {
  const float* u = P.qd_prev + ((size_t)w * P.m.B + b) * 6;
  return mk(u[3], u[4], u[5]);
}

// Features of contact c for the learned warm start (ostrich_amd/learned.py): gap,
// normal, tangential relative velocity in the contact basis, the matched previous
// impulse, mu, dt and the mass ratio, made dimensionless with |g|, dt and m_eff.
__device__ void write_features(const WsParams& P, int w, int c, int b0, int b1, int s0, int s1, v3 n, v3 t1, v3 t2,
                               float gmag, float lnm, float lt1, float lt2) {
  const int K = P.L.K, B = P.m.B, SC = P.m.SC;
  const float* f = P.cf + (size_t)w * CF * K + c;
  const tf q0 = body_pose(P, w, b0), q1 = body_pose(P, w, b1);
  const v3 p0 = tpoint(q0, cvec(P, w, c, CF_P0X)) - n * f[CF_TH0 * K];
  const v3 p1 = tpoint(q1, cvec(P, w, c, CF_P1X)) + n * f[CF_TH1 * K];
  const float gap = dot(n, p0 - p1);
  v3 v0 = mk(0.f, 0.f, 0.f), v1 = v0;
  if (b0 >= 0) {
    const float* cm = P.m.body_com + ((size_t)w * B + b0) * 3;
    v0 = lin_vel(P, w, b0) + cross(ang_vel(P, w, b0), p0 - tpoint(q0, mk(cm[0], cm[1], cm[2])));
  }
  if (b1 >= 0) {
    const float* cm = P.m.body_com + ((size_t)w * B + b1) * 3;
    v1 = lin_vel(P, w, b1) + cross(ang_vel(P, w, b1), p1 - tpoint(q1, mk(cm[0], cm[1], cm[2])));
  }
  const v3 vr = v0 - v1;
  const float im0 = b0 >= 0 ? P.m.body_inv_mass[w * B + b0] : 0.f;
  const float im1 = b1 >= 0 ? P.m.body_inv_mass[w * B + b1] : 0.f;
  const float meff = (im0 + im1) > 0.f ? 1.f / (im0 + im1) : 1.f;
  const float g = gmag > 0.f ? gmag : 1.f, dt = P.dt;
  const float sc = meff * g;
  float* o = P.feat + ((size_t)w * K + c) * 12;
  o[0] = gap / (g * dt * dt);
  o[1] = n.x; o[2] = n.y; o[3] = n.z;
  o[4] = dot(vr, t1) / (g * dt);
  o[5] = dot(vr, t2) / (g * dt);
  o[6] = lnm / sc; o[7] = lt1 / sc; o[8] = lt2 / sc;
  o[9] = 0.5f * (P.m.shape_mu[w * SC + s0] + P.m.shape_mu[w * SC + s1]);
  o[10] = dt / 0.05f;
  o[11] = (im0 + im1) > 0.f ? im1 / (im0 + im1) : 0.5f;
  P.fscale[(size_t)w * K + c] = sc;
}

// one int per body pair (bodies -1 .. 65534): equal keys, equal pairs
__device__ __forceinline__ int pair_key(int b0, int b1) { return ((b0 + 1) << 16) | (b1 + 1); }

__global__ void __launch_bounds__(64) ws_apply_kernel(WsParams P) {
  extern __shared__ __align__(16) float ws_lds[];
  const int w = blockIdx.x, lane = threadIdx.x;
  const int B = P.m.B, K = P.L.K, Nc = P.L.Nc, SC = P.m.SC;
  // the previous step's contacts of this world, staged once for the scans below (every
  // current contact walks all of them): midpoint (one 128-bit load), body pair as one
  // key (one load to reject another pair's contact), normal impulse
  float4* ppos = reinterpret_cast<float4*>(ws_lds);          // [K]
  int* pkey = reinterpret_cast<int*>(ppos + K);              // [K]
  float* pln = reinterpret_cast<float*>(pkey + K);           // [K]
  float* support = pln + K;                                  // [B]
  int* unmatched = reinterpret_cast<int*>(support + B);      // [B]
  int* matched = unmatched + B;                              // [K]
  for (int b = lane; b < B; b += 64) { support[b] = 0.f; unmatched[b] = 0; }
  {
    const int n_prev0 = min(P.prev_count[w], K);
    for (int j = lane; j < n_prev0; j += 64) {
      const size_t kj = (size_t)w * K + j;
      ppos[j] = make_float4(P.prev_p[kj * 3], P.prev_p[kj * 3 + 1], P.prev_p[kj * 3 + 2], 0.f);
      pkey[j] = pair_key(P.prev_b0[kj], P.prev_b1[kj]);
      pln[j] = P.prev_ln[kj];
    }
  }
  for (int i = lane; i < Nc; i += 64) P.seed[(size_t)w * Nc + i] = 0.f;
  if (P.feat)
    for (int e = lane; e < K * 12; e += 64) P.feat[(size_t)w * K * 12 + e] = 0.f;
  if (P.fscale)
    for (int e = lane; e < K; e += 64) P.fscale[(size_t)w * K + e] = 0.f;
  __syncthreads();
  const v3 g = mk(P.m.gravity[0], P.m.gravity[1], P.m.gravity[2]);
  const float gmag = sqrtf(dot(g, g));
  const v3 up = gmag > 1e-6f ? g * (-1.f / gmag) : mk(0.f, 0.f, 1.f);
  const int on = P.L.off_n, of = P.L.off_f;
  const int n_c = min(P.ccount[w], K), n_prev = min(P.prev_count[w], K);
  const float dt = P.dt;
  float* seed = P.seed + (size_t)w * Nc;
  // _match_kernel (:84-297)
  for (int c = lane; c < K; c += 64) {
    matched[c] = 0;
    if (c >= n_c) continue;
    const int s0 = P.ci[(size_t)w * 2 * K + c], s1 = P.ci[(size_t)w * 2 * K + K + c];
    if (s0 == s1) continue;
    const int b0 = s0 >= 0 ? P.m.shape_body[w * SC + s0] : -1;
    const int b1 = s1 >= 0 ? P.m.shape_body[w * SC + s1] : -1;
    const v3 nc = cvec(P, w, c, CF_NX);
    if (P.method == OFB_WS_PAIR_AGGREGATE) {
      // pair_aggregate (_pair_count_current_kernel, _pair_apply_kernel, :943-1010): the
      // pair's previous total normal impulse, summed in contact order, spread evenly over
      // the pair's current contacts; lambda_t stays 0, and the cold start's gravity share
      // sees no unmatched contacts
      float total = 0.f;
      int n_pp = 0, n_cur = 0;
      const int key = pair_key(b0, b1);
      for (int j = 0; j < n_prev; ++j) {
        if (pkey[j] == key) { total += pln[j]; ++n_pp; }
      }
      for (int j = 0; j < n_c; ++j) {
        const int t0 = P.ci[(size_t)w * 2 * K + j], t1 = P.ci[(size_t)w * 2 * K + K + j];
        if (t0 == t1) continue;
        const int c0 = t0 >= 0 ? P.m.shape_body[w * SC + t0] : -1;
        const int c1 = t1 >= 0 ? P.m.shape_body[w * SC + t1] : -1;
        if (c0 == b0 && c1 == b1) ++n_cur;
      }
      float lnm = 0.f;
      if (n_pp > 0 && n_cur > 0) {
        lnm = total / (float)n_cur;
        seed[on + c] = lnm;
        matched[c] = 1;
      }
      if (P.feat) {
        v3 ct1, ct2;
        orthogonal_basis(nc, ct1, ct2);
        write_features(P, w, c, b0, b1, s0, s1, nc, ct1, ct2, gmag, lnm, 0.f, 0.f);
      }
      continue;
    }
    const v3 pc = midpoint(P, w, c, b0, b1, nc);
    v3 vt = mk(0.f, 0.f, 0.f);
    if (b0 >= 0) vt = vt + lin_vel(P, w, b0);
    if (b1 >= 0) vt = vt - lin_vel(P, w, b1);
    vt = vt - nc * dot(vt, nc);
    const float thr = fmaxf(kWsMinThreshold, kWsAlpha * sqrtf(dot(vt, vt)) * dt);
    float best_d = thr * thr;
    int best = -1;
    const int key = pair_key(b0, b1);
    for (int j = 0; j < n_prev; ++j) {
      if (pkey[j] != key) continue;
      const float4 q = ppos[j];
      v3 pp = mk(q.x, q.y, q.z) + vt * dt;
      v3 d = pc - pp;
      float dsq = dot(d, d);
      if (dsq < best_d) { best_d = dsq; best = j; }
    }
    const float ndu = dot(nc, up);
    const bool sup0 = -ndu > kWsNDotUpMin, sup1 = ndu > kWsNDotUpMin;
    v3 ct1, ct2;
    orthogonal_basis(nc, ct1, ct2);
    float lnm = 0.f, lt1 = 0.f, lt2 = 0.f;
    if (best >= 0) {
      size_t kb = (size_t)w * K + best;
      v3 pn = mk(P.prev_n[kb * 3], P.prev_n[kb * 3 + 1], P.prev_n[kb * 3 + 2]);
      v3 pt1, pt2;
      orthogonal_basis(pn, pt1, pt2);
      v3 ltw = pt1 * P.prev_lt[kb * 2] + pt2 * P.prev_lt[kb * 2 + 1];
      lnm = P.prev_ln[kb];
      lt1 = dot(ltw, ct1);
      lt2 = dot(ltw, ct2);
    }
    if (P.feat) write_features(P, w, c, b0, b1, s0, s1, nc, ct1, ct2, gmag, lnm, lt1, lt2);
    if (best < 0) {
      if (b0 >= 0 && sup0) atomicAdd(&unmatched[b0], 1);
      if (b1 >= 0 && sup1) atomicAdd(&unmatched[b1], 1);
      continue;
    }
    seed[on + c] = lnm;
    seed[of + 2 * c] = lt1;
    seed[of + 2 * c + 1] = lt2;
    matched[c] = 1;
    if (b0 >= 0 && sup0) atomicAdd(&support[b0], -ndu * lnm);
    if (b1 >= 0 && sup1) atomicAdd(&support[b1], ndu * lnm);
  }
  __syncthreads();
  // _cold_start_kernel (:300-462)
  for (int c = lane; c < n_c; c += 64) {
    if (matched[c]) continue;
    const int s0 = P.ci[(size_t)w * 2 * K + c], s1 = P.ci[(size_t)w * 2 * K + K + c];
    if (s0 == s1) continue;
    const int b0 = s0 >= 0 ? P.m.shape_body[w * SC + s0] : -1;
    const int b1 = s1 >= 0 ? P.m.shape_body[w * SC + s1] : -1;
    const v3 n = cvec(P, w, c, CF_NX);
    const v3 p = midpoint(P, w, c, b0, b1, n);
    v3 v0 = mk(0.f, 0.f, 0.f), v1 = v0;
    // velocity about the body origin, as the reference computes it (:383-392)
    if (b0 >= 0) v0 = lin_vel(P, w, b0) + cross(ang_vel(P, w, b0), p - body_pose(P, w, b0).p);
    if (b1 >= 0) v1 = lin_vel(P, w, b1) + cross(ang_vel(P, w, b1), p - body_pose(P, w, b1).p);
    const v3 vr = v1 - v0;
    const float vn = dot(vr, n);
    const v3 vtt = vr - n * vn;
    float l_imp = 0.f;
    if (P.cold_impact && vn < 0.f) {
      float im = (b0 >= 0 ? P.m.body_inv_mass[w * B + b0] : 0.f) + (b1 >= 0 ? P.m.body_inv_mass[w * B + b1] : 0.f);
      if (im > 1e-9f) l_imp = (1.f / im) * (-vn) / dt;
    }
    const float ndu = dot(n, up);
    float l_grav = 0.f;
    if (P.cold_gravity) {
      if (b0 >= 0 && -ndu > kWsNDotUpMin && unmatched[b0] > 0) {
        float res = P.m.body_mass[w * B + b0] * gmag - support[b0];
        if (res > 0.f) l_grav = fmaxf(l_grav, res / ((float)unmatched[b0] * (-ndu)));
      }
      if (b1 >= 0 && ndu > kWsNDotUpMin && unmatched[b1] > 0) {
        float res = P.m.body_mass[w * B + b1] * gmag - support[b1];
        if (res > 0.f) l_grav = fmaxf(l_grav, res / ((float)unmatched[b1] * ndu));
      }
    }
    const float ln = fmaxf(l_imp, l_grav);
    if (ln > 0.f) seed[on + c] = ln;
    if (P.v_t_threshold > 0.f && ln > 0.f) {
      float vsq = dot(vtt, vtt);
      if (vsq > P.v_t_threshold * P.v_t_threshold) {
        v3 th = vtt * (-1.f / sqrtf(vsq));
        float mu = 0.5f * (P.m.shape_mu[w * SC + s0] + P.m.shape_mu[w * SC + s1]);
        v3 ft = th * (mu * ln);
        v3 t1, t2;
        orthogonal_basis(n, t1, t2);
        seed[of + 2 * c] = dot(ft, t1);
        seed[of + 2 * c + 1] = dot(ft, t2);
      }
    }
  }
}

// learned seed: contact rows from the network's (n, t1, t2) outputs times m_eff |g|,
// joint and control rows 0 (as the warm starter leaves them)
__global__ void ws_learned_seed_kernel(int W, int K, RowLayout L, const int32_t* __restrict__ ccount,
                                       const float* __restrict__ y, const float* __restrict__ fscale,
                                       float* __restrict__ seed2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * L.Nc) return;
  const int w = t / L.Nc, i = t % L.Nc;
  const int n_c = min(ccount[w], K);
  float v = 0.f;
  int c = -1, comp = 0;
  if (i >= L.off_n && i < L.off_n + K) { c = i - L.off_n; comp = 0; }
  else if (i >= L.off_f && i < L.off_f + 2 * K) { c = (i - L.off_f) >> 1; comp = 1 + ((i - L.off_f) & 1); }
  if (c >= 0 && c < n_c) {
    const size_t k = (size_t)w * K + c;
    v = y[k * 3 + comp] * fscale[k];
  }
  seed2[t] = v;
}

// _snapshot_kernel (:465-519); one thread per (world, slot)
__global__ void ws_snapshot_kernel(WsParams P) {
  const int K = P.L.K, W = P.m.W, SC = P.m.SC;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * K) return;
  const int w = t / K, c = t % K;
  const int n_c = min(P.ccount[w], K);
  if (c == 0) P.prev_count[w] = n_c;
  if (c >= n_c) return;
  const size_t k = (size_t)w * K + c;
  const int s0 = P.ci[(size_t)w * 2 * K + c], s1 = P.ci[(size_t)w * 2 * K + K + c];
  if (s0 == s1) { P.prev_b0[k] = -2; P.prev_b1[k] = -2; return; }  // never matched
  const int b0 = s0 >= 0 ? P.m.shape_body[w * SC + s0] : -1;
  const int b1 = s1 >= 0 ? P.m.shape_body[w * SC + s1] : -1;
  const v3 n = cvec(P, w, c, CF_NX);
  const v3 mp = midpoint(P, w, c, b0, b1, n);
  P.prev_b0[k] = b0; P.prev_b1[k] = b1;
  P.prev_p[k * 3] = mp.x; P.prev_p[k * 3 + 1] = mp.y; P.prev_p[k * 3 + 2] = mp.z;
  P.prev_n[k * 3] = n.x; P.prev_n[k * 3 + 1] = n.y; P.prev_n[k * 3 + 2] = n.z;
  const float* lam = P.lambda + (size_t)w * P.L.Nc;
  P.prev_ln[k] = lam[P.L.off_n + c];
  P.prev_lt[k * 2] = lam[P.L.off_f + 2 * c];
  P.prev_lt[k * 2 + 1] = lam[P.L.off_f + 2 * c + 1];
}

WsParams ws_params(ofb_solver* s, const ofb_state_in* in, float dt) {
  WsParams P;
  P.m = s->m; P.L = s->L; P.dt = dt;
  P.method = s->cfg.warm_start_method;
  P.cold_gravity = s->cfg.warm_start_cold_gravity; P.cold_impact = s->cfg.warm_start_cold_impact;
  P.v_t_threshold = s->cfg.warm_start_cold_friction_v;
  P.q_prev = in->body_q; P.qd_prev = in->body_qd;
  P.ccount = s->ccount; P.cf = s->cf; P.ci = s->ci; P.lambda = s->lambda;
  P.prev_count = s->ws_count; P.prev_b0 = s->ws_b0; P.prev_b1 = s->ws_b1; P.prev_p = s->ws_p; P.prev_n = s->ws_n;
  P.prev_ln = s->ws_ln; P.prev_lt = s->ws_lt; P.seed = s->lambda_seed;
  P.feat = s->ws_feat; P.fscale = s->ws_fscale;
  return P;
}

}  // namespace

int warm_start_apply(ofb_solver* s, const ofb_state_in* in, float dt, hipStream_t stream) {
  if (s->m.B > 65534) {  // the match keys a body pair in one int
    set_error("the warm start supports at most 65534 bodies per world");
    return -2;
  }
  WsParams P = ws_params(s, in, dt);
  size_t lds = (size_t)(7 * s->L.K + 2 * s->m.B) * 4;
  hipLaunchKernelGGL(ws_apply_kernel, dim3(s->m.W), dim3(64), lds, stream, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("warm start launch failed: %s", hipGetErrorString(e)); return -4; }
  return 0;
}

int warm_start_learned_seed(ofb_solver* s, const float* y, hipStream_t stream) {
  const int n = s->m.W * s->L.Nc;
  hipLaunchKernelGGL(ws_learned_seed_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, s->m.W, s->L.K, s->L,
                     s->ccount, y, s->ws_fscale, s->lambda_seed2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("learned seed launch failed: %s", hipGetErrorString(e)); return -4; }
  return 0;
}

int warm_start_snapshot(ofb_solver* s, const ofb_state_in* in, float dt, hipStream_t stream) {
  WsParams P = ws_params(s, in, dt);
  const int n = s->m.W * s->L.K;
  hipLaunchKernelGGL(ws_snapshot_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("warm start snapshot failed: %s", hipGetErrorString(e)); return -4; }
  return 0;
}

}  // namespace ofb
