// fb_reduce.hip -- per-pair contact reduction on the packed contacts (collision/top_k.py,
// fps.py, cluster.py), one wavefront per world.
//
// The reference runs each policy as a (world, contact) launch whose pair leaders (the
// smallest contact index of each (body0, body1) pair) work serially, then compacts the
// survivors with an atomic slot counter.  Here a world's contacts are staged in LDS
// (bodies, depth, midpoint, normal) and:
//   * top_k: every lane ranks its contacts within their pair (depth descending, ties
//     by index) and keeps rank < K;
//   * fps: pairs in leader order; the seed is the pair's first deepest contact, then
//     K - 1 rounds each add the contact whose smallest squared midpoint distance to
//     the picked ones is largest (ties: smallest index), found by a wave argmax;
//   * cluster: pairs in leader order; pass 1 marks the pair's representatives (no
//     deeper contact of the pair within the normal and position thresholds); pass 2
//     walks every marked contact of the world in index order and drops it when K or
//     more of the pair's marked representatives are deeper -- the reference's pass 2
//     also visits earlier pairs' representatives (cluster.py:205-238), kept as is;
// and survivors are compacted in index order (the order the reference's serial scatter
// gives), with the world's count set to the number kept.
// Depths and midpoints use the step-start poses, as the reference's reducers read
// body_pose_prev during load_data (base_engine.py:317-320).  The reducer's geometry
// (depths, midpoints, normal dots, squared distances) is evaluated in FP64 from the FP32
// inputs: its decisions are comparisons, and contact patches are full of near-ties (the
// corners of a flat face sit at equal depths and equal distances), which FP32 rounding
// would decide by accident; in FP64 they fall as the exact geometry decides them (exact
// ties by contact index, as in the reference).

#include <hip/hip_runtime.h>

#include <climits>

#include "fb_solver.h"

namespace ofb {

namespace {

constexpr int kRedWave = 64;
constexpr int kRedWords = 17;  // depth, mid xyz, n xyz (FP64: 14 words), b0, b1, keep

struct RedView {
  double* depth;
  double* mid;  // [3][n]
  double* nrm;  // [3][n]
  int* b0;
  int* b1;
  int* keep;
};

__device__ __forceinline__ RedView red_view(float* base, int K) {
  RedView v;
  v.depth = reinterpret_cast<double*>(base);
  v.mid = v.depth + K;
  v.nrm = v.mid + 3 * K;
  v.b0 = reinterpret_cast<int*>(v.nrm + 3 * K);
  v.b1 = v.b0 + K;
  v.keep = v.b1 + K;
  return v;
}

// floats of LDS per wavefront, even so that every wavefront's doubles stay 8-byte aligned
__host__ __device__ inline int red_stride(int K) { return (kRedWords * K + 1) & ~1; }

struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 dmk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return dmk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return dmk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return dmk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
  return dmk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// p' = X.p + rotate(X.q, p), quaternion xyzw (the solver's transform convention)
__device__ __forceinline__ d3 dtpoint(const float* X, d3 p) {
  const d3 u = dmk(X[3], X[4], X[5]);
  const double w = X[6];
  const double c = 2.0 * w * w - 1.0, d = 2.0 * ddot(u, p);
  return dmk(X[0], X[1], X[2]) + (p * c + u * d + dcross(u, p) * (2.0 * w));
}

__device__ __forceinline__ void red_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a is deeper than b (ties: the smaller index), top_k.py:134-139
__device__ __forceinline__ bool red_deeper(double da, int ia, double db, int ib) {
  return da > db || (da == db && ia < ib);
}

// wave argmax of (key, -idx): every lane returns the winner's index (INT_MAX when no lane
// has a candidate, key = -inf)
__device__ __forceinline__ int red_argmax(double key, int idx) {
  for (int off = 1; off < kRedWave; off <<= 1) {
    const double k2 = __shfl_xor(key, off);
    const int i2 = __shfl_xor(idx, off);
    if (k2 > key || (k2 == key && i2 < idx)) { key = k2; idx = i2; }
  }
  return idx;
}

__device__ __forceinline__ int red_sum(int v) {
  for (int off = 1; off < kRedWave; off <<= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void __launch_bounds__(256) reduce_contacts_kernel(int W, int B, int SC, int K, int policy, int kpair,
                                                              float ndot_th, float pos_th,
                                                              const float* __restrict__ body_q,
                                                              const int32_t* __restrict__ shape_body,
                                                              int32_t* __restrict__ ccount, float* __restrict__ cf,
                                                              int32_t* __restrict__ ci) {
  extern __shared__ __align__(16) float red_lds[];
  const int lane = threadIdx.x & (kRedWave - 1), wave = threadIdx.x / kRedWave;
  const int w = blockIdx.x * (blockDim.x / kRedWave) + wave;
  if (w >= W) return;
  const int n = min(ccount[w], K);
  if (n <= 0) return;
  const RedView V = red_view(red_lds + (size_t)wave * red_stride(K), K);
  float* f = cf + (size_t)w * CF * K;
  int32_t* s = ci + (size_t)w * 2 * K;
  // stage: bodies, depth and midpoint (_contact_depth top_k.py:55-87, fps.py:50-86)
  for (int c = lane; c < n; c += kRedWave) {
    const int s0 = s[c], s1 = s[K + c];
    const int b0 = s0 >= 0 ? shape_body[w * SC + s0] : -1;
    const int b1 = s1 >= 0 ? shape_body[w * SC + s1] : -1;
    const d3 nn = dmk(f[CF_NX * K + c], f[CF_NY * K + c], f[CF_NZ * K + c]);
    d3 pa = dmk(f[CF_P0X * K + c], f[CF_P0Y * K + c], f[CF_P0Z * K + c]);
    d3 pb = dmk(f[CF_P1X * K + c], f[CF_P1Y * K + c], f[CF_P1Z * K + c]);
    if (b0 >= 0) pa = dtpoint(body_q + ((size_t)w * B + b0) * 7, pa);
    if (b1 >= 0) pb = dtpoint(body_q + ((size_t)w * B + b1) * 7, pb);
    pa = pa - nn * (double)f[CF_TH0 * K + c];
    pb = pb + nn * (double)f[CF_TH1 * K + c];
    const d3 m = (pa + pb) * 0.5;
    V.b0[c] = b0; V.b1[c] = b1;
    V.depth[c] = ddot(nn, pb - pa);
    V.mid[c] = m.x; V.mid[K + c] = m.y; V.mid[2 * K + c] = m.z;
    V.nrm[c] = nn.x; V.nrm[K + c] = nn.y; V.nrm[2 * K + c] = nn.z;
    V.keep[c] = 0;
  }
  red_sync();
  auto same = [&](int i, int j) { return V.b0[i] == V.b0[j] && V.b1[i] == V.b1[j]; };
  auto mid = [&](int i) { return dmk(V.mid[i], V.mid[K + i], V.mid[2 * K + i]); };
  if (policy == OFB_REDUCE_TOP_K) {  // rank_within_pair_kernel (top_k.py:90-143)
    for (int c = lane; c < n; c += kRedWave) {
      int rank = 0;
      for (int j = 0; j < n; ++j)
        if (j != c && same(j, c) && red_deeper(V.depth[j], j, V.depth[c], c)) ++rank;
      V.keep[c] = rank < kpair ? 1 : 0;
    }
    red_sync();
  } else {
    const double pos_sq = (double)pos_th * (double)pos_th;
    for (int p = 0; p < n; ++p) {  // pair leaders in contact order
      bool leader = true;
      for (int j = lane; j < p; j += kRedWave)
        if (same(j, p)) leader = false;
      if (__any(!leader)) continue;
      if (policy == OFB_REDUCE_FPS) {  // fps_per_pair_kernel (fps.py:89-225)
        double key = -__builtin_huge_val();
        int idx = INT_MAX;
        for (int j = lane; j < n; j += kRedWave)
          if (same(j, p) && red_deeper(V.depth[j], j, key, idx)) { key = V.depth[j]; idx = j; }
        int pick = red_argmax(key, idx);
        for (int r = 0; r < kpair && pick != INT_MAX; ++r) {
          if (lane == 0) V.keep[pick] = 1;
          red_sync();
          if (r + 1 == kpair) break;
          key = -__builtin_huge_val();
          idx = INT_MAX;
          for (int j = lane; j < n; j += kRedWave) {
            if (!same(j, p) || V.keep[j]) continue;
            const d3 mj = mid(j);
            double md = 1.0e30;
            for (int k = 0; k < n; ++k) {
              if (!V.keep[k] || !same(k, p)) continue;
              const d3 d = mj - mid(k);
              md = fmin(md, ddot(d, d));
            }
            if (md > key || (md == key && j < idx)) { key = md; idx = j; }
          }
          pick = red_argmax(key, idx);
        }
      } else {  // cluster_per_pair_kernel (cluster.py:94-238)
        for (int j = lane; j < n; j += kRedWave) {
          if (!same(j, p)) continue;
          const d3 nj = dmk(V.nrm[j], V.nrm[K + j], V.nrm[2 * K + j]);
          const d3 mj = mid(j);
          bool rep = true;
          for (int k = 0; k < n; ++k) {
            if (k == j || !same(k, p)) continue;
            const d3 nk = dmk(V.nrm[k], V.nrm[K + k], V.nrm[2 * K + k]);
            const d3 d = mj - mid(k);
            if (ddot(nj, nk) > (double)ndot_th && ddot(d, d) < pos_sq && red_deeper(V.depth[k], k, V.depth[j], j))
              rep = false;
          }
          V.keep[j] = rep ? 1 : 0;
        }
        red_sync();
        for (int j = 0; j < n; ++j) {
          if (!V.keep[j]) continue;
          int cnt = 0;
          for (int k = lane; k < n; k += kRedWave)
            if (k != j && V.keep[k] && same(k, p) && red_deeper(V.depth[k], k, V.depth[j], j)) ++cnt;
          if (red_sum(cnt) >= kpair) {
            if (lane == 0) V.keep[j] = 0;
            red_sync();
          }
        }
      }
      red_sync();
    }
  }
  // stable compaction (scatter in contact order), in place: a survivor only moves to a
  // lower slot, and each group of 64 reads all its sources before writing
  const unsigned long long below = (1ull << lane) - 1ull;
  int base = 0;
  for (int c0 = 0; c0 < n; c0 += kRedWave) {
    const int c = c0 + lane;
    const bool kp = c < n && V.keep[c];
    const unsigned long long m = __ballot(kp);
    const int dst = base + __popcll(m & below);
    float v[CF];
    int32_t i0 = 0, i1 = 0;
    if (kp) {
#pragma unroll
      for (int k = 0; k < CF; ++k) v[k] = f[k * K + c];
      i0 = s[c]; i1 = s[K + c];
    }
    red_sync();
    if (kp && dst != c) {
#pragma unroll
      for (int k = 0; k < CF; ++k) f[k * K + dst] = v[k];
      s[dst] = i0; s[K + dst] = i1;
    }
    base += __popcll(m);
    red_sync();
  }
  if (lane == 0) ccount[w] = base;
}

}  // namespace

int reduce_contacts(ofb_solver* s, const float* body_q, hipStream_t stream) {
  const int policy = s->cfg.contact_reduction;
  if (policy == OFB_REDUCE_NONE) return 0;
  const int W = s->m.W, K = s->L.K, T = 256, wpb = T / kRedWave;
  const size_t lds = (size_t)wpb * red_stride(K) * sizeof(float);
  hipLaunchKernelGGL(reduce_contacts_kernel, dim3((W + wpb - 1) / wpb), dim3(T), lds, stream, W, s->m.B, s->m.SC, K,
                     policy, s->cfg.contact_max_per_pair, s->cfg.cluster_normal_dot, s->cfg.cluster_pos, body_q,
                     s->m.shape_body, s->ccount, s->cf, s->ci);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("contact reduction launch failed: %s", hipGetErrorString(e));
    return -4;
  }
  return 0;
}

}  // namespace ofb
