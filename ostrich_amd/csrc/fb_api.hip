// fb_api.hip -- C ABI entry points (include/ostrich_fb.h): solver lifetime, contact
// packing and step dispatch.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "fb_solver.h"

namespace ofb {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

#define OFB_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error("%s failed: %s", #expr, hipGetErrorString(e_));                     \
      return -4;                                                                     \
    }                                                                                \
  } while (0)

// Line-search candidates (_compute_linesearch_step_size_array, core/engine_data.py:15-39):
// `nc` log-spaced values in [min_step, upper] and `no` linear ones in 1 -/+ window/2,
// sorted ascending, the one closest to 1 set to exactly 1; computed in double, stored
// as float.
static int linesearch_steps(const ofb_config* c, float* out, int cap) {
  const int nc = c->linesearch_conservative_count, no = c->linesearch_optimistic_count;
  if (nc < 1 || no < 1 || nc + no > cap) return -1;
  std::vector<double> v;
  const double a = std::log10((double)c->linesearch_min_step), b = std::log10((double)c->linesearch_conservative_upper);
  for (int i = 0; i < nc; ++i) {
    double e = nc == 1 ? a : (i == nc - 1 ? b : a + i * ((b - a) / (nc - 1)));
    v.push_back(std::pow(10.0, e));
  }
  const double h = (double)c->linesearch_optimistic_window / 2.0, lo = 1.0 - h, hi = 1.0 + h;
  for (int i = 0; i < no; ++i) v.push_back(no == 1 ? lo : (i == no - 1 ? hi : lo + i * ((hi - lo) / (no - 1))));
  std::sort(v.begin(), v.end());
  size_t best = 0;
  for (size_t i = 1; i < v.size(); ++i)
    if (std::fabs(v[i] - 1.0) < std::fabs(v[best] - 1.0)) best = i;
  v[best] = 1.0;
  for (size_t i = 0; i < v.size(); ++i) out[i] = (float)v[i];
  return (int)v.size();
}

static int joint_row_count(int type) {  // core/model.py:88-99
  static const int m[8] = {5, 5, 3, 6, 0, 1, 6, 0};
  return (type >= 0 && type < 8) ? m[type] : 0;
}

// ---------------------------------------------------------------------------------
// Contact packing (core/contacts.py:25-99).  The reference assigns slots with a
// per-world atomic counter, which orders a world's contacts by thread id on Warp's CPU
// device and arbitrarily on a GPU.  Here the order is the input order, on every device,
// in two passes over the valid prefix of the flat arrays (the producer's count):
//   1. each contact finds its world and reserves a member slot with a per-world atomic
//      counter (the count is order-independent; the member order is not);
//   2. one wavefront per world ranks its members by contact index (ascending = input
//      order) and writes contact r to slot r.
// A world with more than K contacts keeps its first K in input order: its wavefront
// scans the world keys of the valid prefix with ballots (a rare path, flagged with
// FLAG_CONTACT_OVERFLOW).  The counter keeps counting past K, as the reference's does
// (contacts.py:73-76).
// ---------------------------------------------------------------------------------
constexpr int kPackWave = 64;  // lanes per wavefront (gfx950)

__device__ __forceinline__ int contact_world(int a, int b, const int32_t* __restrict__ shape_world) {
  if (a == b) return -1;
  int w0 = a >= 0 ? shape_world[a] : -1;
  int w1 = b >= 0 ? shape_world[b] : -1;
  return w0 >= 0 ? w0 : w1;
}

__global__ void pack_reset_kernel(int W, int32_t* __restrict__ ccount, int32_t* __restrict__ flags) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  ccount[w] = 0;
  flags[w] = 0;
}

__global__ void pack_members_kernel(int n_max, const int32_t* __restrict__ count, const int32_t* __restrict__ s0,
                                    const int32_t* __restrict__ s1, const int32_t* __restrict__ shape_world, int W,
                                    int K, int32_t* __restrict__ ccount, int32_t* __restrict__ members,
                                    int32_t* __restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = i < n_max && i < count[0];
  int w = -1;
  if (live) {
    w = contact_world(s0[i], s1[i], shape_world);
    keys[i] = w;
  }
  const bool take = live && w >= 0 && w < W;
  // Contacts usually arrive grouped by world, so the lanes of a wave form runs of one
  // world: the first lane of each run reserves the run's slots with one atomic add (one
  // per lane contended on the same counter), and the run's lanes take consecutive slots.
  // A lane that takes no slot is a run of its own.  pack_world_kernel orders each world's
  // members by input index afterwards, so the slot order is immaterial.
  const int key = take ? w : -1 - lane;
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || key != prev;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
  const int first = 63 - __clzll(heads & upto);  // this lane's run starts here
  const unsigned long long after = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
  const int end = after ? __ffsll(after) - 1 : 64;  // the next run starts here
  int base = 0;
  if (take && lane == first) base = atomicAdd(ccount + w, end - lane);
  base = __shfl(base, first);
  if (take) {
    const int slot = base + (lane - first);
    if (slot < K) members[(size_t)w * K + slot] = i;
  }
}

__device__ __forceinline__ void pack_write(int w, int slot, int i, int G, int S, int K, const float* __restrict__ p0,
                                           const float* __restrict__ p1, const float* __restrict__ nrm,
                                           const int32_t* __restrict__ s0, const int32_t* __restrict__ s1,
                                           const float* __restrict__ m0, const float* __restrict__ m1,
                                           float* __restrict__ cf, int32_t* __restrict__ ci) {
  float* f = cf + (size_t)w * CF * K + slot;
  f[CF_P0X * K] = p0[i * 3 + 0]; f[CF_P0Y * K] = p0[i * 3 + 1]; f[CF_P0Z * K] = p0[i * 3 + 2];
  f[CF_P1X * K] = p1[i * 3 + 0]; f[CF_P1Y * K] = p1[i * 3 + 1]; f[CF_P1Z * K] = p1[i * 3 + 2];
  // Newton emits the normal from shape0 to shape1; the solver uses the opposite sign
  // (core/contacts.py:80-83)
  f[CF_NX * K] = -nrm[i * 3 + 0]; f[CF_NY * K] = -nrm[i * 3 + 1]; f[CF_NZ * K] = -nrm[i * 3 + 2];
  f[CF_TH0 * K] = m0[i];
  f[CF_TH1 * K] = m1[i];
  int a = s0[i], b = s1[i];
  auto local = [&](int s) { return s < G ? S + s : (s - G) % S; };  // contacts.py:10-22
  ci[(size_t)w * 2 * K + slot] = a >= 0 ? local(a) : a;
  ci[(size_t)w * 2 * K + K + slot] = b >= 0 ? local(b) : b;
}

// one wavefront per world; LDS: K member indices per wavefront
__global__ void __launch_bounds__(256) pack_world_kernel(int W, int G, int S, int K, int n_max,
                                                         const int32_t* __restrict__ count,
                                                         const int32_t* __restrict__ ccount,
                                                         const int32_t* __restrict__ members,
                                                         const int32_t* __restrict__ keys,
                                                         const float* __restrict__ p0, const float* __restrict__ p1,
                                                         const float* __restrict__ nrm,
                                                         const int32_t* __restrict__ s0,
                                                         const int32_t* __restrict__ s1,
                                                         const float* __restrict__ m0, const float* __restrict__ m1,
                                                         float* __restrict__ cf, int32_t* __restrict__ ci,
                                                         int32_t* __restrict__ flags) {
  extern __shared__ int32_t mem_lds[];
  const int lane = threadIdx.x & (kPackWave - 1), wave = threadIdx.x / kPackWave;
  const int w = blockIdx.x * (blockDim.x / kPackWave) + wave;
  if (w >= W) return;
  const int c = ccount[w];
  if (c <= 0) return;
  if (c <= K) {
    int32_t* m = mem_lds + wave * K;
    for (int e = lane; e < c; e += kPackWave) m[e] = members[(size_t)w * K + e];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int e = lane; e < c; e += kPackWave) {
      const int i = m[e];
      int rank = 0;
      for (int q = 0; q < c; ++q) rank += m[q] < i ? 1 : 0;
      pack_write(w, rank, i, G, S, K, p0, p1, nrm, s0, s1, m0, m1, cf, ci);
    }
    return;
  }
  // overflow: the first K contacts of the world in input order
  if (lane == 0) flags[w] |= FLAG_CONTACT_OVERFLOW;
  const int n = min(n_max, count[0]);
  const unsigned long long below = (1ull << lane) - 1ull;
  int found = 0;
  for (int base = 0; base < n && found < K; base += kPackWave) {
    const int i = base + lane;
    const bool mine = i < n && keys[i] == w;
    const unsigned long long mask = __ballot(mine);
    const int rank = found + __popcll(mask & below);
    if (mine && rank < K) pack_write(w, rank, i, G, S, K, p0, p1, nrm, s0, s1, m0, m1, cf, ci);
    found += __popcll(mask);
  }
}

static bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone;
}

static int pack_contacts(ofb_solver* s, const ofb_contacts* c, hipStream_t stream) {
  const int W = s->m.W, n = c->max_contacts;
  if (n > s->rigid_contact_cap) {
    set_error("contacts.max_contacts=%d exceeds config.rigid_contact_max=%d", n, s->rigid_contact_cap);
    return -2;
  }
  const int T = 256;
  hipLaunchKernelGGL(pack_reset_kernel, dim3((W + T - 1) / T), dim3(T), 0, stream, W, s->ccount, s->st_flags);
  if (n == 0) return 0;
  const int K = s->L.K;
  hipLaunchKernelGGL(pack_members_kernel, dim3((n + T - 1) / T), dim3(T), 0, stream, n, c->count, c->shape0,
                     c->shape1, s->m.shape_world, W, K, s->ccount, s->pk_members, s->pk_keys);
  const int wpb = T / kPackWave;
  hipLaunchKernelGGL(pack_world_kernel, dim3((W + wpb - 1) / wpb), dim3(T), (size_t)wpb * K * sizeof(int32_t), stream,
                     W, s->m.G, s->m.S, K, n, c->count, s->ccount, s->pk_members, s->pk_keys, c->point0, c->point1,
                     c->normal, c->shape0, c->shape1, c->margin0, c->margin1, s->cf, s->ci, s->st_flags);
  OFB_CHECK(hipGetLastError());
  return 0;
}

// component-major packed contacts -> row-major [W*K(*3)] arrays (diagnostics)
__global__ void unpack_contacts_kernel(int W, int K, const float* __restrict__ cf, const int32_t* __restrict__ ci,
                                       int32_t* shape0, int32_t* shape1, float* point0, float* point1, float* normal,
                                       float* th0, float* th1) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * K) return;
  int w = t / K, c = t % K;
  const float* f = cf + (size_t)w * CF * K + c;
  if (shape0) shape0[t] = ci[(size_t)w * 2 * K + c];
  if (shape1) shape1[t] = ci[(size_t)w * 2 * K + K + c];
  if (th0) th0[t] = f[CF_TH0 * K];
  if (th1) th1[t] = f[CF_TH1 * K];
  for (int d = 0; d < 3; ++d) {
    if (point0) point0[t * 3 + d] = f[(CF_P0X + d) * K];
    if (point1) point1[t * 3 + d] = f[(CF_P1X + d) * K];
    if (normal) normal[t * 3 + d] = f[(CF_NX + d) * K];
  }
}

template <class T>
static int dev_copy(T** dst, const T* src, size_t n) {
  *dst = nullptr;
  if (n == 0) return 0;
  if (!src) { set_error("model array missing"); return -1; }
  OFB_CHECK(hipMalloc((void**)dst, n * sizeof(T)));
  OFB_CHECK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyDefault));
  return 0;
}

template <class T>
static int dev_alloc(T** dst, size_t n) {
  *dst = nullptr;
  if (n == 0) n = 1;
  OFB_CHECK(hipMalloc((void**)dst, n * sizeof(T)));
  OFB_CHECK(hipMemset(*dst, 0, n * sizeof(T)));
  return 0;
}

}  // namespace ofb

using namespace ofb;

extern "C" {

const char* ofb_last_error(void) { return g_err; }

int ofb_destroy(ofb_solver* s) {
  if (!s) return 0;
  DeviceModel& m = s->m;
  void* ptrs[] = {m.body_mass, m.body_inv_mass, m.body_inertia, m.body_inv_inertia, m.body_com, m.joint_type,
                  m.joint_parent, m.joint_child, m.joint_qd_start, m.joint_enabled, m.joint_X_p, m.joint_X_c,
                  m.joint_compliance, m.joint_axis, m.joint_dof_mode, m.joint_target_ke, m.joint_target_kd,
                  m.shape_body, m.shape_mu, m.shape_friction_axis, m.shape_mu_perp, m.shape_mu_stiction_scale,
                  m.shape_v_stribeck, m.shape_stribeck_lateral_only, m.shape_world, s->ccount, s->cf, s->ci,
                  s->pk_keys, s->pk_members, s->topo,
                  s->lambda, s->lambda_prev, s->lambda_seed, s->ws_count, s->ws_b0, s->ws_b1, s->ws_p, s->ws_n,
                  s->ws_ln, s->ws_lt, s->ws_feat, s->ws_fscale, s->ws_y, s->ws_off, s->lambda_seed2, s->ls_steps, s->st_nr,
                  s->st_pcr, s->st_best, s->st_flags, s->st_merit};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->plain) plain_destroy(s);
  if (s->events_ready) {
    (void)hipEventDestroy(s->ev_solve0); (void)hipEventDestroy(s->ev_solve1);
    (void)hipEventDestroy(s->ev_kernel0); (void)hipEventDestroy(s->ev_kernel1);
  }
  fused_classes_free(s);
  delete s;
  return 0;
}

int ofb_create(const ofb_model* model, const ofb_config* config, int32_t path, ofb_solver** out_solver) {
  if (!model || !config || !out_solver) { set_error("null argument"); return -1; }
  *out_solver = nullptr;
  const int W = model->num_worlds, B = model->bodies_per_world, J = model->joints_per_world,
            D = model->dofs_per_world, S = model->shapes_per_world, G = model->num_global_shapes;
  if (W < 1 || B < 1 || J < 0 || D < 0 || S < 0 || G < 0) { set_error("invalid model dimensions"); return -2; }
  if (J > kMaxJoints) { set_error("at most %d joints per world are supported (got %d)", kMaxJoints, J); return -2; }
  if (config->nr_max_iters < 1 || config->nr_backtrack_min_iter < 0 ||
      config->nr_backtrack_min_iter >= config->nr_max_iters) {
    set_error("invalid Newton iteration limits (engine_config.py:110-123)");
    return -2;
  }
  if (config->lin_max_iters < 1) { set_error("linear.max_iters must be >= 1"); return -2; }
  if (config->max_contacts_per_world < 1) { set_error("contacts.max_per_world must be >= 1"); return -2; }
  if (config->contact_reduction < OFB_REDUCE_NONE || config->contact_reduction > OFB_REDUCE_CLUSTER) {
    set_error("unknown contact reduction policy %d", config->contact_reduction);
    return -2;
  }
  if (config->large_world_slots != 0 && config->large_world_slots != 2 && config->large_world_slots != 4) {
    set_error("large_world_slots must be 2 or 4 (got %d)", config->large_world_slots);
    return -2;
  }
  if (config->warm_start_method != OFB_WS_POSITION_MATCH && config->warm_start_method != OFB_WS_PAIR_AGGREGATE) {
    set_error("unknown warm-start method %d", config->warm_start_method);
    return -2;
  }
  if (config->contact_reduction != OFB_REDUCE_NONE && config->contact_max_per_pair < 1) {
    set_error("contacts.reduction.max_per_pair must be >= 1 (collision/config.py:49-52)");
    return -2;
  }
  if (config->preconditioner != OFB_PRECOND_JACOBI && config->preconditioner != OFB_PRECOND_PER_BODY_PAIR) {
    set_error("unknown preconditioner"); return -2;
  }
  if (path != OFB_PATH_FUSED && path != OFB_PATH_PLAIN) { set_error("unknown execution path"); return -2; }
  if (path == OFB_PATH_FUSED && config->termination != OFB_TERM_PER_WORLD) {
    set_error("the fused path retires worlds independently: use termination=per_world (global needs the plain path)");
    return -2;
  }
  if (path == OFB_PATH_FUSED && B > 64) { set_error("the fused path supports at most 64 bodies per world"); return -2; }
  if (config->linesearch && path != OFB_PATH_FUSED) {
    set_error("the merit line search runs on the fused path");
    return -2;
  }
  if (path == OFB_PATH_PLAIN && config->preconditioner == OFB_PRECOND_PER_BODY_PAIR) {
    set_error("the plain path preconditions with Jacobi only: use the fused path for per_body_pair");
    return -2;
  }
  ofb_solver* s = new ofb_solver;
  std::memset(s, 0, sizeof(*s));
  s->cfg = *config;
  if (s->cfg.large_world_slots == 0) s->cfg.large_world_slots = 2;
  s->path = path;
  s->keep_lambda_prev = 1;
  DeviceModel& m = s->m;
  m.W = W; m.B = B; m.J = J; m.D = D; m.S = S; m.G = G; m.SC = S + G;
  for (int i = 0; i < 3; ++i) m.gravity[i] = model->gravity[i];
  m.flat_shapes = G + W * S;
  const size_t WB = (size_t)W * B, WJ = (size_t)W * J, WD = (size_t)W * D, WS = (size_t)W * (S + G);
  // host copies of the topology of world 0 (identical worlds, core/model.py:200-203)
  std::vector<int32_t> jt(J), jq(J), dm(D > 0 ? D : 1);
  int rc = 0;
  if (J > 0) {
    if (hipMemcpy(jt.data(), model->joint_type, J * sizeof(int32_t), hipMemcpyDefault) != hipSuccess ||
        hipMemcpy(jq.data(), model->joint_qd_start, J * sizeof(int32_t), hipMemcpyDefault) != hipSuccess) {
      set_error("cannot read the joint topology");
      delete s;
      return -4;
    }
  }
  if (D > 0 && hipMemcpy(dm.data(), model->joint_dof_mode, D * sizeof(int32_t), hipMemcpyDefault) != hipSuccess) {
    set_error("cannot read the dof modes");
    delete s;
    return -4;
  }
  RowLayout& L = s->L;
  L.Nj = 0; L.Nctrl = 0;
  std::vector<int32_t> topo_j, topo_k, topo_c;
  for (int j = 0; j < J; ++j) {
    L.joint_off[j] = L.Nj;
    int rcn = joint_row_count(jt[j]);
    for (int k = 0; k < rcn; ++k) { topo_j.push_back(j); topo_k.push_back(k); }
    L.Nj += rcn;
  }
  for (int j = 0; j < J; ++j) {
    L.ctrl_off[j] = -1;
    if ((jt[j] == 0 || jt[j] == 1) && jq[j] >= 0 && jq[j] < D && dm[jq[j]] != 0) {
      L.ctrl_off[j] = L.Nctrl++;
      topo_c.push_back(j);
    }
  }
  L.K = config->max_contacts_per_world;
  L.Nc = L.Nj + L.Nctrl + 3 * L.K;
  L.off_ctrl = L.Nj;
  L.off_n = L.Nj + L.Nctrl;
  L.off_f = L.off_n + L.K;
  std::vector<int32_t> topo(topo_j);
  topo.insert(topo.end(), topo_k.begin(), topo_k.end());
  topo.insert(topo.end(), topo_c.begin(), topo_c.end());
  if (topo.empty()) topo.push_back(0);
#define COPY(field, n) if ((rc = dev_copy(&m.field, model->field, n)) != 0) { ofb_destroy(s); return rc; }
  COPY(body_mass, WB); COPY(body_inv_mass, WB); COPY(body_inertia, WB * 9); COPY(body_inv_inertia, WB * 9);
  COPY(body_com, WB * 3);
  COPY(joint_type, WJ); COPY(joint_parent, WJ); COPY(joint_child, WJ); COPY(joint_qd_start, WJ);
  COPY(joint_enabled, WJ); COPY(joint_X_p, WJ * 7); COPY(joint_X_c, WJ * 7); COPY(joint_compliance, WJ);
  COPY(joint_axis, WD * 3); COPY(joint_dof_mode, WD); COPY(joint_target_ke, WD); COPY(joint_target_kd, WD);
  COPY(shape_body, WS); COPY(shape_mu, WS); COPY(shape_friction_axis, WS * 3); COPY(shape_mu_perp, WS);
  COPY(shape_mu_stiction_scale, WS); COPY(shape_v_stribeck, WS); COPY(shape_stribeck_lateral_only, WS);
  COPY(shape_world, (size_t)m.flat_shapes);
#undef COPY
#define ALLOC(ptr, n) if ((rc = dev_alloc(&s->ptr, n)) != 0) { ofb_destroy(s); return rc; }
  const size_t K = L.K;
  ALLOC(ccount, W); ALLOC(cf, (size_t)W * CF * K); ALLOC(ci, (size_t)W * 2 * K);
  ALLOC(lambda, (size_t)W * L.Nc); ALLOC(lambda_prev, (size_t)W * L.Nc); ALLOC(lambda_seed, (size_t)W * L.Nc);
  ALLOC(st_nr, W); ALLOC(st_pcr, W); ALLOC(st_best, W); ALLOC(st_flags, W); ALLOC(st_merit, W);
  ALLOC(topo, topo.size());
  int cap = config->rigid_contact_max > 0 ? config->rigid_contact_max : W * L.K;
  s->rigid_contact_cap = cap;
  ALLOC(pk_keys, cap); ALLOC(pk_members, (size_t)W * K);
  if (config->linesearch) {
    int n = linesearch_steps(config, s->ls_host, 512);
    if (n < 1) { set_error("invalid line-search step counts (at most 512 candidates)"); ofb_destroy(s); return -2; }
    s->ls_count = n;
    ALLOC(ls_steps, (size_t)n);
    if (hipMemcpy(s->ls_steps, s->ls_host, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("hipMemcpy failed (line search)");
      ofb_destroy(s);
      return -4;
    }
  }
  if (config->warm_start) {
    ALLOC(ws_count, W); ALLOC(ws_b0, (size_t)W * K); ALLOC(ws_b1, (size_t)W * K); ALLOC(ws_p, (size_t)W * K * 3);
    ALLOC(ws_n, (size_t)W * K * 3); ALLOC(ws_ln, (size_t)W * K); ALLOC(ws_lt, (size_t)W * K * 2);
    ALLOC(ws_feat, (size_t)W * K * 12); ALLOC(ws_fscale, (size_t)W * K); ALLOC(ws_y, (size_t)W * K * 3);
    ALLOC(ws_off, (size_t)W + 1);
    ALLOC(lambda_seed2, (size_t)W * L.Nc);
    if (hipMemset(s->ws_count, 0, (size_t)W * sizeof(int32_t)) != hipSuccess) {
      set_error("hipMemset failed (warm start)");
      ofb_destroy(s);
      return -4;
    }
  }
#undef ALLOC
  if (hipMemcpy(s->topo, topo.data(), topo.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("topology upload failed");
    ofb_destroy(s);
    return -4;
  }
  s->fused_worlds_per_block = 4;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    s->num_simds = 4 * cus;
  }
  if (hipEventCreate(&s->ev_solve0) != hipSuccess || hipEventCreate(&s->ev_solve1) != hipSuccess ||
      hipEventCreate(&s->ev_kernel0) != hipSuccess || hipEventCreate(&s->ev_kernel1) != hipSuccess) {
    set_error("hipEventCreate failed");
    ofb_destroy(s);
    return -4;
  }
  s->events_ready = 1;
  if (path == OFB_PATH_PLAIN) {
    if ((rc = plain_create(s)) != 0) { ofb_destroy(s); return rc; }
  } else {
    if ((rc = fused_classes_alloc(s)) != 0) { ofb_destroy(s); return rc; }
  }
  *out_solver = s;
  return 0;
}

int ofb_dims(const ofb_solver* s, int32_t* n_c, int32_t* offset_ctrl, int32_t* offset_n, int32_t* offset_f) {
  if (!s) { set_error("null solver"); return -1; }
  if (n_c) *n_c = s->L.Nc;
  if (offset_ctrl) *offset_ctrl = s->L.off_ctrl;
  if (offset_n) *offset_n = s->L.off_n;
  if (offset_f) *offset_f = s->L.off_f;
  return 0;
}

int ofb_step(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl,
             const ofb_contacts* contacts, float dt, void* stream) {
  if (!s || !in || !out || !ctl || !contacts) { set_error("null argument"); return -1; }
  if (!(dt > 0.f)) { set_error("dt must be positive"); return -2; }
  if (!in->body_q || !in->body_qd || !in->body_f || !out->body_q || !out->body_qd) {
    set_error("state pointers must be set");
    return -1;
  }
  if (s->m.D > 0 && (!ctl->joint_target_pos || !ctl->joint_target_vel)) { set_error("control pointers must be set"); return -1; }
  hipStream_t st = (hipStream_t)stream;
  int rc = pack_contacts(s, contacts, st);
  if (rc) return rc;
  if ((rc = reduce_contacts(s, in->body_q, st)) != 0) return rc;  // base_engine.py:317-320
  // the warm starter's seed reaches the solve only through seed_iterate (engine.py:52-59)
  const bool ws = s->cfg.warm_start && s->cfg.seed_iterate;
  // timing events are left out of a step captured into a graph
  const bool timed = !stream_capturing(st);
  s->solve_timed = timed ? 1 : 0;
  s->kernel_timed = 0;
  if (timed) OFB_CHECK(hipEventRecord(s->ev_solve0, st));
  if (ws && (rc = warm_start_apply(s, in, dt, st)) != 0) return rc;
  if (ws && s->ws_net && s->path == OFB_PATH_FUSED) {
    if ((rc = mlp_forward_packed(s->ws_net, s->ws_feat, s->ccount, s->m.W, s->L.K, s->ws_off, s->ws_y, st)) != 0)
      return rc;
    if ((rc = warm_start_learned_seed(s, s->ws_y, st)) != 0) return rc;
  }
  rc = s->path == OFB_PATH_FUSED ? fused_step(s, in, out, ctl, dt, st) : plain_step(s, in, out, ctl, dt, st);
  if (rc) return rc;
  if (ws && (rc = warm_start_snapshot(s, in, dt, st)) != 0) return rc;
  if (timed) OFB_CHECK(hipEventRecord(s->ev_solve1, st));
  return 0;
}

int ofb_warm_start_reset(ofb_solver* s, void* stream) {
  if (!s) { set_error("null solver"); return -1; }
  if (!s->ws_count) return 0;
  OFB_CHECK(hipMemsetAsync(s->ws_count, 0, (size_t)s->m.W * sizeof(int32_t), (hipStream_t)stream));
  return 0;
}

// the snapshot buffers in the order of the opaque block: count [W] i32, body pairs
// [W*K] i32 twice, midpoints and normals [W*K*3], normal impulses [W*K], tangential
// impulses [W*K*2] (f32)
static int ws_parts(const ofb_solver* s, void* (&p)[7], size_t (&b)[7]) {
  const size_t W = s->m.W, WK = W * s->L.K;
  p[0] = s->ws_count; p[1] = s->ws_b0; p[2] = s->ws_b1; p[3] = s->ws_p; p[4] = s->ws_n; p[5] = s->ws_ln;
  p[6] = s->ws_lt;
  b[0] = W * 4; b[1] = WK * 4; b[2] = WK * 4; b[3] = WK * 12; b[4] = WK * 12; b[5] = WK * 4; b[6] = WK * 8;
  return s->ws_count ? 7 : 0;
}

int ofb_warm_start_state_size(const ofb_solver* s, int64_t* bytes) {
  if (!s || !bytes) { set_error("null argument"); return -1; }
  void* p[7];
  size_t b[7];
  const int n = ws_parts(s, p, b);
  int64_t t = 0;
  for (int i = 0; i < n; ++i) t += (int64_t)b[i];
  *bytes = t;
  return 0;
}

int ofb_warm_start_state_get(const ofb_solver* s, void* dst, void* stream) {
  if (!s || !dst) { set_error("null argument"); return -1; }
  void* p[7];
  size_t b[7];
  const int n = ws_parts(s, p, b);
  char* d = static_cast<char*>(dst);
  for (int i = 0; i < n; ++i) {
    OFB_CHECK(hipMemcpyAsync(d, p[i], b[i], hipMemcpyDeviceToDevice, (hipStream_t)stream));
    d += b[i];
  }
  return 0;
}

int ofb_warm_start_state_set(ofb_solver* s, const void* src, void* stream) {
  if (!s || !src) { set_error("null argument"); return -1; }
  void* p[7];
  size_t b[7];
  const int n = ws_parts(s, p, b);
  const char* d = static_cast<const char*>(src);
  for (int i = 0; i < n; ++i) {
    OFB_CHECK(hipMemcpyAsync(p[i], d, b[i], hipMemcpyDeviceToDevice, (hipStream_t)stream));
    d += b[i];
  }
  return 0;
}

int ofb_get_linesearch_steps(const ofb_solver* s, float* steps, int32_t* count) {
  if (!s || !count) { set_error("null argument"); return -1; }
  *count = s->ls_count;
  if (steps) std::memcpy(steps, s->ls_host, (size_t)s->ls_count * sizeof(float));
  return 0;
}

int ofb_set_warm_start_net(ofb_solver* s, const ofb_mlp* net) {
  if (!s) { set_error("null solver"); return -1; }
  if (net && !s->cfg.warm_start) { set_error("the learned warm start needs warm_start enabled"); return -2; }
  if (net && s->path != OFB_PATH_FUSED) { set_error("the learned warm start runs on the fused path"); return -2; }
  s->ws_net = net;
  return 0;
}

int ofb_get_contact_features(const ofb_solver* s, float* features, float* scale, void* stream) {
  if (!s) { set_error("null solver"); return -1; }
  if (!s->ws_feat) { set_error("contact features need warm_start enabled"); return -2; }
  hipStream_t st = (hipStream_t)stream;
  const size_t n = (size_t)s->m.W * s->L.K;
  if (features) OFB_CHECK(hipMemcpyAsync(features, s->ws_feat, n * 12 * 4, hipMemcpyDeviceToDevice, st));
  if (scale) OFB_CHECK(hipMemcpyAsync(scale, s->ws_fscale, n * 4, hipMemcpyDeviceToDevice, st));
  return 0;
}

int ofb_get_lambda_seed(const ofb_solver* s, float* seed_out, void* stream) {
  if (!s || !seed_out) { set_error("null argument"); return -1; }
  OFB_CHECK(hipMemcpyAsync(seed_out, s->lambda_seed, (size_t)s->m.W * s->L.Nc * 4, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return 0;
}

int ofb_last_solve_ms(const ofb_solver* s, float* ms) {
  if (!s || !ms) { set_error("null argument"); return -1; }
  if (!s->solve_timed) { set_error("the last step was not timed (captured into a graph, or none yet)"); return -2; }
  OFB_CHECK(hipEventSynchronize(s->ev_solve1));
  OFB_CHECK(hipEventElapsedTime(ms, s->ev_solve0, s->ev_solve1));
  return 0;
}

int ofb_get_stats(const ofb_solver* s, int32_t* nr_iters, int32_t* pcr_iters, float* merit, int32_t* best_iter,
                  int32_t* flags, void* stream) {
  if (!s) { set_error("null solver"); return -1; }
  hipStream_t st = (hipStream_t)stream;
  const size_t W = s->m.W;
  if (nr_iters) OFB_CHECK(hipMemcpyAsync(nr_iters, s->st_nr, W * 4, hipMemcpyDeviceToDevice, st));
  if (pcr_iters) OFB_CHECK(hipMemcpyAsync(pcr_iters, s->st_pcr, W * 4, hipMemcpyDeviceToDevice, st));
  if (merit) OFB_CHECK(hipMemcpyAsync(merit, s->st_merit, W * 4, hipMemcpyDeviceToDevice, st));
  if (best_iter) OFB_CHECK(hipMemcpyAsync(best_iter, s->st_best, W * 4, hipMemcpyDeviceToDevice, st));
  if (flags) OFB_CHECK(hipMemcpyAsync(flags, s->st_flags, W * 4, hipMemcpyDeviceToDevice, st));
  return 0;
}

int ofb_get_lambda(const ofb_solver* s, float* lambda_out, void* stream) {
  if (!s || !lambda_out) { set_error("null argument"); return -1; }
  OFB_CHECK(hipMemcpyAsync(lambda_out, s->lambda, (size_t)s->m.W * s->L.Nc * 4, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return 0;
}

int ofb_get_lambda_prev(const ofb_solver* s, float* lambda_prev_out, void* stream) {
  if (!s || !lambda_prev_out) { set_error("null argument"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("lambda_prev_iter is recorded by the fused path only"); return -2; }
  if (!s->keep_lambda_prev) {
    set_error("lambda_prev_iter is not kept by this solver (ofb_set_keep_lambda_prev(s, 1) before the step)");
    return -2;
  }
  OFB_CHECK(hipMemcpyAsync(lambda_prev_out, s->lambda_prev, (size_t)s->m.W * s->L.Nc * 4, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return 0;
}

int ofb_set_shape_mu(ofb_solver* s, const float* shape_mu, void* stream) {
  if (!s || !shape_mu) { set_error("null argument"); return -1; }
  OFB_CHECK(hipMemcpyAsync(s->m.shape_mu, shape_mu, (size_t)s->m.W * s->m.SC * 4, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return 0;
}

int ofb_step_backward(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl,
                      const ofb_contacts* contacts, const float* lambda, const float* lambda_prev,
                      const ofb_adjoint_in* gin, const ofb_adjoint_out* gout, float dt, void* stream) {
  if (!s || !in || !out || !ctl || !contacts || !lambda || !lambda_prev || !gin || !gout) {
    set_error("null argument");
    return -1;
  }
  if (s->path != OFB_PATH_FUSED) { set_error("step_backward runs on the fused path"); return -2; }
  if (!(dt > 0.f)) { set_error("dt must be positive"); return -2; }
  if (!in->body_q || !in->body_qd || !in->body_f || !out->body_q || !out->body_qd) {
    set_error("state pointers must be set");
    return -1;
  }
  if (!gin->body_q_grad || !gin->body_qd_grad) { set_error("incoming gradients must be set"); return -1; }
  if (s->m.D > 0 && (!ctl->joint_target_pos || !ctl->joint_target_vel)) { set_error("control pointers must be set"); return -1; }
  if (!(s->cfg.adjoint_pose_eps > 0.f) && s->cfg.adjoint_pose_vjp) {
    set_error("adjoint_pose_eps must be positive");
    return -2;
  }
  hipStream_t st = (hipStream_t)stream;
  int rc = pack_contacts(s, contacts, st);
  if (rc) return rc;
  if ((rc = reduce_contacts(s, in->body_q, st)) != 0) return rc;
  const bool timed = !stream_capturing(st);
  s->solve_timed = timed ? 1 : 0;
  s->kernel_timed = 0;
  if (timed) OFB_CHECK(hipEventRecord(s->ev_solve0, st));
  rc = fused_adjoint(s, in, out, ctl, lambda, lambda_prev, gin, gout, dt, st);
  if (rc) return rc;
  if (timed) OFB_CHECK(hipEventRecord(s->ev_solve1, st));
  return 0;
}

int ofb_last_kernel_ms(const ofb_solver* s, float* ms) {
  if (!s || !ms) { set_error("null argument"); return -1; }
  if (!s->kernel_timed) {
    set_error("the last step launched no timed paired-class kernel (plain path, pairing off, or captured)");
    return -2;
  }
  OFB_CHECK(hipEventSynchronize(s->ev_kernel1));
  OFB_CHECK(hipEventElapsedTime(ms, s->ev_kernel0, s->ev_kernel1));
  return 0;
}

int ofb_set_keep_lambda_prev(ofb_solver* s, int32_t keep) {
  if (!s) { set_error("null solver"); return -1; }
  s->keep_lambda_prev = keep ? 1 : 0;
  return 0;
}

int ofb_set_class_timing(ofb_solver* s, int32_t enabled) {
  if (!s) { set_error("null solver"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("class timing applies to the fused path only"); return -2; }
  s->cls_timing = enabled ? 1 : 0;
  s->cls_timed = 0;
  return 0;
}

int ofb_get_class_times(const ofb_solver* s, int32_t capacity, int32_t* rows_cap, int32_t* worlds, float* ms,
                        int32_t* count) {
  if (!s || !count) { set_error("null argument"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("class timing applies to the fused path only"); return -2; }
  return fused_class_times(s, capacity, rows_cap, worlds, ms, count);
}

int ofb_get_class_shapes(const ofb_solver* s, int32_t capacity, int32_t* slots, int32_t* lanes, int32_t* max_rows,
                         int32_t* count) {
  if (!s || !count) { set_error("null argument"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("row classes apply to the fused path only"); return -2; }
  return fused_class_shapes(s, capacity, slots, lanes, max_rows, count);
}

int ofb_set_fused_launch(ofb_solver* s, int32_t waves_per_block, int32_t pair_small_worlds) {
  if (!s) { set_error("null solver"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("launch settings apply to the fused path only"); return -2; }
  return fused_set_launch(s, waves_per_block, pair_small_worlds);
}

int ofb_set_fused_layout(ofb_solver* s, int32_t large_world_slots) {
  if (!s) { set_error("null solver"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("the large-world layout applies to the fused path only"); return -2; }
  return fused_set_layout(s, large_world_slots);
}

int ofb_get_fused_layout(const ofb_solver* s, int32_t* large_world_slots) {
  if (!s) { set_error("null solver"); return -1; }
  if (large_world_slots) *large_world_slots = s->cfg.large_world_slots;
  return 0;
}

int ofb_set_iteration_history(ofb_solver* s, float* merit, int32_t* pcr_iters, int32_t capacity) {
  if (!s) { set_error("null solver"); return -1; }
  if (s->path != OFB_PATH_FUSED) { set_error("the iteration history is written by the fused path"); return -2; }
  if ((merit == nullptr) != (pcr_iters == nullptr)) { set_error("give both history buffers or neither"); return -1; }
  if (merit && capacity < 1) { set_error("history capacity must be >= 1"); return -2; }
  s->hist_merit = merit;
  s->hist_pcr = pcr_iters;
  s->hist_cap = merit ? capacity : 0;
  return 0;
}

int ofb_get_fused_launch(const ofb_solver* s, int32_t* waves_per_block, int32_t* pair_small_worlds) {
  if (!s) { set_error("null solver"); return -1; }
  if (waves_per_block) *waves_per_block = s->fused_worlds_per_block;
  if (pair_small_worlds) *pair_small_worlds = s->cfg.pair_small_worlds;
  return 0;
}

int ofb_set_lambda_seed(ofb_solver* s, const float* seed, void* stream) {
  if (!s || !seed) { set_error("null argument"); return -1; }
  OFB_CHECK(hipMemcpyAsync(s->lambda_seed, seed, (size_t)s->m.W * s->L.Nc * 4, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return 0;
}

int ofb_get_packed_contacts(const ofb_solver* s, int32_t* count, int32_t* shape0, int32_t* shape1, float* point0,
                            float* point1, float* normal, float* thickness0, float* thickness1, void* stream) {
  if (!s) { set_error("null solver"); return -1; }
  hipStream_t st = (hipStream_t)stream;
  const int W = s->m.W, K = s->L.K;
  if (count) OFB_CHECK(hipMemcpyAsync(count, s->ccount, (size_t)W * 4, hipMemcpyDeviceToDevice, st));
  const int n = W * K, T = 256;
  hipLaunchKernelGGL(unpack_contacts_kernel, dim3((n + T - 1) / T), dim3(T), 0, st, W, K, s->cf, s->ci, shape0,
                     shape1, point0, point1, normal, thickness0, thickness1);
  OFB_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
