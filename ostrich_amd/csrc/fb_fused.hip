// fb_fused.hip -- host side of the fused persistent solver: per-step bucketing of worlds
// into row classes, LDS sizing, and the class launches (kernel in fb_fused_kernel.h).

#include <mutex>
#include <algorithm>

#include "fb_adjoint_kernel.h"

namespace ofb {

// Row classes.  Worlds are bucketed on the device every step by active rows n_r; a class
// is a kernel instance with a row capacity:
//  * code 0: worlds of at most 32 rows, two per wavefront, assembled (dense) Schur matrix
//    in VGPRs (only with pair_small_worlds);
//  * code 1: at most 64 rows, one world per wavefront, dense;
//  * code R = 2..16: at most 64 R rows, one world per wavefront, R row slots per lane,
//    matrix-free operator;
//  * code multiwave_code(NW, R): at most 64 NW R rows, one world per workgroup of NW
//    wavefronts, R = 2 or 4 row slots per lane, matrix-free (Jacobi only).
// A world runs in the first class whose capacity holds it.  Tables:
//  * Jacobi forward: the paired and dense classes, R = 2, then worlds of 2, 4 and 8
//    waves.  Large worlds set a step's duration (a wave's time is its slowest world's),
//    and on several waves each lane walks two rows instead of up to sixteen, without the
//    register spills of the many-slot instances.
//  * block-Jacobi forward: one wavefront per world up to R = 16 (the per-pair Cholesky
//    works inside one wave); the classes from R = 3 on hold few worlds each and are
//    launched as one, in the instance of the largest world the solver can hold.  There
//    are no 8- and 10-row instances: compiled, they returned wrong iterates for worlds of
//    129-640 rows (against the oracle and against the 7-, 12- and 16-row instances on the
//    same worlds), so worlds above 448 rows run in the 12- or 16-row one.
//  * adjoint: one wavefront per world, R = 1..16 (no pairing, no merging).
constexpr int kMaxClasses = 10;
constexpr int kPairRows = 32;
#ifndef OFB_SIDE_STREAMS
#define OFB_SIDE_STREAMS 3
#endif
constexpr int kSideStreams = OFB_SIDE_STREAMS;
static_assert(kSideStreams >= 1 && kSideStreams <= 8, "the solver holds at most 8 side streams");
constexpr int kMW = kMultiWave;

struct ClassTable {
  int n;                   // classes
  int rows[kMaxClasses];   // row capacity of each class
  int code[kMaxClasses];   // kernel instance of each class
  int side[kMaxClasses];   // side stream of each class (-1: the step's stream); class 0 runs last on its stream
  int merge_from;          // classes from this index on are launched as one (n: none)
};
// Streams of the Jacobi classes (measured on 4096 uneven skid-steer worlds and 1024
// tracked-flipper worlds; four streams, as many as a process's hardware queues): the
// dense class, the R = 2 class and the worlds of two waves on side streams of their own;
// the worlds of eight and of four waves, then the paired class, on the step's stream.
// The skid-steer batch has a few worlds of four waves and none of eight, so the paired
// class (the shortest launch) follows the four-wave launch; in the tracked-flipper batch
// the eight-wave worlds (LDS of a whole CU each) start first and take the CUs before the
// other classes fill them.  The dense class ahead of the paired one on the step's stream
// stepped 0.04 ms slower from a graph and 0.15 ms slower eagerly (skid-steer), and the
// eight- and four-wave launches on a side stream 0.5 ms slower (tracked flipper; on
// separate streams, 0.2 ms faster there but 0.05 ms slower from a graph on skid-steer).
#ifndef OFB_JACOBI_SIDES
#define OFB_JACOBI_SIDES -1, 2, 0, 1, -1, -1
#endif
#ifndef OFB_JACOBI_CODES
#define OFB_JACOBI_CODES 0, 1, 2, kMW + 2, kMW + 4, kMW + 8
#endif
static const ClassTable kTabJacobi = {6, {32, 64, 128, 256, 512, 1024}, {OFB_JACOBI_CODES}, {OFB_JACOBI_SIDES}, 6};
// The wide layout of large worlds (large_world_slots = 4): worlds of 129-1024 rows on
// two wavefronts with two, three or four row slots per lane (129-256, 257-384, 385-512
// rows) and 513-1024 on four with four.  A world's solve takes longer than on the
// default layout's 4 or 8 wavefronts, but a batch of many large worlds keeps about twice
// as many of them resident, so it is the faster layout when such worlds outnumber what
// the chip holds at once (contact-rich tracked robots) and the slower one when a few of
// them set the step (the uneven skid-steer batch).  Worlds of 65-128 rows join the
// 129-256-row class (a batch of many large worlds has few of them), which keeps the
// four big classes on four streams: the four-wavefront class first on the step's stream
// (its worlds need the most LDS), the dense class after the 129-256-row one.  Measured on
// 1024 tracked robots: 2.19-2.23 ms per step against 2.33-2.35 ms for six classes with
// the 65-128-row one and 257-512 rows in one four-slot class.
#ifndef OFB_JACOBI_WIDE_SIDES
#define OFB_JACOBI_WIDE_SIDES -1, 0, 0, 2, 1, -1
#endif
#ifndef OFB_JACOBI_WIDE_CODES
#define OFB_JACOBI_WIDE_CODES \
  0, 1, multiwave_code(2, 2), multiwave_code(2, 3), multiwave_code(2, 4), multiwave_code(4, 4)
#endif
#ifndef OFB_JACOBI_WIDE_ROWS
#define OFB_JACOBI_WIDE_ROWS 32, 64, 256, 384, 512, 1024
#endif
#ifndef OFB_JACOBI_WIDE_N
#define OFB_JACOBI_WIDE_N 6
#endif
static const ClassTable kTabJacobiWide = {OFB_JACOBI_WIDE_N, {OFB_JACOBI_WIDE_ROWS}, {OFB_JACOBI_WIDE_CODES},
                                          {OFB_JACOBI_WIDE_SIDES}, OFB_JACOBI_WIDE_N};

static const ClassTable kTabBlockJacobi = {10, {32, 64, 128, 192, 256, 320, 384, 448, 768, 1024},
                                           {0, 1, 2, 3, 4, 5, 6, 7, 12, 16}, {-1, 0, 1, 2, 2, 2, 2, 2, 2, 2}, 3};
static const ClassTable kTabAdjoint = {10, {32, 64, 128, 192, 256, 320, 384, 448, 768, 1024},
                                       {0, 1, 2, 3, 4, 5, 6, 7, 12, 16}, {-1, 0, 0, 0, 0, 0, 0, 0, 0, 0}, 10};

static const ClassTable& fwd_table(const ofb_solver* s) {
  if (s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR) return kTabBlockJacobi;
  return s->cfg.large_world_slots == 4 ? kTabJacobiWide : kTabJacobi;
}

// the capacities passed to the classifier by value
struct ClassCaps {
  int n, first, merge_from;  // classes, first class a world may take (1: no pairing), merge index
  int rows[kMaxClasses];
};
static ClassCaps class_caps(const ClassTable& t, bool pairing) {
  ClassCaps c;
  c.n = t.n;
  c.first = pairing ? 0 : 1;
  c.merge_from = t.merge_from;
  for (int i = 0; i < kMaxClasses; ++i) c.rows[i] = i < t.n ? t.rows[i] : 0;
  return c;
}

__global__ void reset_classes_kernel(int32_t* counts) {
  if (threadIdx.x < kMaxClasses) counts[threadIdx.x] = 0;
}

// One pass over the worlds: each wave ranks its worlds per class with ballots, the
// block sums its waves in LDS, and one global atomic per class and block reserves the
// block's range of each list (one atomic per world on a single counter serialises).
// The order of a class list is not significant: every world is solved on its own.
__global__ void __launch_bounds__(256) classify_worlds_kernel(int W, const int32_t* __restrict__ ccount, int K,
                                                              int base_rows, ClassCaps caps,
                                                              int32_t* __restrict__ counts,
                                                              int32_t* __restrict__ lists) {
  __shared__ int wcnt[4][kMaxClasses];  // worlds per (wave, class)
  __shared__ int bbase[kMaxClasses];    // block's base in each list
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  int c = -1;
  if (w < W) {
    const int n_r = base_rows + 3 * min(ccount[w], K);
    c = caps.first;
    while (c < caps.n - 1 && caps.rows[c] < n_r) ++c;
    if (c > caps.merge_from) c = caps.merge_from;
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  int rank = 0;
  for (int k = 0; k < kMaxClasses; ++k) {
    const unsigned long long m = __ballot(c == k);
    if (c == k) rank = __popcll(m & below);
    if (lane == 0) wcnt[wave][k] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < kMaxClasses) {
    const int k = threadIdx.x;
    int tot = 0;
    for (int v = 0; v < (int)(blockDim.x / WAVE); ++v) tot += wcnt[v][k];
    bbase[k] = tot > 0 ? atomicAdd(counts + k, tot) : 0;
  }
  __syncthreads();
  if (c >= 0) {
    int off = bbase[c];
    for (int v = 0; v < wave; ++v) off += wcnt[v][c];
    lists[(size_t)c * W + off + rank] = w;
  }
}

// classes a world of this solver can fall in (capacities are increasing)
static bool class_used(const ofb_solver* s, const ClassTable& t, int c, bool pairing) {
  if (c >= t.n) return false;
  if (c == 0) return pairing;
  return c == 1 || t.rows[c - 1] < s->L.Nc;
}

// largest class a world of this solver can fall in
static int top_class(const ofb_solver* s, const ClassTable& t) {
  int c = 1;
  while (c < t.n - 1 && t.rows[c] < s->L.Nc) ++c;
  return c;
}

// kernel instance of the forward launch of class list c
static int launch_code(const ofb_solver* s, int c) {
  const ClassTable& t = fwd_table(s);
  return c == t.merge_from ? t.code[std::max(c, top_class(s, t))] : t.code[c];
}
// forward launches: classes past the merge index are merged into it
static bool class_launched(const ofb_solver* s, int c) {
  const ClassTable& t = fwd_table(s);
  return c <= t.merge_from && class_used(s, t, c, s->cfg.pair_small_worlds != 0);
}

// lanes per world, row slots per lane and row capacity of a kernel instance
static int class_lanes(int code) {
  return code == 0 ? kPairRows : code > kMW ? WAVE * ((code - kMW) % kMultiWaveR) : WAVE;
}
static int class_slots(int code) { return code == 0 ? 1 : code > kMW ? 2 + (code - kMW) / kMultiWaveR : code; }
static int class_rows(int code) { return class_lanes(code) * class_slots(code); }

// LDS layout of one world of a class: a world of the class has at most
// (rows_cap - joint and control rows) / 3 contacts
static FusedLayout class_layout(const ofb_solver* s, int code) {
  const bool bj = s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR;
  const int cap = class_rows(code);
  const int base = s->L.Nj + s->L.Nctrl;
  const int cs_cap = std::min(s->L.K, std::max(0, (cap - base) / 3));
  return fused_layout(s->m.B, s->m.J, cap, cs_cap, bj, code <= 1, class_lanes(code));
}
static int lds_words(const ofb_solver* s, int code) { return class_layout(s, code).words; }

// worlds per workgroup of a class: a world of several waves has its workgroup to itself;
// otherwise at most the tuned number of waves' worth and at most kGroupThreads threads
// (the instances' launch bound), a whole number of waves, within 64 KB of LDS per
// workgroup, and among those the choice that keeps the most worlds resident in a CU's
// 160 KB (larger workgroups on ties)
static int choose_wpb(const ofb_solver* s, int code) {
  if (class_lanes(code) > WAVE) return 1;
  const int per_wave = WAVE / class_lanes(code);
  const size_t world_bytes = (size_t)lds_words(s, code) * sizeof(float);
  int best = per_wave, best_res = -1;
  const int max_waves = std::min(s->fused_worlds_per_block, kGroupThreads / WAVE);
  for (int n = 1; n <= max_waves; ++n) {
    const int wpb = n * per_wave;
    const size_t bytes = wpb * world_bytes;
    if (bytes > 65536 && n > 1) break;
    const int res = (int)(163840 / bytes) * wpb;
    if (res >= best_res) { best_res = res; best = wpb; }
  }
  return best;
}

// LDS floats per wave of the adjoint kernel: the forward's bodies and lambda_prev, the
// adjoint scratch, then the block-Jacobi region
static int adj_lds_words(const ofb_solver* s, int R) {
  int base = s->m.B * BS + s->L.Nc + 1 + adj_words(s->m.B, s->m.D, s->m.SC);
  return base + (s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR ? bj_words(WAVE * R) : 0);
}

static int adj_class_op(int R, int op, const AdjParams& P, int blocks, int threads, size_t lds, hipStream_t st,
                        bool bj) {
  int rc = bj ? adjoint_class_op_bj1_set0(R, op, P, blocks, threads, lds, st)
              : adjoint_class_op_bj0_set0(R, op, P, blocks, threads, lds, st);
  if (rc == -1)
    rc = bj ? adjoint_class_op_bj1_set1(R, op, P, blocks, threads, lds, st)
            : adjoint_class_op_bj0_set1(R, op, P, blocks, threads, lds, st);
  return rc;
}

static int class_op(int code, int op, const FusedParams& P, int blocks, int threads, size_t lds, hipStream_t st,
                    bool bj) {
  int rc = bj ? fused_class_op_bj1_set0(code, op, P, blocks, threads, lds, st)
              : fused_class_op_bj0_set0(code, op, P, blocks, threads, lds, st);
  if (rc == -1)
    rc = bj ? fused_class_op_bj1_set1(code, op, P, blocks, threads, lds, st)
            : fused_class_op_bj0_set1(code, op, P, blocks, threads, lds, st);
  if (rc == -1 && !bj) rc = fused_class_op_bj0_set2(code, op, P, blocks, threads, lds, st);
  return rc;
}

static int launch_fused(int code, FusedParams P, const ofb_solver* s, int cls, hipStream_t stream) {
  P.wlist = s->cls_lists + (size_t)cls * P.m.W;
  P.wcount = s->cls_counts + cls;
  P.rows_cap = class_rows(code);
  P.lay = class_layout(s, code);
  const int wpb = s->cls_wpb[cls];
  const int per_block = class_lanes(code) > WAVE ? class_lanes(code) : wpb * class_lanes(code);
  if (per_block > (class_lanes(code) > WAVE ? class_lanes(code) : kGroupThreads)) {
    set_error("fused launch of %d threads per workgroup exceeds the instance's bound", per_block);
    return -4;
  }
  size_t lds_bytes = (size_t)wpb * P.lay.words * sizeof(float);
  int blocks = (P.m.W + wpb - 1) / wpb;
  return class_op(code, 0, P, blocks, per_block, lds_bytes, stream,
                  s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR);
}

// lets a forward class instance use more than 64 KB of LDS when its smallest workgroup
// (one wave, or one world of several waves) needs it
static int allow_lds_fwd(int code, const ofb_solver* s) {
  const bool bj = s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR;
  const int per_wave = class_lanes(code) > WAVE ? 1 : WAVE / class_lanes(code);
  const size_t bytes = (size_t)lds_words(s, code) * sizeof(float) * per_wave;
  if (bytes > 163840) {
    set_error("fused kernel needs %zu bytes of LDS per workgroup (at most 160 KB)", bytes);
    return -4;
  }
  if (bytes > 65536) {
    FusedParams P{};
    if (class_op(code, 1, P, 0, 0, bytes, nullptr, bj) != 0) {
      set_error("fused kernel needs %zu bytes of LDS per workgroup", bytes);
      return -4;
    }
  }
  return 0;
}

// the same for the adjoint instance with R row slots per lane
static int allow_lds_adj(int R, const ofb_solver* s) {
  const bool bj = s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR;
  const size_t bytes = (size_t)(up4(adj_lds_words(s, R)) + rs_words(WAVE, s->m.B)) * sizeof(float);
  if (bytes > 65536) {
    AdjParams A{};
    if (adj_class_op(R, 1, A, 0, 0, bytes, nullptr, bj) != 0) {
      set_error("adjoint kernel needs %zu bytes of LDS per world", bytes);
      return -4;
    }
  }
  return 0;
}

// workgroup shape of every class launch for the solver's current launch settings
static int fused_configure(ofb_solver* s) {
  for (int c = 0; c < kMaxClasses; ++c) {
    s->cls_wpb[c] = 1;
    if (!class_launched(s, c)) continue;
    int rc = allow_lds_fwd(launch_code(s, c), s);
    if (rc) return rc;
    s->cls_wpb[c] = choose_wpb(s, launch_code(s, c));
  }
  for (int c = 1; c < kTabAdjoint.n; ++c) {
    if (!class_used(s, kTabAdjoint, c, false)) continue;
    int rc = allow_lds_adj(kTabAdjoint.code[c], s);
    if (rc) return rc;
  }
  return 0;
}

// launch settings of the fused path: the largest number of waves per workgroup and whether
// worlds of at most 32 rows run two per wavefront.  The workgroup size changes nothing
// in a world's arithmetic; pairing changes the lane-group width of the reductions (a
// rounding-level difference).
int fused_set_launch(ofb_solver* s, int waves_per_block, int pair_small_worlds) {
  if (waves_per_block < 1 || waves_per_block > 16) {
    set_error("waves per workgroup must be in [1, 16] (got %d)", waves_per_block);
    return -2;
  }
  s->fused_worlds_per_block = waves_per_block;
  s->cfg.pair_small_worlds = pair_small_worlds ? 1 : 0;
  return fused_configure(s);
}

// Row slots per lane of the Jacobi worlds of more than 256 rows (2: the default layout,
// 4: the wide one); like pairing, a rounding-level difference (other reduction widths).
int fused_set_layout(ofb_solver* s, int large_world_slots) {
  if (large_world_slots != 2 && large_world_slots != 4) {
    set_error("large_world_slots must be 2 or 4 (got %d)", large_world_slots);
    return -2;
  }
  s->cfg.large_world_slots = large_world_slots;
  return fused_configure(s);
}

// The side streams are shared by every solver of the process on a device: created with
// the first solver there and kept until the process ends.  A process spreads its streams
// over a few hardware queues (GPU_MAX_HW_QUEUES, 4 by default); with streams of their
// own, the solvers created later in a process (a scratch solver of the autotuner, one
// solver per workload) found their side streams on the queues of other streams, and the
// classes of a step then ran one after another.  Solvers that step at the same time from
// different host threads queue their classes behind one another on the shared streams;
// each step orders its own launches with its own events, so results are unaffected.
constexpr int kMaxDevices = 64;
static std::mutex g_side_mu;
static hipStream_t g_side[kMaxDevices][kSideStreams];
static bool g_side_made[kMaxDevices];

static int shared_side_streams(hipStream_t (&out)[8]) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("no current HIP device for the side streams");
    return -4;
  }
  std::lock_guard<std::mutex> lock(g_side_mu);
  if (!g_side_made[dev]) {
    for (int i = 0; i < kSideStreams; ++i) {
      if (hipStreamCreateWithFlags(&g_side[dev][i], hipStreamNonBlocking) != hipSuccess) {
        for (int j = 0; j < i; ++j) (void)hipStreamDestroy(g_side[dev][j]);
        set_error("hipStreamCreate failed");
        return -4;
      }
    }
    g_side_made[dev] = true;
  }
  for (int i = 0; i < kSideStreams; ++i) out[i] = g_side[dev][i];
  return 0;
}

int fused_classes_alloc(ofb_solver* s) {
  int rc = fused_configure(s);
  if (rc) return rc;
  if (hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess) {
    set_error("hipEventCreate failed");
    return -4;
  }
  s->side_ready = 1;
  if ((rc = shared_side_streams(s->side_stream)) != 0) return rc;
  for (int i = 0; i < kSideStreams; ++i) {
    s->side_ready = 2 + 2 * i;
    if (hipEventCreateWithFlags(&s->ev_join[i], hipEventDisableTiming) != hipSuccess) {
      set_error("hipEventCreate failed");
      return -4;
    }
    s->side_ready = 3 + 2 * i;
  }
  for (int c = 0; c < kMaxClasses; ++c) {
    if (hipEventCreate(&s->ev_cls0[c]) != hipSuccess || hipEventCreate(&s->ev_cls1[c]) != hipSuccess) {
      set_error("hipEventCreate failed (class timing)");
      return -4;
    }
    s->cls_events_ready = c + 1;
  }
  if (hipMalloc((void**)&s->cls_counts, kMaxClasses * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&s->cls_lists, (size_t)kMaxClasses * s->m.W * sizeof(int32_t)) != hipSuccess) {
    set_error("hipMalloc failed (world classes)");
    return -4;
  }
  return 0;
}

void fused_classes_free(ofb_solver* s) {
  // side_ready counts what fused_classes_alloc set up: the fork event, then a (shared,
  // kept) stream and its join event per side stream
  if (s->side_ready >= 1) (void)hipEventDestroy(s->ev_fork);
  for (int i = 0; i < kSideStreams; ++i) {
    if (s->side_ready >= 3 + 2 * i) (void)hipEventDestroy(s->ev_join[i]);
  }
  s->side_ready = 0;
  for (int c = 0; c < s->cls_events_ready; ++c) {
    (void)hipEventDestroy(s->ev_cls0[c]);
    (void)hipEventDestroy(s->ev_cls1[c]);
  }
  s->cls_events_ready = 0;
  if (s->cls_counts) (void)hipFree(s->cls_counts);
  if (s->cls_lists) (void)hipFree(s->cls_lists);
  s->cls_counts = nullptr;
  s->cls_lists = nullptr;
}

// Row capacity, world count and duration of every forward class of the last step (eager
// steps with class timing on; ms = -1 for a class that was not launched or not timed).
int fused_class_times(const ofb_solver* s, int capacity, int32_t* rows_cap, int32_t* worlds, float* ms,
                      int32_t* count) {
  const ClassTable& t = fwd_table(s);
  *count = t.n;
  if (capacity < kMaxClasses) { set_error("class times need room for %d classes", kMaxClasses); return -2; }
  int32_t n[kMaxClasses];
  if (hipMemcpy(n, s->cls_counts, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("class count read-back failed");
    return -4;
  }
  for (int c = 0; c < t.n; ++c) {
    if (rows_cap) rows_cap[c] = class_rows(launch_code(s, c));
    if (worlds) worlds[c] = n[c];
    if (!ms) continue;
    ms[c] = -1.f;
    if (s->cls_timed & (1 << c)) {
      hipEvent_t e0 = c == 0 ? s->ev_kernel0 : s->ev_cls0[c], e1 = c == 0 ? s->ev_kernel1 : s->ev_cls1[c];
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms + c, e0, e1) != hipSuccess) {
        set_error("class event read failed");
        return -4;
      }
    }
  }
  return 0;
}

// Kernel shape of every forward class: row slots per lane and lanes per world of the
// instance that class's launch runs, and the largest row count of a world the class list
// holds (0: no world can fall in it, or it is merged into an earlier launch).  A world
// of n_r rows is in the first class whose largest row count is >= n_r.
int fused_class_shapes(const ofb_solver* s, int capacity, int32_t* slots, int32_t* lanes, int32_t* max_rows,
                       int32_t* count) {
  const ClassTable& t = fwd_table(s);
  *count = t.n;
  if (capacity < kMaxClasses) { set_error("class shapes need room for %d classes", kMaxClasses); return -2; }
  for (int c = 0; c < t.n; ++c) {
    const int code = launch_code(s, c);
    if (slots) slots[c] = class_slots(code);
    if (lanes) lanes[c] = class_lanes(code);
    if (max_rows) max_rows[c] = !class_launched(s, c) ? 0 : c == t.merge_from ? s->L.Nc : t.rows[c];
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
  P.lambda_prev_out = s->keep_lambda_prev ? s->lambda_prev : nullptr;
  P.lambda_seed2 = (s->cfg.warm_start && s->ws_net) ? s->lambda_seed2 : nullptr;
  P.ls_steps = s->ls_steps;
  P.ls_count = s->cfg.linesearch ? s->ls_count : 0;
  P.st_nr = s->st_nr; P.st_pcr = s->st_pcr; P.st_best = s->st_best; P.st_flags = s->st_flags; P.st_merit = s->st_merit;
  P.hist_merit = s->hist_merit; P.hist_pcr = s->hist_pcr; P.hist_cap = s->hist_cap;
  const int W = s->m.W;
  const ClassTable& t = fwd_table(s);
  if (s->L.Nc > t.rows[t.n - 1]) {
    set_error("fused path supports at most %d constraint rows per world (got %d)", t.rows[t.n - 1], s->L.Nc);
    return -3;
  }
  hipLaunchKernelGGL(reset_classes_kernel, dim3(1), dim3(64), 0, stream, s->cls_counts);
  hipLaunchKernelGGL(classify_worlds_kernel, dim3((W + 255) / 256), dim3(256), 0, stream, W, s->ccount, s->L.K,
                     s->L.Nj + s->L.Nctrl, class_caps(t, s->cfg.pair_small_worlds != 0), s->cls_counts,
                     s->cls_lists);
  // The one-world-per-wave classes (and the worlds of several waves) run on side
  // streams, launched first so that their waves are resident beside the paired class's,
  // the largest worlds first; the table spreads the classes over the side streams so
  // that the large classes do not wait for one another to drain.
  if (hipEventRecord(s->ev_fork, stream) != hipSuccess) {
    set_error("fused step: stream fork failed");
    return -4;
  }
  for (int i = 0; i < kSideStreams; ++i) {
    if (hipStreamWaitEvent(s->side_stream[i], s->ev_fork, 0) != hipSuccess) {
      set_error("fused step: stream fork failed");
      return -4;
    }
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cap) != hipSuccess) cap = hipStreamCaptureStatusNone;
  const bool cls_timed = s->cls_timing && cap == hipStreamCaptureStatusNone;
  s->cls_timed = 0;
  for (int c = std::min(t.merge_from, t.n - 1); c >= 1; --c) {  // the slowest worlds first
    if (!class_launched(s, c)) continue;  // no world can fall in this class
    hipStream_t side = t.side[c] < 0 ? stream : s->side_stream[t.side[c]];
    if (cls_timed && hipEventRecord(s->ev_cls0[c], side) != hipSuccess) {
      set_error("fused step: event record failed");
      return -4;
    }
    if (const int rc = launch_fused(launch_code(s, c), P, s, c, side)) {
      if (rc == -1) set_error("no fused kernel instance for class code %d", launch_code(s, c));
      return -4;
    }
    if (cls_timed) {
      if (hipEventRecord(s->ev_cls1[c], side) != hipSuccess) {
        set_error("fused step: event record failed");
        return -4;
      }
      s->cls_timed |= 1 << c;
    }
  }
  // the paired class, which holds most worlds of a typical batch, last on its stream,
  // bracketed by the kernel events (not recorded while the stream is being captured)
  hipStream_t st0 = t.side[0] < 0 ? stream : s->side_stream[t.side[0]];
  const bool timed = cap == hipStreamCaptureStatusNone && class_launched(s, 0);
  if (timed && hipEventRecord(s->ev_kernel0, st0) != hipSuccess) {
    set_error("fused step: event record failed");
    return -4;
  }
  if (class_launched(s, 0)) {
    if (const int rc = launch_fused(t.code[0], P, s, 0, st0)) {
      if (rc == -1) set_error("no fused kernel instance for class code %d", t.code[0]);
      return -4;
    }
  }
  if (timed && hipEventRecord(s->ev_kernel1, st0) != hipSuccess) {
    set_error("fused step: event record failed");
    return -4;
  }
  s->kernel_timed = timed ? 1 : 0;
  if (timed && cls_timed) s->cls_timed |= 1;  // the paired class's events are the kernel events
  for (int i = 0; i < kSideStreams; ++i) {
    if (hipEventRecord(s->ev_join[i], s->side_stream[i]) != hipSuccess) {
      set_error("fused step: stream join failed");
      return -4;
    }
  }
  for (int i = 0; i < kSideStreams; ++i) {
    if (hipStreamWaitEvent(stream, s->ev_join[i], 0) != hipSuccess) {
      set_error("fused step: stream join failed");
      return -4;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("fused kernel launch failed: %s", hipGetErrorString(e)); return -4; }
  return 0;
}

// waves per world of the adjoint's pose central differences (at most 4: 256 threads)
#ifndef OFB_ADJ_MAX_SPLIT
#define OFB_ADJ_MAX_SPLIT 4
#endif

int fused_adjoint(ofb_solver* s, const ofb_state_in* in, const ofb_state_out* out, const ofb_control* ctl,
                  const float* lambda, const float* lambda_prev, const ofb_adjoint_in* gin,
                  const ofb_adjoint_out* gout, float dt, hipStream_t stream) {
  AdjParams P{};
  P.m = s->m;
  P.L = s->L;
  P.cfg = s->cfg;
  P.dt = dt;
  P.target_pos = ctl->joint_target_pos; P.target_vel = ctl->joint_target_vel;
  P.ccount = s->ccount; P.cf = s->cf; P.ci = s->ci; P.topo = s->topo;
  P.q_prev = in->body_q; P.qd_prev = in->body_qd; P.body_f = in->body_f;
  P.q = out->body_q; P.qd = out->body_qd;
  P.lam = lambda; P.lam_prev = lambda_prev;
  P.g_q = gin->body_q_grad; P.g_qd = gin->body_qd_grad;
  P.g_q_prev = gout->body_q_grad; P.g_qd_prev = gout->body_qd_grad; P.g_f = gout->body_f_grad;
  P.g_tpos = gout->joint_target_pos_grad; P.g_tvel = gout->joint_target_vel_grad;
  P.g_ke = gout->joint_target_ke_grad; P.g_kd = gout->joint_target_kd_grad;
  P.g_mu = gout->shape_mu_grad; P.w_lam = gout->adjoint_lambda; P.w_u = gout->adjoint_body;
  P.st_pcr = s->st_pcr; P.st_flags = s->st_flags;
  const int W = s->m.W;
  const ClassTable& t = kTabAdjoint;
  if (s->L.Nc > t.rows[t.n - 1]) {
    set_error("fused path supports at most %d constraint rows per world (got %d)", t.rows[t.n - 1], s->L.Nc);
    return -3;
  }
  const bool bj = s->cfg.preconditioner == OFB_PRECOND_PER_BODY_PAIR;
  // the adjoint runs one world per wavefront: no world is put in the paired class
  hipLaunchKernelGGL(reset_classes_kernel, dim3(1), dim3(64), 0, stream, s->cls_counts);
  hipLaunchKernelGGL(classify_worlds_kernel, dim3((W + 255) / 256), dim3(256), 0, stream, W, s->ccount, s->L.K,
                     s->L.Nj + s->L.Nctrl, class_caps(t, false), s->cls_counts, s->cls_lists);
  for (int c = 1; c < t.n; ++c) {
    if (!class_used(s, t, c, false)) continue;
    const int R = t.code[c];
    P.wlist = s->cls_lists + (size_t)c * W;
    P.wcount = s->cls_counts + c;
    P.rows_cap = WAVE * R;
    // [forward's bodies, lambda_prev, adjoint scratch, block-Jacobi | lane-private body
    // sums | transfer area | row-local difference table]
    // (each wave's region and its body sums start on 16-byte boundaries: the sums are
    // read with 128-bit loads)
    P.rs_off = rs_enabled(s->m.B) ? up4(adj_lds_words(s, R)) : -1;
    P.lds_per_wave = up4(adj_lds_words(s, R)) + rs_words(WAVE, s->m.B);
    P.xfer_off = P.lds_per_wave;
    P.split = 1;
    P.fd_off = -1;
    int wpb = s->fused_worlds_per_block;
    if (s->cfg.adjoint_pose_vjp) {
      // Row-local central differences (worlds of at most 64 rows) need a table of
      // 30 words per row slot after the transfer area.
      const int base = P.lds_per_wave;
      const int fd_words = adj_fd_words(P.rows_cap, s->m.B);
      const int words_fd = up4(base + 2 * P.rows_cap + fd_words);
      const bool row_fd = R == 1 && (size_t)words_fd * sizeof(float) <= 65536;
      const int words = row_fd ? words_fd : up4(base + 2 * P.rows_cap);
      // Fewer worlds than resident wave slots (the kernel's registers allow two waves per
      // SIMD): give each world `split` waves for the pose central differences (one world
      // per workgroup, each wave its own LDS copy).  1024 worlds on 1024 SIMDs: split 2.
      int split = 1;
      while (split < OFB_ADJ_MAX_SPLIT && (size_t)W * split * 2 <= (size_t)2 * s->num_simds) split *= 2;
      while (split > 1 && (size_t)split * words * sizeof(float) > 65536) split /= 2;
      if (row_fd) {
        P.fd_off = base + 2 * P.rows_cap;
        P.lds_per_wave = words;
      }
      if (split > 1) {
        P.split = split;
        P.lds_per_wave = words;
        wpb = 1;
      }
    }
    while (wpb > 1 && (size_t)wpb * P.lds_per_wave * sizeof(float) > 65536) --wpb;
    const int waves = wpb * P.split;
    size_t lds_bytes = (size_t)waves * P.lds_per_wave * sizeof(float);
    if (adj_class_op(R, 0, P, (W + wpb - 1) / wpb, waves * WAVE, lds_bytes, stream, bj) != 0) {
      set_error("no adjoint kernel instance for %d row slots per lane", R);
      return -4;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("adjoint kernel launch failed: %s", hipGetErrorString(e)); return -4; }
  return 0;
}

}  // namespace ofb
