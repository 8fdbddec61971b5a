// fb_mlp.hip -- the per-contact warm-start network on gfx950 matrix cores.
//
// 12 -> 64 -> 64 -> 3 with ReLU, BF16 weights and activations, FP32 accumulation on
// v_mfma_f32_16x16x16_bf16 (one wavefront computes a 16x16 output tile per instruction
// per 16-wide K step).  Inputs are padded to 16 features and outputs to 16 columns.  One
// workgroup is one wavefront and owns a block of 64 rows (contacts): X, H1, H2 and the
// backward's dZ tiles are staged in LDS as BF16 so every GEMM - forward and backward,
// including the transposed ones - is the same LDS-tile MMA with different strides.
//
// Training: the loss is the mean over rows of sum_c (y_c - t_c)^2 (c < 3); every block
// writes its partial weight gradient to its own slot, a fixed-order reduction sums the
// slots (so training is deterministic), and Adam updates FP32 master weights and
// re-rounds the BF16 copies.  The learned warm start has no counterpart in the
// reference; the network's arithmetic is restated in numpy by the oracle (oracle/mlp.py).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstring>
#include <mutex>
#include <vector>

#include "fb_solver.h"

namespace ofb {

namespace mlp {

constexpr int IN = 12, INP = 16, H = 64, OUT = 3, OUTP = 16, ROWS = 64;
// parameter layout (floats): W1 [INP x H], b1 [H], W2 [H x H], b2 [H], W3 [H x OUTP], b3 [OUTP]
constexpr int OW1 = 0, OB1 = OW1 + INP * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * OUTP;
constexpr int NPARAM = OB3 + OUTP;
// LDS row strides (bf16 elements): padded by 8 to spread banks
constexpr int LX = INP + 8, LH = H + 8, LY = OUTP + 8;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One 16x16 tile of C = A B over K (multiple of 16): A(i,k) = A[i*sai + k*sak],
// B(k,j) = B[k*sbk + j*sbj].  Lane l supplies A row l%16 and B column l%16 for
// k = k0 + 4(l/16) .. +3 and receives C[4(l/16) + t][l%16] in acc[t].
__device__ __forceinline__ f4 tile_mma(const uint16_t* A, int sai, int sak, const uint16_t* B, int sbk, int sbj,
                                       int K, int lane) {
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const int r = lane & 15, q = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    s4 a, b;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 + 4 * q + t;
      a[t] = (short)A[r * sai + k * sak];
      b[t] = (short)B[k * sbk + r * sbj];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

struct Net {
  float* w;       // [NPARAM] FP32 master
  uint16_t* wb;   // [NPARAM] BF16 copy (biases read from w)
  float* m;       // Adam moments
  float* v;
  float* gpart;   // [blocks][NPARAM] per-block gradient slots
  float* grad;    // [NPARAM]
  float* lpart;   // [blocks] per-block loss
  int gcap;       // slot capacity in blocks
  int step;
};

// forward of one 64-row block into LDS: X (bf16) -> H1, H2 (bf16, post-ReLU); Y fp32 out
__device__ void block_forward(const uint16_t* __restrict__ wb, const float* __restrict__ w, uint16_t* sX,
                              uint16_t* sH1, uint16_t* sH2, float* sY, int lane) {
  const int r = lane & 15, q = lane >> 4;
  for (int ti = 0; ti < ROWS / 16; ++ti)
    for (int tj = 0; tj < H / 16; ++tj) {
      f4 acc = tile_mma(sX + ti * 16 * LX, LX, 1, wb + OW1 + tj * 16, H, 1, INP, lane);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = ti * 16 + 4 * q + t, j = tj * 16 + r;
        sH1[i * LH + j] = to_bf16(fmaxf(acc[t] + w[OB1 + j], 0.f));
      }
    }
  __syncthreads();
  for (int ti = 0; ti < ROWS / 16; ++ti)
    for (int tj = 0; tj < H / 16; ++tj) {
      f4 acc = tile_mma(sH1 + ti * 16 * LH, LH, 1, wb + OW2 + tj * 16, H, 1, H, lane);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = ti * 16 + 4 * q + t, j = tj * 16 + r;
        sH2[i * LH + j] = to_bf16(fmaxf(acc[t] + w[OB2 + j], 0.f));
      }
    }
  __syncthreads();
  for (int ti = 0; ti < ROWS / 16; ++ti) {
    f4 acc = tile_mma(sH2 + ti * 16 * LH, LH, 1, wb + OW3, OUTP, 1, H, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = ti * 16 + 4 * q + t, j = r;
      sY[i * LY + j] = acc[t] + w[OB3 + j];
    }
  }
  __syncthreads();
}

__device__ void load_x(const float* __restrict__ x, int n, int row0, uint16_t* sX, int lane) {
  for (int e = lane; e < ROWS * INP; e += 64) {
    const int i = e / INP, k = e % INP, g = row0 + i;
    sX[i * LX + k] = (g < n && k < IN) ? to_bf16(x[(size_t)g * IN + k]) : (uint16_t)0;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(64) mlp_forward_kernel(const float* __restrict__ w, const uint16_t* __restrict__ wb,
                                                         const float* __restrict__ x, int n, float* __restrict__ y) {
  __shared__ uint16_t sX[ROWS * LX], sH1[ROWS * LH], sH2[ROWS * LH];
  __shared__ float sY[ROWS * LY];
  const int lane = threadIdx.x, row0 = blockIdx.x * ROWS;
  load_x(x, n, row0, sX, lane);
  block_forward(wb, w, sX, sH1, sH2, sY, lane);
  for (int e = lane; e < ROWS * OUT; e += 64) {
    const int i = e / OUT, c = e % OUT, g = row0 + i;
    if (g < n) y[(size_t)g * OUT + c] = sY[i * LY + c];
  }
}

// Forward over the valid contact slots only: dense row g of the packed list is slot c of
// world w, where off[w] <= g < off[w + 1] (off = exclusive scan of the worlds' contact
// counts, off[W] = total); x and y stay in the slot layout [W * K] rows.  Every row's
// result depends on that row alone, so it equals the forward over all W * K slots.
__global__ void __launch_bounds__(64) mlp_forward_packed_kernel(const float* __restrict__ w,
                                                                const uint16_t* __restrict__ wb,
                                                                const float* __restrict__ x,
                                                                const int32_t* __restrict__ off, int W, int K,
                                                                float* __restrict__ y) {
  __shared__ uint16_t sX[ROWS * LX], sH1[ROWS * LH], sH2[ROWS * LH];
  __shared__ float sY[ROWS * LY];
  __shared__ int sRow[ROWS];
  const int lane = threadIdx.x, row0 = blockIdx.x * ROWS;
  const int n = off[W];
  if (row0 >= n) return;  // uniform over the block
  {
    const int g = row0 + lane;  // ROWS == 64 == blockDim.x: one packed row per lane
    int r = -1;
    if (g < n) {
      int lo = 0, hi = W;  // largest world with off[world] <= g
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
      }
      r = lo * K + (g - off[lo]);
    }
    sRow[lane] = r;
  }
  __syncthreads();
  for (int e = lane; e < ROWS * INP; e += 64) {
    const int i = e / INP, k = e % INP, r = sRow[i];
    sX[i * LX + k] = (r >= 0 && k < IN) ? to_bf16(x[(size_t)r * IN + k]) : (uint16_t)0;
  }
  __syncthreads();
  block_forward(wb, w, sX, sH1, sH2, sY, lane);
  for (int e = lane; e < ROWS * OUT; e += 64) {
    const int i = e / OUT, c = e % OUT, r = sRow[i];
    if (r >= 0) y[(size_t)r * OUT + c] = sY[i * LY + c];
  }
}

// exclusive scan of min(count[w], K) over the worlds (one workgroup); off[W] = total
__global__ void __launch_bounds__(1024) mlp_offsets_kernel(int W, int K, const int32_t* __restrict__ count,
                                                           int32_t* __restrict__ off) {
  __shared__ int part[1024];
  const int t = threadIdx.x, per = (W + 1023) / 1024;
  int s = 0;
  for (int i = t * per; i < min(W, (t + 1) * per); ++i) s += min(count[i], K);
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int i = t * per; i < min(W, (t + 1) * per); ++i) {
    off[i] = run;
    run += min(count[i], K);
  }
  if (t == 1023) off[W] = part[1023];
}

// forward + backward of one block; writes the block's gradient slot and loss
__global__ void __launch_bounds__(64) mlp_grad_kernel(const float* __restrict__ w, const uint16_t* __restrict__ wb,
                                                      const float* __restrict__ x, const float* __restrict__ tgt,
                                                      int n, float inv_n, float* __restrict__ gpart,
                                                      float* __restrict__ lpart) {
  __shared__ uint16_t sX[ROWS * LX], sH1[ROWS * LH], sH2[ROWS * LH], sD2[ROWS * LH], sD1[ROWS * LH];
  __shared__ uint16_t sDY[ROWS * LY];
  __shared__ float sY[ROWS * LY];
  const int lane = threadIdx.x, row0 = blockIdx.x * ROWS;
  const int r = lane & 15, q = lane >> 4;
  float* gp = gpart + (size_t)blockIdx.x * NPARAM;
  load_x(x, n, row0, sX, lane);
  block_forward(wb, w, sX, sH1, sH2, sY, lane);
  // dY = 2 (y - t) / n on the three outputs of valid rows; loss partial
  float lsum = 0.f;
  for (int e = lane; e < ROWS * OUTP; e += 64) {
    const int i = e / OUTP, c = e % OUTP, g = row0 + i;
    float d = 0.f;
    if (g < n && c < OUT) {
      float diff = sY[i * LY + c] - tgt[(size_t)g * OUT + c];
      lsum += diff * diff;
      d = 2.f * diff * inv_n;
    }
    sDY[i * LY + c] = to_bf16(d);
  }
  for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (lane == 0) lpart[blockIdx.x] = lsum * inv_n;
  __syncthreads();
  // dW3 [H x OUTP] = H2^T dY ; db3 = sum_rows dY
  for (int ti = 0; ti < H / 16; ++ti) {
    f4 acc = tile_mma(sH2 + ti * 16, 1, LH, sDY, LY, 1, ROWS, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) gp[OW3 + (ti * 16 + 4 * q + t) * OUTP + r] = acc[t];
  }
  for (int c = lane; c < OUTP; c += 64) {
    float s = 0.f;
    for (int i = 0; i < ROWS; ++i) s += from_bf16(sDY[i * LY + c]);
    gp[OB3 + c] = s;
  }
  // dZ2 = (dY W3^T) * [H2 > 0]
  for (int ti = 0; ti < ROWS / 16; ++ti)
    for (int tj = 0; tj < H / 16; ++tj) {
      f4 acc = tile_mma(sDY + ti * 16 * LY, LY, 1, wb + OW3 + tj * 16 * OUTP, 1, OUTP, OUTP, lane);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = ti * 16 + 4 * q + t, j = tj * 16 + r;
        sD2[i * LH + j] = sH2[i * LH + j] != 0 ? to_bf16(acc[t]) : (uint16_t)0;
      }
    }
  __syncthreads();
  // dW2 = H1^T dZ2 ; db2
  for (int ti = 0; ti < H / 16; ++ti)
    for (int tj = 0; tj < H / 16; ++tj) {
      f4 acc = tile_mma(sH1 + ti * 16, 1, LH, sD2 + tj * 16, LH, 1, ROWS, lane);
#pragma unroll
      for (int t = 0; t < 4; ++t) gp[OW2 + (ti * 16 + 4 * q + t) * H + tj * 16 + r] = acc[t];
    }
  for (int c = lane; c < H; c += 64) {
    float s = 0.f;
    for (int i = 0; i < ROWS; ++i) s += from_bf16(sD2[i * LH + c]);
    gp[OB2 + c] = s;
  }
  // dZ1 = (dZ2 W2^T) * [H1 > 0]
  for (int ti = 0; ti < ROWS / 16; ++ti)
    for (int tj = 0; tj < H / 16; ++tj) {
      f4 acc = tile_mma(sD2 + ti * 16 * LH, LH, 1, wb + OW2 + tj * 16 * H, 1, H, H, lane);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = ti * 16 + 4 * q + t, j = tj * 16 + r;
        sD1[i * LH + j] = sH1[i * LH + j] != 0 ? to_bf16(acc[t]) : (uint16_t)0;
      }
    }
  __syncthreads();
  // dW1 [INP x H] = X^T dZ1 ; db1
  for (int tj = 0; tj < H / 16; ++tj) {
    f4 acc = tile_mma(sX, 1, LX, sD1 + tj * 16, LH, 1, ROWS, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) gp[OW1 + (4 * q + t) * H + tj * 16 + r] = acc[t];
  }
  for (int c = lane; c < H; c += 64) {
    float s = 0.f;
    for (int i = 0; i < ROWS; ++i) s += from_bf16(sD1[i * LH + c]);
    gp[OB1 + c] = s;
  }
}

// fixed-order sum of the block slots; the padded input rows / output columns stay zero
__global__ void mlp_reduce_kernel(const float* __restrict__ gpart, int blocks, float* __restrict__ grad,
                                  const float* __restrict__ lpart, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) {
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += gpart[(size_t)b * NPARAM + p];
    bool pad = (p >= OW1 && p < OB1 && (p - OW1) / H >= IN) || (p >= OW3 && p < OB3 && (p - OW3) % OUTP >= OUT) ||
               (p >= OB3 && p - OB3 >= OUT);
    grad[p] = pad ? 0.f : s;
  }
  if (p == 0 && loss) {
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += lpart[b];
    *loss = s;
  }
}

__global__ void mlp_adam_kernel(float* __restrict__ w, uint16_t* __restrict__ wb, float* __restrict__ m,
                                float* __restrict__ v, const float* __restrict__ g, float lr, float b1, float b2,
                                float eps, float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float gg = g[p];
  float mm = b1 * m[p] + (1.f - b1) * gg;
  float vv = b2 * v[p] + (1.f - b2) * gg * gg;
  m[p] = mm;
  v[p] = vv;
  float nw = w[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
  w[p] = nw;
  wb[p] = to_bf16(nw);
}

__global__ void mlp_round_kernel(const float* __restrict__ w, uint16_t* __restrict__ wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) wb[p] = to_bf16(w[p]);
}

}  // namespace mlp
}  // namespace ofb

using namespace ofb;
using namespace ofb::mlp;

struct ofb_mlp {
  Net n;
};

static int mlp_grad_launch(ofb_mlp* net, const float* x, const float* t, int n, hipStream_t st, float* loss) {
  const int blocks = (n + ROWS - 1) / ROWS;
  if (blocks > net->n.gcap) {
    if (net->n.gpart) (void)hipFree(net->n.gpart);
    if (net->n.lpart) (void)hipFree(net->n.lpart);
    net->n.gpart = nullptr; net->n.lpart = nullptr;
    if (hipMalloc((void**)&net->n.gpart, (size_t)blocks * NPARAM * 4) != hipSuccess ||
        hipMalloc((void**)&net->n.lpart, (size_t)blocks * 4) != hipSuccess) {
      set_error("hipMalloc failed (mlp gradient slots)");
      net->n.gcap = 0;
      return -4;
    }
    net->n.gcap = blocks;
  }
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, st, net->n.w, net->n.wb, x, t, n, 1.f / (float)n,
                     net->n.gpart, net->n.lpart);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, net->n.gpart, blocks,
                     net->n.grad, net->n.lpart, loss);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ---- gradient without allocation (ofb_mlp_gradient_noalloc) -------------------------
// The per-block gradient slots and losses of at most kNoAllocMaxRows rows live in a
// file-scope device array that every GPU the library runs on holds until the process
// ends (3.2 MB).  On the host its name is not a device address: host code uses only the
// address hipGetSymbolAddress returns for the current device, looked up once per device
// on the entry point's first call there and kept in g_scratch_ptr.  Every block of the
// gradient kernel writes its whole slot (all NPARAM entries and its loss), so nothing is
// cleared between calls.  Calls that overlap on one GPU share the array and race.
constexpr int kNoAllocMaxRows = 8192;
constexpr int kNoAllocBlocks = kNoAllocMaxRows / ROWS;
static __device__ __align__(256) float scratch[(size_t)kNoAllocBlocks * NPARAM + kNoAllocBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices];

// this device's copy of `scratch` (null with the error set on failure)
static float* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("no current HIP device (or device index >= %d)", kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress of the gradient scratch failed on device %d: %s", dev, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  return g_scratch_ptr[dev];
}

namespace ofb {
int mlp_forward_packed(const ofb_mlp* net, const float* x, const int32_t* count, int W, int K, int32_t* off,
                       float* y, hipStream_t st) {
  static_assert(ROWS == 64, "one packed row per lane of the 64-lane block");
  hipLaunchKernelGGL(mlp_offsets_kernel, dim3(1), dim3(1024), 0, st, W, K, count, off);
  const int blocks = (int)(((size_t)W * K + ROWS - 1) / ROWS);  // upper bound; blocks past the total exit
  hipLaunchKernelGGL(mlp_forward_packed_kernel, dim3(blocks), dim3(64), 0, st, net->n.w, net->n.wb, x, off, W, K,
                     y);
  if (hipGetLastError() != hipSuccess) { set_error("mlp forward launch failed"); return -4; }
  return 0;
}
}  // namespace ofb

extern "C" {

int ofb_mlp_create(ofb_mlp** out) {
  if (!out) { set_error("null argument"); return -1; }
  ofb_mlp* net = new ofb_mlp;
  std::memset(net, 0, sizeof(*net));
  const size_t b = (size_t)NPARAM * 4;
  if (hipMalloc((void**)&net->n.w, b) != hipSuccess || hipMalloc((void**)&net->n.wb, NPARAM * 2) != hipSuccess ||
      hipMalloc((void**)&net->n.m, b) != hipSuccess || hipMalloc((void**)&net->n.v, b) != hipSuccess ||
      hipMalloc((void**)&net->n.grad, b) != hipSuccess) {
    set_error("hipMalloc failed (mlp)");
    return -4;
  }
  if (hipMemset(net->n.w, 0, b) != hipSuccess || hipMemset(net->n.wb, 0, NPARAM * 2) != hipSuccess ||
      hipMemset(net->n.m, 0, b) != hipSuccess || hipMemset(net->n.v, 0, b) != hipSuccess) {
    set_error("hipMemset failed (mlp)");
    return -4;
  }
  *out = net;
  return 0;
}

int ofb_mlp_destroy(ofb_mlp* net) {
  if (!net) return 0;
  void* p[] = {net->n.w, net->n.wb, net->n.m, net->n.v, net->n.grad, net->n.gpart, net->n.lpart};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete net;
  return 0;
}

int ofb_mlp_num_params(void) { return NPARAM; }

int ofb_mlp_set_params(ofb_mlp* net, const float* params, void* stream) {
  if (!net || !params) { set_error("null argument"); return -1; }
  hipStream_t st = (hipStream_t)stream;
  if (hipMemcpyAsync(net->n.w, params, (size_t)NPARAM * 4, hipMemcpyDefault, st) != hipSuccess) {
    set_error("copy of mlp parameters failed");
    return -4;
  }
  hipLaunchKernelGGL(mlp_round_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, net->n.w, net->n.wb);
  net->n.step = 0;
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int ofb_mlp_get_params(const ofb_mlp* net, float* params, void* stream) {
  if (!net || !params) { set_error("null argument"); return -1; }
  if (hipMemcpyAsync(params, net->n.w, (size_t)NPARAM * 4, hipMemcpyDefault, (hipStream_t)stream) != hipSuccess) {
    set_error("copy of mlp parameters failed");
    return -4;
  }
  return 0;
}

int ofb_mlp_forward(const ofb_mlp* net, const float* x, int n, float* y, void* stream) {
  if (!net || !x || !y || n < 0) { set_error("invalid argument"); return -1; }
  if (n == 0) return 0;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + ROWS - 1) / ROWS), dim3(64), 0, (hipStream_t)stream, net->n.w,
                     net->n.wb, x, n, y);
  if (hipGetLastError() != hipSuccess) { set_error("mlp forward launch failed"); return -4; }
  return 0;
}

int ofb_mlp_gradient(ofb_mlp* net, const float* x, const float* target, int n, float* grad, float* loss,
                     void* stream) {
  if (!net || !x || !target || n < 1) { set_error("invalid argument"); return -1; }
  hipStream_t st = (hipStream_t)stream;
  int rc = mlp_grad_launch(net, x, target, n, st, loss);
  if (rc) return rc;
  if (grad && hipMemcpyAsync(grad, net->n.grad, (size_t)NPARAM * 4, hipMemcpyDefault, st) != hipSuccess) {
    set_error("copy of mlp gradient failed");
    return -4;
  }
  return 0;
}

int ofb_mlp_gradient_noalloc(ofb_mlp* net, const float* x, const float* target, int n, float* grad, float* loss,
                             void* stream) {
  if (!net || !x || !target || n < 1) { set_error("invalid argument"); return -1; }
  if (n > kNoAllocMaxRows) {
    set_error("ofb_mlp_gradient_noalloc takes at most %d rows (got %d); use ofb_mlp_gradient", kNoAllocMaxRows, n);
    return -2;
  }
  float* sp = scratch_on_current_device();
  if (!sp) return -4;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = (n + ROWS - 1) / ROWS;
  float* gpart = sp;
  float* lpart = sp + (size_t)kNoAllocBlocks * NPARAM;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, st, net->n.w, net->n.wb, x, target, n,
                     1.f / (float)n, gpart, lpart);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, gpart, blocks,
                     grad ? grad : net->n.grad, lpart, loss);
  if (hipGetLastError() != hipSuccess) { set_error("mlp gradient launch failed"); return -4; }
  return 0;
}

int ofb_mlp_train_step(ofb_mlp* net, const float* x, const float* target, int n, float lr, float* loss,
                       void* stream) {
  if (!net || !x || !target || n < 1 || !(lr > 0.f)) { set_error("invalid argument"); return -1; }
  hipStream_t st = (hipStream_t)stream;
  int rc = mlp_grad_launch(net, x, target, n, st, loss);
  if (rc) return rc;
  net->n.step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  float bc1 = 1.f - powf(b1, (float)net->n.step), bc2 = 1.f - powf(b2, (float)net->n.step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, net->n.w, net->n.wb, net->n.m,
                     net->n.v, net->n.grad, lr, b1, b2, 1e-8f, bc1, bc2);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // extern "C"
