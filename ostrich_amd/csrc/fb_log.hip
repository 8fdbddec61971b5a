// fb_log.hip -- device-resident step logs: every capture copies a fixed set of device
// arrays into slot `step` of their [capacity, ...] buffers, with `step` read on the device
// (a counter the capture can advance), so the capture records into a graph with the
// step it follows.  The behaviour of the reference's history buffers
// (core/history_group.py:12-80: buffer[idx] = x when idx < capacity, nothing otherwise)
// behind its simulation and dataset loggers (logging/sim_logger.py:12-110,
// logging/dataset_logger.py:16-95).  Elements are 4 bytes (float32 / int32).

#include <hip/hip_runtime.h>

#include <vector>

#include "fb_solver.h"

namespace {

struct LogEntry {
  const uint32_t* src;
  uint32_t* dst;
  long long words;  // per step
};

constexpr int kLogThreads = 256;

__global__ void __launch_bounds__(kLogThreads) log_copy_kernel(const LogEntry* table, const int32_t* step, int cap) {
  const int idx = step[0];
  if (idx < 0 || idx >= cap) return;
  const LogEntry e = table[blockIdx.y];
  uint32_t* dst = e.dst + (size_t)idx * e.words;
  for (long long i = (long long)blockIdx.x * kLogThreads + threadIdx.x; i < e.words;
       i += (long long)gridDim.x * kLogThreads)
    dst[i] = e.src[i];
}

__global__ void log_advance_kernel(int32_t* step) {
  if (threadIdx.x == 0) step[0] = step[0] + 1;
}

}  // namespace

struct ofb_log {
  int n, cap;
  long long max_words;
  LogEntry* table;  // device
};

extern "C" {

int ofb_log_destroy(ofb_log* log) {
  if (!log) return 0;
  if (log->table) (void)hipFree(log->table);
  delete log;
  return 0;
}

int ofb_log_create(int32_t num_arrays, const void* const* src, void* const* dst, const int64_t* words_per_step,
                   int32_t capacity, ofb_log** out) {
  if (!out || num_arrays < 1 || !src || !dst || !words_per_step) {
    ofb::set_error("invalid log description");
    return -1;
  }
  *out = nullptr;
  if (capacity < 1 || num_arrays > 65535) {
    ofb::set_error("log capacity must be >= 1 and at most 65535 arrays");
    return -2;
  }
  std::vector<LogEntry> t(num_arrays);
  long long mx = 1;
  for (int a = 0; a < num_arrays; ++a) {
    if (!src[a] || !dst[a] || words_per_step[a] < 1) {
      ofb::set_error("log array %d: null pointer or empty", a);
      return -2;
    }
    t[a] = LogEntry{(const uint32_t*)src[a], (uint32_t*)dst[a], (long long)words_per_step[a]};
    if (t[a].words > mx) mx = t[a].words;
  }
  ofb_log* log = new ofb_log();
  log->n = num_arrays;
  log->cap = capacity;
  log->max_words = mx;
  if (hipMalloc((void**)&log->table, sizeof(LogEntry) * num_arrays) != hipSuccess ||
      hipMemcpy(log->table, t.data(), sizeof(LogEntry) * num_arrays, hipMemcpyHostToDevice) != hipSuccess) {
    ofb_log_destroy(log);
    ofb::set_error("hipMalloc failed (log)");
    return -4;
  }
  *out = log;
  return 0;
}

int ofb_log_capture(ofb_log* log, int32_t* step_index, int32_t advance, void* stream) {
  if (!log || !step_index) {
    ofb::set_error("null argument");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long chunks = (log->max_words + kLogThreads - 1) / kLogThreads;
  const unsigned gx = (unsigned)(chunks < 1024 ? chunks : 1024);
  hipLaunchKernelGGL(log_copy_kernel, dim3(gx, log->n), dim3(kLogThreads), 0, st, log->table, step_index, log->cap);
  if (advance) hipLaunchKernelGGL(log_advance_kernel, dim3(1), dim3(64), 0, st, step_index);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ofb::set_error("log launch failed: %s", hipGetErrorString(e));
    return -4;
  }
  return 0;
}

}  // extern "C"
