// Shared helpers for the ceres_mi355x HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ceres_mi355x.h"

namespace ct {

// ---------------------------------------------------------------------------
// error reporting (thread-local message, negative status codes)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define CT_CHECK_ARG(cond, msg)              \
  do {                                       \
    if (!(cond)) {                           \
      ::ct::set_error(std::string(msg));     \
      return CT_ERR_ARGUMENT;                \
    }                                        \
  } while (0)

#define CT_HIP(call)                                                                            \
  do {                                                                                          \
    hipError_t err_ = (call);                                                                   \
    if (err_ != hipSuccess) {                                                                   \
      ::ct::set_error(std::string(#call) + " failed: " + hipGetErrorString(err_));              \
      return CT_ERR_HIP;                                                                        \
    }                                                                                           \
  } while (0)

#define CT_LAUNCH_CHECK()                                                                       \
  do {                                                                                          \
    hipError_t err_ = hipGetLastError();                                                        \
    if (err_ != hipSuccess) {                                                                   \
      ::ct::set_error(std::string("kernel launch failed: ") + hipGetErrorString(err_));          \
      return CT_ERR_HIP;                                                                        \
    }                                                                                           \
  } while (0)

#define CT_TRY(expr)            \
  do {                          \
    int st_ = (expr);           \
    if (st_ != CT_OK) return st_; \
  } while (0)

// ---------------------------------------------------------------------------
// workspace carving: 256-byte aligned slices of one caller buffer
// ---------------------------------------------------------------------------
struct Carver {
  char* base;
  size_t offset = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    offset = (offset + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + offset) : nullptr;
    offset += count * sizeof(T);
    return p;
  }
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ int wave_min_int(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
  return v;
}

// Deterministic block sum: fixed shuffle tree inside each wave, then waves
// summed in index order by thread 0.  Result valid in thread 0 only.
template <int kThreads>
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double total = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kThreads / kWave; ++w) total += scratch[w];
  }
  __syncthreads();
  return total;
}

template <int kThreads>
__device__ __forceinline__ double block_max(double v, double* scratch) {
  v = wave_max(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double total = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kThreads / kWave; ++w) total = fmax(total, scratch[w]);
  }
  __syncthreads();
  return total;
}

__device__ __forceinline__ bool is_finite(double v) { return isfinite(v); }

}  // namespace ct
