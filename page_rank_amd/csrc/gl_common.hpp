// Shared host/device helpers of the gfx950 PageRank library: error reporting, launch
// sizing, XCD-aware stripe scheduling and wave64 cross-lane primitives.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/graph_link_hip.h"

namespace gl {

// ---- error reporting (thread-local last error, never throws across the ABI) ----
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

struct DeviceInfo {
  int num_cus = 256;
  int lds_bytes = 160 * 1024;
  int clock_khz = 0;
  int64_t hbm_bytes = 0;
};
// Cached per current device.
const DeviceInfo& device_info();

// A non-blocking stream (default priority) with two events, leased by one call at a time.
struct StreamSet {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr;   // recorded on the caller's stream, waited on by `stream`
  hipEvent_t ev_out = nullptr;  // recorded on `stream`, waited on by the caller's stream
};
// Lease a set of the current device (nullptr and the last error set on failure).
StreamSet* acquire_stream_set();
void release_stream_set(StreamSet* s);
struct StreamLease {
  StreamSet* s;
  StreamLease() : s(acquire_stream_set()) {}
  ~StreamLease() { release_stream_set(s); }
  StreamLease(const StreamLease&) = delete;
  StreamLease& operator=(const StreamLease&) = delete;
};

}  // namespace gl

#define GL_HIP(call)                                                        \
  do {                                                                      \
    hipError_t gl_e_ = (call);                                              \
    if (gl_e_ != hipSuccess) return gl::hip_fail(gl_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define GL_LAUNCH_CHECK(name) GL_HIP(hipGetLastError())

#define GL_REQUIRE(cond, ...)          \
  do {                                 \
    if (!(cond)) {                     \
      gl::set_error(__VA_ARGS__);      \
      return GL_ERR_INVALID;           \
    }                                  \
  } while (0)

namespace gl {

constexpr int kWave = 64;
constexpr int kNumXcd = 8;  // MI355X: 8 accelerator complex dies, each with its own L2

// Persistent-launch schedule over `n` work items (stripes). Workgroups are dispatched
// round-robin over the XCDs (workgroup w runs on XCD w % 8), so each XCD is given one
// contiguous range of items and its waves interleave inside that range: stripes that sit
// next to each other (and read neighbouring X rows) share one L2.
struct Sched {
  int64_t first, end, stride;
};
__device__ __forceinline__ Sched xcd_sched(int64_t n) {
  const int64_t nwg = gridDim.x;  // host guarantees a multiple of kNumXcd
  const int64_t per = nwg / kNumXcd;
  const int x = blockIdx.x % kNumXcd;
  const int64_t i = blockIdx.x / kNumXcd;
  const int wpb = blockDim.x / kWave;
  // the wave index is uniform across the wave: keep it (and the schedule) in scalar registers
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t lo = n * x / kNumXcd;
  const int64_t hi = n * (x + 1) / kNumXcd;
  Sched s;
  s.first = lo + i * wpb + wid;
  s.end = hi;
  s.stride = per * wpb;
  return s;
}

// LDS written by some lanes of a wave, then read by other lanes of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row_newbcast:n (gfx90a+ DPP): every lane of a 16-lane row reads lane n of that row.
template <int N>
__device__ __forceinline__ float row_bcast(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + N, 0xF, 0xF, true));
}

__device__ __forceinline__ float shfl_xor(float v, int m) {
  return __shfl_xor(v, m, kWave);
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Wave-uniform read of an array that no kernel of the launch writes, through the scalar
// cache (s_load). A plain load of a uniform address compiles to a vector load and a
// readfirstlane, and the wait before that readfirstlane drains every vector memory operation
// the wave has in flight (prefetches and stores alike), once per stripe in the stripe loops.
// The constant address space tells the compiler the data cannot change during the kernel.
template <class T>
__device__ __forceinline__ T sload(const T* p, int64_t i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

}  // namespace gl
