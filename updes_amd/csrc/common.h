// Internal helpers shared by the rbfx translation units (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/rbfx.h"
#include "host_logic.h"

namespace rbfx {

// Thread-local last error message (rbfx_last_error).
void set_error(const char* fmt, ...);
const char* last_error();

#define RBFX_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::rbfx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                  \
      return RBFX_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

#define RBFX_CHECK_LAUNCH()                                                   \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ::rbfx::set_error("kernel launch failed: %s (%s:%d)",                   \
                        hipGetErrorString(e_), __FILE__, __LINE__);           \
      return RBFX_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

#define RBFX_REQUIRE(cond, ...)                                               \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::rbfx::set_error(__VA_ARGS__);                                         \
      return RBFX_ERR_ARG;                                                    \
    }                                                                         \
  } while (0)

struct DeviceInfo {
  int device = -1;
  int num_cus = 0;
  size_t lds_per_block = 0;
  // Ticks of the device's constant-rate real-time counter (wall_clock64,
  // hipDeviceAttributeWallClockRate) in kWaitSeconds: the time bound of
  // every device-side wait.
  uint64_t wait_ticks = 0;
};
// Cached per current device.
const DeviceInfo& device_info();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }


// Every device-side wait gives up after kWaitSeconds of real time, read from
// the constant-rate counter (wall_clock64), whatever each poll costs; it
// then sets a status word that the host turns into RBFX_ERR_TIMEOUT.
constexpr double kWaitSeconds = 4.0;

// Real-time deadline of one wait.  The counter is read once per
// kWaitCheckEvery polls (a read is a scalar memory op of ~100 cycles).
constexpr uint32_t kWaitCheckEvery = 16;
struct WaitClock {
  uint64_t t0, ticks;
  uint32_t polls = 0;
  __device__ explicit WaitClock(uint64_t budget) : t0(wall_clock64()), ticks(budget) {}
  // true once the budget is spent
  __device__ bool expired() {
    if (++polls % kWaitCheckEvery) return false;
    return wall_clock64() - t0 > ticks;
  }
};

}  // namespace rbfx

#include <type_traits>

namespace rbfx {
// Data exchanged between workgroups of one kernel (pivot records, solved
// slices of a triangular sweep) goes through agent-coherent relaxed loads and
// stores, which bypass the non-coherent cache levels.  The writer waits for
// completion of its stores (s_waitcnt 0) before it publishes a flag, and the
// reader reads the data after it has seen the flag; no cache-wide write-back
// or invalidation (what agent-scope release/acquire fences cost, several us
// per exchange on MI355X, measured) is involved.
template <typename T>
__device__ inline void st_coh(T* p, T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32/64-bit only");
  typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type U;
  U u;
  __builtin_memcpy(&u, &v, sizeof(T));
  __hip_atomic_store(reinterpret_cast<U*>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T ld_coh(const T* p) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32/64-bit only");
  typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type U;
  U u = __hip_atomic_load(reinterpret_cast<const U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  T v;
  __builtin_memcpy(&v, &u, sizeof(T));
  return v;
}
}  // namespace rbfx

namespace rbfx {
// Optional per-launch timing with HIP events on the launch stream
// (rbfx_profile_begin / rbfx_profile_end).  When disabled, prof_start()
// returns -1 and prof_stop() does nothing.
int prof_start(const char* name, hipStream_t s);
void prof_stop(int id, hipStream_t s, double flops, double bytes);
// Workspace bytes of the transposed matrix-free product over n rows.
size_t matvec_t_ws_bytes(int64_t n);
// Destroys the LU's per-device panel streams and events.
int release_panel_streams();
}  // namespace rbfx
