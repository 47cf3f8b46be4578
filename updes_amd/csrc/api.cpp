// Library-level entry points: ABI version, error reporting, device queries.
#include <stdarg.h>
#include <atomic>
#include <stdio.h>

#include "common.h"

namespace rbfx {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

static DeviceInfo g_info[64];
static int g_clock_khz[64];
// Wait bound in seconds (rbfx_set_wait_seconds); negative: kWaitSeconds.
static std::atomic<double> g_wait_seconds{-1.0};

const DeviceInfo& device_info() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  DeviceInfo& d = g_info[dev];
  if (d.device != dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      d.num_cus = prop.multiProcessorCount;
      d.lds_per_block = prop.sharedMemPerBlock;
    } else {
      d.num_cus = 256;
      d.lds_per_block = 65536;
    }
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
      khz = 100000;  // CDNA real-time counter: 100 MHz
    g_clock_khz[dev] = khz;
    d.device = dev;
  }
  const double secs = g_wait_seconds.load(std::memory_order_relaxed);
  d.wait_ticks = (uint64_t)((secs < 0.0 ? kWaitSeconds : secs) * 1e3 * g_clock_khz[dev]);
  return d;
}

}  // namespace rbfx

extern "C" int rbfx_abi_version(void) { return RBFX_ABI_VERSION; }

extern "C" const char* rbfx_last_error(void) { return rbfx::last_error(); }

extern "C" int rbfx_set_wait_seconds(double seconds) {
  rbfx::g_wait_seconds.store(seconds, std::memory_order_relaxed);
  return RBFX_OK;
}

extern "C" int rbfx_init(void) {
  int n = 0;
  RBFX_HIP(hipGetDeviceCount(&n));
  RBFX_REQUIRE(n > 0, "no HIP device");
  (void)rbfx::device_info();
  return RBFX_OK;
}

// ---- per-launch timing ------------------------------------------------------
#include <mutex>
#include <string>
#include <vector>

namespace rbfx {
namespace {
struct ProfRec {
  std::string name;
  hipEvent_t a = nullptr, b = nullptr;
  double flops = 0, bytes = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

int prof_start(const char* name, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on) return -1;
  ProfRec r;
  r.name = name;
  r.a = take_event();
  r.b = take_event();
  if (!r.a || !r.b) return -1;
  if (hipEventRecord(r.a, s) != hipSuccess) return -1;
  g_prof.push_back(r);
  return (int)g_prof.size() - 1;
}

void prof_stop(int id, hipStream_t s, double flops, double bytes) {
  if (id < 0) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (id >= (int)g_prof.size()) return;
  g_prof[id].flops = flops;
  g_prof[id].bytes = bytes;
  (void)hipEventRecord(g_prof[id].b, s);
}
}  // namespace rbfx

extern "C" int rbfx_profile_begin(void) {
  std::lock_guard<std::mutex> lk(rbfx::g_prof_mu);
  for (auto& r : rbfx::g_prof) {
    rbfx::g_pool.push_back(r.a);
    rbfx::g_pool.push_back(r.b);
  }
  rbfx::g_prof.clear();
  rbfx::g_prof_on = true;
  return RBFX_OK;
}

extern "C" int rbfx_profile_end(rbfx_prof_entry* out, int max_entries, int* n_entries) {
  std::vector<rbfx_prof_entry> agg;
  {
    std::lock_guard<std::mutex> lk(rbfx::g_prof_mu);
    rbfx::g_prof_on = false;
    for (auto& r : rbfx::g_prof) {
      RBFX_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      RBFX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      rbfx_prof_entry* e = nullptr;
      for (auto& x : agg)
        if (r.name == x.name) e = &x;
      if (!e) {
        rbfx_prof_entry z{};
        snprintf(z.name, sizeof(z.name), "%s", r.name.c_str());
        agg.push_back(z);
        e = &agg.back();
      }
      e->launches += 1;
      e->total_ms += ms;
      e->flops += r.flops;
      e->bytes += r.bytes;
    }
  }
  const int n = (int)agg.size();
  for (int i = 0; i < n && i < max_entries; ++i) out[i] = agg[i];
  if (n_entries) *n_entries = n;
  return RBFX_OK;
}

extern "C" int rbfx_shutdown(void) {
  {
    std::lock_guard<std::mutex> lk(rbfx::g_prof_mu);
    for (auto& r : rbfx::g_prof) {
      rbfx::g_pool.push_back(r.a);
      rbfx::g_pool.push_back(r.b);
    }
    rbfx::g_prof.clear();
    rbfx::g_prof_on = false;
    for (hipEvent_t e : rbfx::g_pool) (void)hipEventDestroy(e);
    rbfx::g_pool.clear();
  }
  return rbfx::release_panel_streams();
}
