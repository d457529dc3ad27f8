// context.h -- the device context of libgsh (gsh_ctx) and the host helpers the entry-point
// sources share: HIP status checks and the bounded wait on a stream event.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <string>
#include <thread>

#include "abi_util.h"
#include "device.h"

namespace gsh {

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? GSH_E_OOM : GSH_E_DEVICE;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) gsh::hip_check((x), #x)

using Clock = std::chrono::steady_clock;

// Waits for an event with a time limit (pcg_loop.cpp:444-456 raises instead of hanging).
inline void wait_event(hipEvent_t ev, double timeout_s, const char* what) {
  const auto t0 = Clock::now();
  int spins = 0;
  for (;;) {
    const hipError_t r = hipEventQuery(ev);
    if (r == hipSuccess) return;
    if (r != hipErrorNotReady) hip_check(r, what);
    if (std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s)
      throw Error(GSH_E_TIMEOUT, std::string(what) + ": the GPU did not finish within the time limit");
    if (++spins > 64) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace gsh

struct gsh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // meshes, solvers and servers on this context; gsh_ctx_destroy frees the context once the
  // last of them is destroyed, so objects destroyed after the context never see freed state
  std::atomic<int> refs{0};
  std::atomic<bool> closed{false};
  gsh::PcgState* pinned = nullptr;   // host-visible readback slots
  int pinned_cap = 0;
  int32_t* pinned_flag = nullptr;    // host-visible word for small device-to-host readbacks
  hipEvent_t ev[2] = {nullptr, nullptr};
  void set() const { HIPCHK(hipSetDevice(device)); }
  gsh::PcgState* pinned_slots(int per_slot) {
    if (2 * per_slot > pinned_cap) {
      if (pinned) (void)hipHostFree(pinned);
      pinned = nullptr;
      const int cap = std::max(64, 2 * per_slot);
      HIPCHK(hipHostMalloc((void**)&pinned, sizeof(gsh::PcgState) * cap, hipHostMallocDefault));
      pinned_cap = cap;
    }
    return pinned;
  }
};
