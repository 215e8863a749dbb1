// Allocation-free launch: tfqx_plan_launch_noalloc.
//
// Some callers cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  This entry point
// runs a plan with its workspace in a file-scope device array instead of a
// caller-provided buffer: it copies the plan's device image there and
// launches the plan's kernels on it.  The caller still provides the output.
//
// Each GPU the library runs on holds its own copy of the array (kNoAllocBytes)
// until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and cached; host code never
// passes the array's name to a kernel or a copy, only that address.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents.  The copied image re-initialises everything the kernels need to
// start at zero (queue counters, accumulators), and the kernels clear their
// own scratch (gradient accumulators) before use.  Calls that overlap on one
// GPU (different streams, or host threads without ordering) race on the
// array: callers serialise them, e.g. by using one stream.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <mutex>
#include <string>

#include "tfqx.h"

namespace tfqx {
int64_t PlanLaunchBytes(const tfqx_plan* p);  // api.cc
}  // namespace tfqx

namespace {

constexpr size_t kNoAllocBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

int Fail(char* err, int64_t len, int code, const std::string& msg) {
  if (err && len > 0) {
    const size_t n = std::min<size_t>(size_t(len - 1), msg.size());
    memcpy(err, msg.data(), n);
    err[n] = '\0';
  }
  return code;
}

// The array's address on the current device (looked up on first use there).
int DeviceScratch(void** out, char* err, int64_t len) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return Fail(err, len, TFQX_DEVICE_ERROR, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return Fail(err, len, TFQX_DEVICE_ERROR, "device index out of range: " + std::to_string(dev));
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return Fail(err, len, TFQX_DEVICE_ERROR,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = ptr;
  }
  *out = g_ptr[dev];
  return TFQX_OK;
}

}  // namespace

extern "C" {

int tfqx_plan_launch_noalloc(tfqx_plan* p, void* output, void* stream, char* err,
                             int64_t err_len) {
  if (!p) return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "null plan");
  int64_t stats[10] = {0};
  tfqx_plan_stats(p, stats, 10);
  if (stats[6] != 0)
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT,
                "tfqx_plan_launch_noalloc: plans on the HBM-streaming path (circuits wider than "
                "the batch kernel takes) are not supported; use tfqx_plan_launch");
  const int64_t need = tfqx::PlanLaunchBytes(p);
  if (need < 0 || size_t(need) > kNoAllocBytes)
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT,
                "tfqx_plan_launch_noalloc: the plan needs " + std::to_string(need) +
                    " workspace bytes, more than the static scratch of " +
                    std::to_string(kNoAllocBytes) + " bytes; use tfqx_plan_launch");
  void* ws = nullptr;
  int rc = DeviceScratch(&ws, err, err_len);
  if (rc != TFQX_OK) return rc;
  rc = tfqx_plan_upload(p, ws, stream, err, err_len);
  if (rc != TFQX_OK) return rc;
  return tfqx_plan_launch(p, ws, output, stream, err, err_len);
}

}  // extern "C"
