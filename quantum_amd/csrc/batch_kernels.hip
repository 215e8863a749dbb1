// Host side of the persistent batch kernel: LDS sizing, occupancy and launch.
// The kernels themselves are instantiated per mode in batch_mode_*.hip.
#include <hip/hip_runtime.h>

#include "batch_kernels.h"
#include "lds_ops.h"  // kThreads

namespace tfqx {

const void* BatchKernelExpectation(int mode, int kmax);
const void* BatchKernelAdjoint(int mode, int kmax);
const void* BatchKernelInnerProductGrad(int mode, int kmax);
const void* BatchKernelOther(int mode, int kmax);

namespace {

const void* KernelFor(int mode, int kmax) {
  switch (mode) {
    case kModeExpectation: return BatchKernelExpectation(mode, kmax);
    case kModeAdjoint: return BatchKernelAdjoint(mode, kmax);
    case kModeInnerProductGrad: return BatchKernelInnerProductGrad(mode, kmax);
    default: return BatchKernelOther(mode, kmax);
  }
}

}  // namespace

int BatchBuffers(int mode) {
  return (mode == kModeAdjoint || mode == kModeSampledExpectation || mode == kModeInnerProduct ||
          mode == kModeInnerProductGrad)
             ? 2
             : 1;
}

size_t BatchLdsBytes(int lds_log2, int mode) {
  return size_t(BatchBuffers(mode)) * (size_t(1) << lds_log2) * sizeof(float2);
}

hipError_t LaunchBatch(const BatchArgs& args, int grid, hipStream_t stream) {
  // gfx950 workgroups may use up to 160 KiB of LDS without an opt-in
  // attribute (and setting one above 64 KiB is rejected by the runtime).
  const size_t lds = BatchLdsBytes(args.lds_log2, args.mode);
  void* params[] = {const_cast<BatchArgs*>(&args)};
  return hipLaunchKernel(KernelFor(args.mode, args.kmax), dim3(grid), dim3(kThreads), params, lds,
                         stream);
}

int BatchMaxGrid(int lds_log2, int mode, int kmax) {
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t lds = BatchLdsBytes(lds_log2, mode);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, KernelFor(mode, kmax), kThreads, lds);
  if (per_cu < 1) per_cu = 1;
  if (cus < 1) cus = 1;
  return cus * per_cu;
}

}  // namespace tfqx
