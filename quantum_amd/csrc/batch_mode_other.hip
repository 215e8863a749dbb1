// Instantiates the batch kernel for kModeState, kModeSamples, kModeSampledExpectation, kModeInnerProduct.
#include "batch_kernel_impl.h"

namespace tfqx {
namespace {

template <int MODE>
const void* KernelPtr(int kmax) {
  return kmax <= 2 ? reinterpret_cast<const void*>(&batch_kernel<MODE, 2>)
                   : reinterpret_cast<const void*>(&batch_kernel<MODE, 4>);
}

}  // namespace

const void* BatchKernelOther(int mode, int kmax) {
  switch (mode) {
    case kModeState: return KernelPtr<kModeState>(kmax);
    case kModeSamples: return KernelPtr<kModeSamples>(kmax);
    case kModeSampledExpectation: return KernelPtr<kModeSampledExpectation>(kmax);
    default: return KernelPtr<kModeInnerProduct>(kmax);
  }
}

}  // namespace tfqx
