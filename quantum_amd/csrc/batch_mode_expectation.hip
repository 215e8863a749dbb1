// Instantiates the batch kernel for kModeExpectation.
#include "batch_kernel_impl.h"

namespace tfqx {
namespace {

template <int MODE>
const void* KernelPtr(int kmax) {
  return kmax <= 2 ? reinterpret_cast<const void*>(&batch_kernel<MODE, 2>)
                   : reinterpret_cast<const void*>(&batch_kernel<MODE, 4>);
}

}  // namespace

const void* BatchKernelExpectation(int mode, int kmax) {
  switch (mode) {
    default: return KernelPtr<kModeExpectation>(kmax);
  }
}

}  // namespace tfqx
