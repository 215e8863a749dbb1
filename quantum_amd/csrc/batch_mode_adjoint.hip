// Instantiates the batch kernel for kModeAdjoint.
#include "batch_kernel_impl.h"

namespace tfqx {
namespace {

template <int MODE>
const void* KernelPtr(int kmax) {
  return kmax <= 2 ? reinterpret_cast<const void*>(&batch_kernel<MODE, 2>)
                   : reinterpret_cast<const void*>(&batch_kernel<MODE, 4>);
}

}  // namespace

const void* BatchKernelAdjoint(int mode, int kmax) {
  switch (mode) {
    default: return KernelPtr<kModeAdjoint>(kmax);
  }
}

}  // namespace tfqx
