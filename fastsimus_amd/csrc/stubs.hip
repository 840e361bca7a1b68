// Entry points whose kernels are not built yet report FS_ERR_UNSUPPORTED.
#include "fs_device.h"

namespace fs {
int adjoint_fp32(const Problem&, const cf*, float*, float*, float*, double*, char*, hipStream_t) {
  return FS_ERR_UNSUPPORTED;
}
size_t adjoint_workspace(const Problem&) { return 256; }
}  // namespace fs
