// 192 x 192 split GEMM tiles in bf16 (kernel template in xgemm.hpp).
#include "xgemm.hpp"

namespace tci {

void launch_xgemm192_bf16(const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, const float* scales,
                          hipStream_t st, int remap) {
  launch_xgemm_shape<192, 192, TCI_SIGMA_BF16X3>(tasks, segs, ntasks, J0, scales, st, remap);
}

}  // namespace tci
