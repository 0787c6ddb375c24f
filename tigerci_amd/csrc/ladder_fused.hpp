// Fused split-precision ladder task table (ladder_fused_split.hip).
#pragma once

#include "common.hpp"

namespace tci {

// One workgroup: panel a0..a0+63 of the c index of walk Q (a partial sum of
// R_Q over that panel), Cholesky vectors [j0, j1).
struct FusedTask {
  const uint16_t* L;   // split rows of L^J[D,D] of this walk, J-block every jstep elements
  const uint16_t* C;   // split rows of C_Q (row c, k contiguous)
  double* R;           // K-split slice of R_Q (row-major d x d) of this panel and J range
  int64_t jstep;       // elements between consecutive J blocks
  int32_t d, ldx;      // domain size (129..192), split row stride (elements)
  int32_t a0, j0, j1;  // panel start in c, Cholesky vectors [j0, j1)
  int32_t pad_;
};

// R[n][a] += s4 sum_J sum_{c in panel} L_J[n][c] (s3 L_J C^T)[a][c] for each
// task (scales[3], scales[4]; accumulates into zero-filled K-split slices)
void launch_xladder_fused(int precision, const FusedTask* tasks, int64_t ntasks, const float* scales, hipStream_t st,
                          int remap);

// Domains of 65..128 (ladder_fused128.hip): one task is a whole walk (a0 = 0)
// for Cholesky vectors [j0, j1); R[n][a] += s4 sum_J sum_c L_J[n][c]
// (s3 L_J C^T)[a][c] into the task's zero-filled K-split slice.
void launch_xladder_fused128(int precision, const FusedTask* tasks, int64_t ntasks, const float* scales,
                             hipStream_t st, int remap, int waves = 4);

// Domains of 129..192 (ladder_fused192.hip): one task is a whole walk (a0 = 0)
// for Cholesky vectors [j0, j1); C_Q in LDS, L_J read from global memory.
// The L blocks must be followed by 16 KiB of readable memory.
void launch_xladder_fused192(int precision, const FusedTask* tasks, int64_t ntasks, const float* scales,
                             hipStream_t st, int remap);

}  // namespace tci
