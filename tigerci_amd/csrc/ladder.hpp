// Grouped 4-external ladder launch (see ladder.hip).
#pragma once

#include <algorithm>

#include "common.hpp"

namespace tci {

struct LadderItem {
  int32_t P;      // pair walk (-1: padding)
  int32_t r0;     // first row of the 64-row panel
  int32_t J0, J1; // Cholesky vector range
  int32_t chunk;  // partial-sum slot
  int32_t pad_;
};

struct LadderArgs {
  const int64_t* dom_off;
  const int32_t* dom_idx;
  const int64_t* cd_off;
  const double* Cd;       // dense C blocks (FP64)
  const void* Lvv;        // [n_chol][n_v][n_v], FP64 or FP32
  bool lvv_f32;
  int64_t n_v;
  double* partial;        // [n_chunks][sum d^2]
  int64_t chunk_stride;
};

struct LadderLaunch {
  DevBuf d_items;
  int64_t n_per_xcd = 0;
  int64_t n_items = 0;
  int dpad_max = 0;
  void build(const std::vector<LadderItem>& items, const std::vector<double>& cost, int dmax,
             hipStream_t st);
  void launch(const LadderArgs& a, hipStream_t st) const;
};

void ladder_reduce(const double* partial, int64_t chunk_stride, int n_chunks, double* R, int64_t n,
                   hipStream_t st);

}  // namespace tci
