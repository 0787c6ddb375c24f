// Launch-schedule autotuning.
//
// Plans pick K-split fill targets and the workgroup order of their GEMM
// programs by timing the candidates on the plan's own buffers at set-up (the
// FP64 ladder when the plan is built, the split ladder when a split precision
// is first prepared).  Candidates that change K-splits change the summation
// order only; the workgroup order does not change results at all.
// TCI_AUTOTUNE=0 keeps the fixed defaults (reproducible across plans).
#include <cstdio>
#include <cstdlib>

#include "common.hpp"

namespace tci {

bool autotune_enabled() {
  const char* e = std::getenv("TCI_AUTOTUNE");
  return !(e && e[0] == '0');
}

std::string fmt_ms(double ms) {
  char b[32];
  std::snprintf(b, sizeof b, "%.3fms", ms);
  return b;
}

double time_ms(hipStream_t st, int reps, const std::function<void()>& fn) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TCI_HIP(hipEventCreate(&e0));
  TCI_HIP(hipEventCreate(&e1));
  float ms = 0.f;
  try {
    fn();  // untimed: first-touch and instruction-cache effects
    TCI_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) fn();
    TCI_HIP(hipEventRecord(e1, st));
    TCI_HIP(hipEventSynchronize(e1));
    TCI_HIP(hipEventElapsedTime(&ms, e0, e1));
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

int tune_order(GemmProgram& g, hipStream_t st, const std::function<void()>& launch) {
  if (xcd_remap_forced() || g.tasks.empty()) return g.remap;
  double t[2];
  for (int o = 0; o < 2; ++o) {
    g.remap = o;
    t[o] = time_ms(st, 2, launch);
  }
  // plain order unless the XCD-contiguous one is clearly faster
  g.remap = t[1] < 0.97 * t[0] ? 1 : 0;
  return g.remap;
}

std::string tune_block_order(GemmProgram& g, hipStream_t st, const std::function<void()>& launch) {
  if (xcd_remap_forced() || g.tasks.empty()) return "-";
  const std::vector<Task> base = g.tasks;
  const int cur_remap = g.remap < 0 ? 0 : g.remap;
  g.remap = cur_remap;
  const double t0 = time_ms(st, 2, launch);
  double best = t0;
  int bg = 0, bo = cur_remap;
  for (int o = 0; o < 2; ++o)
    for (int G : {0, 4, 8, 16}) {
      if (G == 0 && o == cur_remap) continue;
      g.block_order(G, base, st);
      g.remap = o;
      const double t = time_ms(st, 2, launch);
      if (t < best) {
        best = t;
        bg = G;
        bo = o;
      }
    }
  if (best > 0.97 * t0) {  // keep the order it was built with
    bg = 0;
    bo = cur_remap;
  }
  g.block_order(bg, base, st);
  g.remap = bo;
  TCI_HIP(hipStreamSynchronize(st));
  return (bo ? "x" : "p") + std::to_string(bg);
}

}  // namespace tci
