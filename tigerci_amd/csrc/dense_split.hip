// Split-precision grouped GEMM: tile dispatch and program launch (kernels in
// xgemm.hpp).
#include "xgemm.hpp"

namespace tci {

template <int PREC>
void launch_xgemm_t(int tm, int tn, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, const float* scales,
                    hipStream_t st, int remap) {
  if (tm == 128 && tn == 128) launch_xgemm_shape<128, 128, PREC>(tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 128 && tn == 192) launch_xgemm_shape<128, 192, PREC>(tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 128 && tn == 64) launch_xgemm_shape<128, 64, PREC>(tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 64 && tn == 64) launch_xgemm_shape<64, 64, PREC>(tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 192 && tn == 192 && PREC == TCI_SIGMA_BF16X3)
    launch_xgemm192_bf16(tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 192 && tn == 192) launch_xgemm192_f16(tasks, segs, ntasks, J0, scales, st, remap);
  else throw Error(TCI_ERR_ARG, "unsupported split GEMM tile " + std::to_string(tm) + "x" + std::to_string(tn));
}

void launch_xgemm(int precision, int tm, int tn, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0,
                  const float* scales, hipStream_t st, int remap) {
  TCI_REQUIRE(precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3, "split GEMM needs a split precision");
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many gemm tasks");
  if (precision == TCI_SIGMA_BF16X3) launch_xgemm_t<TCI_SIGMA_BF16X3>(tm, tn, tasks, segs, ntasks, J0, scales, st, remap);
  else launch_xgemm_t<TCI_SIGMA_FP16X3>(tm, tn, tasks, segs, ntasks, J0, scales, st, remap);
  TCI_HIP(hipGetLastError());
}

void GemmProgram::launch_split(int precision, hipStream_t st, int64_t J0, const float* scales) const {
  TCI_REQUIRE(split, "not a split program");
  TCI_REQUIRE(!scaled || scales, "scaled split tasks need a scales array");
  const int rm = remap < 0 || xcd_remap_forced() ? xcd_remap_default() : remap;
  if (tile == 256) {
    launch_xgemm256(precision, waves, d_tasks.as<Task>(), d_segs.as<Seg>(), (int64_t)tasks.size(), J0, scales, st, rm);
    return;
  }
  if (dma) {
    TCI_REQUIRE(precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3, "split GEMM needs a split precision");
    launch_xgemmd(precision, tile, tile_n > 0 ? tile_n : tile, dma, d_tasks.as<Task>(), d_segs.as<Seg>(),
                  (int64_t)tasks.size(), J0, scales, st, rm);
    return;
  }
  launch_xgemm(precision, tile, tile_n > 0 ? tile_n : tile, d_tasks.as<Task>(), d_segs.as<Seg>(),
               (int64_t)tasks.size(), J0, scales, st, rm);
}

}  // namespace tci
