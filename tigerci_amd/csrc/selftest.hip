// Device check of the MFMA lane layouts the kernels rely on.
//
// For each instruction one wave multiplies a known A (16 x K) by a known B
// (K x 16) through the builtin, using the operand placement the kernels use,
// and a scalar loop computes the same product; every element must agree.
//   f64 16x16x4:   A[i][k] lane i+16k;            B[k][j] lane j+16k
//                  D[i][j] lane j+16(i%4), register i/4
//   bf16/f16 16x16x32: A[i][k] lane i+16(k/8), element k%8; B likewise
//                  D[i][j] lane j+16(i/4), register i%4
//   bf16/f16 32x32x16: A[i][k] lane i+32(k/8), element k%8; B likewise
//                  D[i][j] lane j+32((i/4)%2), register 4(i/8)+i%4
// Inputs are small integers so every product and sum is exact.
#include "common.hpp"

namespace tci {
namespace {

__global__ void mfma_layout_kernel(int* __restrict__ bad) {
  const int l = threadIdx.x;
  if (l >= 64) return;
  auto Af = [](int i, int k) { return (double)(((i * 7 + k * 3) % 11) - 5); };
  auto Bf = [](int k, int j) { return (double)(((k * 5 + j * 2) % 9) - 4); };
  int nbad = 0;
  {  // f64
    const int i = l & 15, g = l >> 4;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Af(i, g), Bf(g, i), acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) {
      const int row = g + 4 * r, col = i;
      double ref = 0.0;
      for (int k = 0; k < 4; ++k) ref += Af(row, k) * Bf(k, col);
      if (acc[r] != ref) ++nbad;
    }
  }
  {  // bf16 16x16x32
    const int i = l & 15, g = l >> 4;
    bf16x8 a, b;
    for (int t = 0; t < 8; ++t) {
      a[t] = (__bf16)(float)Af(i, 8 * g + t);
      b[t] = (__bf16)(float)Bf(8 * g + t, i);
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) {
      const int row = 4 * g + r, col = i;
      double ref = 0.0;
      for (int k = 0; k < 32; ++k) ref += Af(row, k) * Bf(k, col);
      if ((double)acc[r] != ref) ++nbad;
    }
  }
  {  // f16 16x16x32
    const int i = l & 15, g = l >> 4;
    h8 a, b;
    for (int t = 0; t < 8; ++t) {
      a[t] = (_Float16)(float)Af(i, 8 * g + t);
      b[t] = (_Float16)(float)Bf(8 * g + t, i);
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) {
      const int row = 4 * g + r, col = i;
      double ref = 0.0;
      for (int k = 0; k < 32; ++k) ref += Af(row, k) * Bf(k, col);
      if ((double)acc[r] != ref) ++nbad;
    }
  }
  {  // bf16 and f16 32x32x16
    typedef float f16v __attribute__((ext_vector_type(16)));
    const int j = l & 31, g = l >> 5;
    bf16x8 a, b;
    h8 ah, bh;
    for (int t = 0; t < 8; ++t) {
      a[t] = (__bf16)(float)Af(j, 8 * g + t);
      b[t] = (__bf16)(float)Bf(8 * g + t, j);
      ah[t] = (_Float16)(float)Af(j, 8 * g + t);
      bh[t] = (_Float16)(float)Bf(8 * g + t, j);
    }
    f16v acc, acch;
    for (int r = 0; r < 16; ++r) acc[r] = acch[r] = 0.f;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    acch = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acch, 0, 0, 0);
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r / 4) + 4 * g + r % 4, col = j;
      double ref = 0.0;
      for (int k = 0; k < 16; ++k) ref += Af(row, k) * Bf(k, col);
      if ((double)acc[r] != ref) ++nbad;
      if ((double)acch[r] != ref) ++nbad;
    }
  }
  bad[l] = nbad;
}

}  // namespace
}  // namespace tci

extern "C" int tci_selftest_mfma(void) {
  using namespace tci;
  try {
    DevBuf b;
    b.alloc(64 * sizeof(int));
    hipLaunchKernelGGL(mfma_layout_kernel, dim3(1), dim3(64), 0, nullptr, b.as<int>());
    TCI_HIP(hipGetLastError());
    int h[64];
    TCI_HIP(hipMemcpy(h, b.p, sizeof(h), hipMemcpyDeviceToHost));
    int tot = 0;
    for (int v : h) tot += v;
    if (tot) {
      set_last_error("MFMA lane layout mismatch in " + std::to_string(tot) + " elements");
      return TCI_ERR_NUMERIC;
    }
    return TCI_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}
