// Shared definitions of the tigersigma HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tigersigma.h"

namespace tci {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define TCI_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      throw ::tci::Error(TCI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e) + \
                                          " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define TCI_REQUIRE(cond, msg)                                                     \
  do {                                                                             \
    if (!(cond)) throw ::tci::Error(TCI_ERR_ARG, msg);                             \
  } while (0)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// Device allocation owned by a plan.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    reset();
    if (n == 0) n = 256;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(TCI_ERR_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed: " +
                                   hipGetErrorString(e));
    }
    bytes = n;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& h, hipStream_t s = nullptr) {
  b.alloc(h.size() * sizeof(T));
  if (!h.empty()) TCI_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of the
// chip (workgroup b runs on XCD b % 8), and every XCD has an L2 of its own.
// Tasks listed next to each other in a program share operand panels (the
// tiles of one output block, the K-splits of one walk), so task lists are
// dealt out in contiguous ranges: XCD x runs tasks [start_x, start_x + count_x)
// in order, and neighbouring tiles meet in one L2 instead of eight.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_task(int64_t b, int64_t G) {
  const int64_t q = G / kXcds, r = G % kXcds, x = b % kXcds, i = b / kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}
// Order of a program that has not been tuned: TCI_XCD_REMAP=0/1 forces plain
// or XCD-contiguous order for every program (and switches tuning of the order
// off); the default is plain order.
int xcd_remap_default();
bool xcd_remap_forced();

// Load through the global address space.  Pointers read from launch tables are
// generic to the compiler, which would otherwise emit flat loads; those also
// count against the LDS wait counter and would make every LDS read wait for
// all outstanding prefetches.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// ---------------------------------------------------------------------------
// Generic grouped gather-GEMM (FP64 MFMA):  C[tile] (+)= sum_seg alpha A_seg B_seg
// An operand element (r,k) is base[J0*jstride + rmap(r)*rs + cmap(k)*cs]; a map
// entry < 0 (or a null base) reads as zero; a null map is the identity.  J0 is
// a launch argument, so one table serves every Cholesky-vector chunk.
struct Op {
  const void* base;   // double, or float when f32 != 0
  int64_t rs, cs;
  const int32_t* rmap;
  const int32_t* cmap;
  int64_t jstride;
  int32_t f32;
};
struct Seg {
  Op A, B;
  int32_t K;
  double alpha;
};
struct Task {
  double* C;
  int64_t ldc;
  int32_t M, N;     // full output extent (bounds)
  int32_t m0, n0;   // this tile's origin
  int32_t seg0, seg1;
  int32_t accumulate;
  int32_t out_split;  // split kernel only: C is a split-row 16-bit matrix, ldc in 16-bit elements
  int32_t orow_div;  // dense kernels: if > 0, logical row r is stored at row (r % div) * (M / div) + r / div
  // dense kernels, blocked output: element (row, col) is stored at
  //   (row / rdiv) rstride + (row % rdiv) ldc + (col / cdiv) cstride + (col % cdiv)
  // (column offset through xidx for split-row output); rdiv / cdiv = 0: no blocking
  int32_t rdiv, cdiv;
  int64_t rstride, cstride;
  int32_t oscale;    // split kernel: result multiplied by scales[oscale] (-1: not scaled)
  int32_t pad_;
};
struct Blocking {
  int32_t rdiv = 0, cdiv = 0;
  int64_t rstride = 0, cstride = 0;
};

// A list of tasks and their segments, kept on the host while being built and
// mirrored on the device.
// tile = 64: 64x64 tiles, 4 waves of 32x32 (small or skinny outputs);
// tile = 128: 128x128 tiles, double-buffered LDS, 4 waves of 64x64 with
// wave-uniform skipping of 16x16 blocks outside the output.
struct GemmProgram {
  int tile = 64;
  int tile_n = 0;      // split programs: tile width in N (0: square tiles)
  bool dense = false;  // contiguous row-major operands only: dense_gemm.hip kernel (tile 64 or 128)
  bool split = false;  // split-row operands, C = sum A B^T: dense_split.hip kernel (tile 64 or 128)
  bool scaled = false; // some task has an output scale (set by upload_tables)
  int dma = 0;         // split programs: LDS stages filled by direct loads (0: register-staged)
  int waves = 4;       // split programs on 256 x 256 tiles: 4 waves of 128 x 128 or 8 of 128 x 64
  int remap = -1;      // workgroup -> task order: 0 plain, 1 XCD-contiguous, -1 default
  double alg_flops = 0.0;  // algorithmic flops of one launch: 2 m n k over valid extents (upload_tables)
  std::vector<Task> tasks;
  std::vector<Seg> segs;
  DevBuf d_tasks, d_segs;
  void clear() { tasks.clear(); segs.clear(); }
  // add one output matrix: tiles over (M,N), all sharing `segs_for_output`
  void add(double* C, int64_t ldc, int M, int N, const std::vector<Seg>& s, bool accumulate,
           bool out_split = false, int orow_div = 0, Blocking blk = Blocking(), int oscale = -1);
  void upload_tables(hipStream_t st);
  void launch(hipStream_t st, int64_t J0 = 0) const;
  // name of the kernel a launch runs, e.g. "xgemmd_kernel<192,192,3>"
  std::string kernel_name() const;
  // Reorder the tiles of every output grid (tasks with one C and one segment
  // range) into blocks of G x G tiles, row blocks outer, keeping the grids in
  // their order; G = 0 restores `base`.  Dealt XCD-contiguously (remap 1), the
  // tiles that run together on one XCD then share G row and G column panels
  // in its L2.  Results do not depend on the order.
  void block_order(int G, const std::vector<Task>& base, hipStream_t st);
  // scales: device array the tasks' oscale index into (may be null when unused)
  void launch_split(int precision, hipStream_t st, int64_t J0 = 0, const float* scales = nullptr) const;
};

// Autotuning helpers (tune.hip).  time_ms: mean duration in ms of `reps` runs
// of fn on st after one untimed run (HIP events).  tune_order: sets g.remap to
// the faster workgroup order of g's launch and returns it (unchanged when the
// order is forced from the environment).  autotune_enabled: TCI_AUTOTUNE != 0.
bool autotune_enabled();
std::string fmt_ms(double ms);
double time_ms(hipStream_t st, int reps, const std::function<void()>& fn);
int tune_order(GemmProgram& g, hipStream_t st, const std::function<void()>& launch);
// tune_block_order: times plain order and XCD-contiguous order with tile
// blocks of G = 0, 4, 8 and 16, keeps the fastest (unless within 3% of the
// current order) and returns a short record such as "x8"
std::string tune_block_order(GemmProgram& g, hipStream_t st, const std::function<void()>& launch);

void launch_ggemm(int tile, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, hipStream_t st,
                  int remap);
void launch_dgemm(int tile, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, hipStream_t st,
                  int remap);
void launch_xgemm(int precision, int tm, int tn, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0,
                  const float* scales, hipStream_t st, int remap);

// ---------------------------------------------------------------------------
// Split-row storage of a matrix for the split-precision GEMM.  A row of K
// values is padded with zeros to Kp = K rounded up to 8 and kept as Kp / 8
// chunks of 16 bf16 values: the high parts hi = bf16(x) of 8 consecutive
// elements, then their low parts lo = bf16(x - hi).  Row stride 2 Kp; every
// chunk is 32-byte aligned when the matrix base is.
inline int64_t xpad(int64_t K) { return (K + 7) & ~(int64_t)7; }
__device__ __forceinline__ int64_t xidx(int64_t k) { return ((k >> 3) << 4) + (k & 7); }
// write the split form of s v[0..7] as one chunk: BF16 parts, or FP16 parts
// when f16 (scaled FP16: s is a power of two that keeps |s v| below 2^14)
__device__ __forceinline__ void put_split8(uint16_t* chunk, const float* v, float s = 1.f, bool f16 = false) {
  if (f16) {
    h8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float x = v[e] * s;
      h[e] = (_Float16)x;
      l[e] = (_Float16)(x - (float)h[e]);
    }
    __attribute__((address_space(1))) h8* p = (__attribute__((address_space(1))) h8*)chunk;
    p[0] = h;
    p[1] = l;
  } else {
    bf16x8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float x = v[e] * s;
      h[e] = (__bf16)x;
      l[e] = (__bf16)(x - (float)h[e]);
    }
    __attribute__((address_space(1))) bf16x8* p = (__attribute__((address_space(1))) bf16x8*)chunk;
    p[0] = h;
    p[1] = l;
  }
}
// scale read from device memory (null: 1)
__device__ __forceinline__ float scale_at(const float* p) { return p ? *p : 1.f; }

}  // namespace tci
