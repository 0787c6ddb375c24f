// tci_htimesc_noalloc: the host-vector htimesc with its two device staging
// vectors in a fixed device array instead of per-call allocations.
//
// `scratch` is a file-scope device array of 2 * kNoallocMaxCsf doubles
// (64 MiB): [0, kNoallocMaxCsf) stages civec, [kNoallocMaxCsf, 2 kNoallocMaxCsf)
// stages sigmavec.  Every GPU the library runs this entry point on holds its
// own copy at its own address for the life of the process.  The host-side
// name `scratch` is not a device address, so the address is looked up once
// per device with hipGetSymbolAddress and only that pointer is handed to
// copies and kernels.
//
// Calls on one device share that device's array: two calls that overlap in
// time on one GPU (two host threads) would race on it, so calls are
// serialised per device by a host mutex; the array keeps the last call's
// contents, and each call overwrites the parts it uses (both staged vectors
// are copied in full before the sigma build) so nothing has to be cleared.
#include "plan.hpp"

#include <mutex>

namespace {

constexpr int64_t kNoallocMaxCsf = (int64_t)1 << 22;  // 4,194,304 CSFs per vector
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[2 * kNoallocMaxCsf];

std::mutex g_lock[kMaxDevices];
double* g_ptr[kMaxDevices] = {};

// device address of this device's `scratch` (looked up on the first call)
double* scratch_ptr(int dev) {
  if (!g_ptr[dev]) {
    void* p = nullptr;
    TCI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TCI_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the staging array");
    g_ptr[dev] = static_cast<double*>(p);
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" int tci_htimesc_noalloc(tci_plan* plan, const double* civec, double* sigmavec, int64_t n) {
  try {
    TCI_REQUIRE(plan && civec && sigmavec, "null argument");
    TCI_REQUIRE(n >= 1 && n <= kNoallocMaxCsf,
                "tci_htimesc_noalloc stages at most 4194304 CSFs per vector; use tci_htimesc for longer vectors");
    TCI_REQUIRE(n == plan->n_csf, "vector length does not match the plan's CSF count");
    int dev = 0;
    TCI_HIP(hipGetDevice(&dev));
    TCI_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> guard(g_lock[dev]);
    double* base = scratch_ptr(dev);
    double* dc = base;
    double* ds = base + kNoallocMaxCsf;
    const size_t bytes = (size_t)n * 8;
    TCI_HIP(hipMemcpy(dc, civec, bytes, hipMemcpyHostToDevice));
    TCI_HIP(hipMemcpy(ds, sigmavec, bytes, hipMemcpyHostToDevice));
    plan->sigma(dc, ds, TCI_SIGMA_FP64 | TCI_ACCUMULATE | TCI_ADD_EREF, nullptr);
    TCI_HIP(hipMemcpy(sigmavec, ds, bytes, hipMemcpyDeviceToHost));
    return TCI_OK;
  } catch (const tci::Error& e) {
    tci::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    tci::set_last_error(e.what());
    return TCI_ERR_ARG;
  } catch (...) {
    tci::set_last_error("unknown error");
    return TCI_ERR_ARG;
  }
}
